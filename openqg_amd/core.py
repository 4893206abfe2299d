"""Drop-in counterparts of OpenQG's `core` and `driver` entry points.

step_qg keeps the argument order of src/qg/core.F90:16 so a Python port of
driver.F90 calls it the same way; suppress_comp_mode mirrors driver.F90:174.
"""
from __future__ import annotations

from .qg import QG


def step_qg(tdt: float, ent, wekp, txis: float, txin: float, qg: QG) -> None:
    """vort_step + solve_pressure + ocqbdy on the GPU (core.F90:16-33)."""
    qg.step(tdt, ent, wekp, txis, txin)


def suppress_comp_mode(qg: QG) -> None:
    """Average the leapfrog time levels (driver.F90:174-189)."""
    qg.suppress()


def init_dynamic_state(qg: QG, p, pm) -> None:
    """Set p, pm and derive q, qm and the constraint state (openqg.F90:222-243)."""
    qg.set_state(p=p, pm=pm)
    qg.init_state()
