"""The oracle's pins are sharp: plausible slips planted in a copy of oracle/cpuref.c fail them.

Each case rewrites one line of the oracle (a sign, an index, a transposed
operand, a dropped factor in the routines of src/qg/vorsubs.F90,
constraint.F90, homog.F90, vorticity.F90, glam/numerics.F90 and the ocean
mixed layers and the coupling that tests/test_oracle_pins.py, tests/test_oml_oracle.py and
tests/test_cpl_oracle.py fix), compiles the copy into a temporary
directory and runs the pin tests against it (OQG_CPUREF_LIB); they must fail.
The oracle in the tree is never modified.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "cpuref.c")
CFLAGS = ["-O1", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-std=c11", "-shared"]

# (case, line in the oracle, planted replacement)
M = [
 ("qcomp beta sign", "E3(q, i, j, k) = E3(q, i, j, k) + b->P.beta * b->yprel[j - 1];\n        for (int i = 1; i <= NXP; i++)\n            for (int k = 1; k <= nl; k++) {",
                     "E3(q, i, j, k) = E3(q, i, j, k) - b->P.beta * b->yprel[j - 1];\n        for (int i = 1; i <= NXP; i++)\n            for (int k = 1; k <= nl; k++) {"),
 ("qcomp A transposed", "for (int l = 1; l <= nl; l++) s += M2(b->amat, k, l) * E3(p, i, j, l);\n                E3(q, i, j, k) = E3(q, i, j, k) - b->P.fnot * s;\n            }\n    }\n}",
                        "for (int l = 1; l <= nl; l++) s += M2(b->amat, l, k) * E3(p, i, j, l);\n                E3(q, i, j, k) = E3(q, i, j, k) - b->P.fnot * s;\n            }\n    }\n}"),
 ("qcomp ddyn layer", "for (int i = 1; i <= NXP; i++) E3(q, i, j, kt) = E3(q, i, j, kt) + E2(b->ddyn, i, j);",
                      "for (int i = 1; i <= NXP; i++) E3(q, i, j, 1) = E3(q, i, j, 1) + E2(b->ddyn, i, j);"),
 ("ocqbdy north diff sign", "E3(q, i, NYP, k) = fac * (E3(p, i, NYP - 1, k) - E3(p, i, NYP, k));",
                            "E3(q, i, NYP, k) = fac * (E3(p, i, NYP, k) - E3(p, i, NYP - 1, k));"),
 ("ocqbdy west index", "E3(q, 1, j, k) = fac * (E3(p, 2, j, k) - E3(p, 1, j, k));",
                       "E3(q, 1, j, k) = fac * (E3(p, 3, j, k) - E3(p, 1, j, k));"),
 ("ocqbdy corner beta once", "E3(q, 1, j, k) = E3(q, 1, j, k) + b->P.beta * b->yprel[j - 1];",
                             "if (j > 1 && j < NYP) E3(q, 1, j, k) = E3(q, 1, j, k) + b->P.beta * b->yprel[j - 1];"),
 ("dpi from dpi", "b->dpi[l] = b->dpip[l] - tdt * b->P.gp[l] * exn;", "b->dpi[l] = b->dpi[l] - tdt * b->P.gp[l] * exn;"),
 ("dpi gp index", "b->dpi[l] = b->dpip[l] - tdt * b->P.gp[l] * exn;", "b->dpi[l] = b->dpip[l] - tdt * b->P.gp[l + 1] * exn;"),
 ("init_constraint cs sign", "b->cs[k - 1] = -b->cs[k - 1] + c * aps;", "b->cs[k - 1] = b->cs[k - 1] + c * aps;"),
 ("init_constraint A transposed", "apsp += M2(b->amat, k, l) * pinsp[l - 1];", "apsp += M2(b->amat, l, k) * pinsp[l - 1];"),
 ("init_constraint cn row", "row[i - 1] = E3(b->p, i, NYP, k) - E3(b->p, i, NYP - 1, k);\n            b->cn[k - 1]",
                            "row[i - 1] = E3(b->p, i, NYP - 1, k) - E3(b->p, i, NYP - 2, k);\n            b->cn[k - 1]"),
 ("ajis weight", "*ajis = (aj5sms + 2.0 * aj9sms) / 6.0;", "*ajis = (2.0 * aj5sms + aj9sms) / 6.0;"),
 ("ajin sign", "*ajin = -(aj5smn + 2.0 * aj9smn) / 6.0;", "*ajin = (aj5smn + 2.0 * aj9smn) / 6.0;"),
 ("ajis dp row", "row[i - 1] = E2(q, i, 1) * E2(dpx, i, 2);", "row[i - 1] = E2(q, i, 1) * E2(dpx, i, 1);"),
 ("step_constraints from cs", "double csnew = b->csp[k] + tdt * rhss[k];", "double csnew = b->cs[k] + tdt * rhss[k];"),
 ("int_Be_n fohfac(1)", "b->int_Be_n[0] = -b->P.dz_sign * fohfac[1]", "b->int_Be_n[0] = -b->P.dz_sign * fohfac[0]"),
 ("txin sign", "rhsn[0] = rhsn[0] - b->P.tau_sign * (f0 / b->P.h[0]) * b->txin;", "rhsn[0] = rhsn[0] + b->P.tau_sign * (f0 / b->P.h[0]) * b->txin;"),
 ("ap3n sign", "rhsn[k] = rhsn[k] + b->ajin[k] + b->ap3n[k] - b->ap5n[k];", "rhsn[k] = rhsn[k] + b->ajin[k] - b->ap3n[k] - b->ap5n[k];"),
 ("ap5s from del2p", "row[i - 1] = E3(d4p, i, 2, k) - E3(d4p, i, 1, k);", "row[i - 1] = E3(del2p, i, 2, k) - E3(del2p, i, 1, k);"),
 ("bdrins layer", "row[i - 1] = E3(b->pm, i, 2, kt) - E3(b->pm, i, 1, kt);", "row[i - 1] = E3(b->pm, i, 2, 1) - E3(b->pm, i, 1, 1);"),
 ("bottom drag to layer kt", "rhss[nl - 1] = rhss[nl - 1] + (f0 / b->P.h[nl - 1]) * b->bdrins;", "rhss[0] = rhss[0] + (f0 / b->P.h[nl - 1]) * b->bdrins;"),
 ("cyclic_homog c_bc2 sign", "c2[m - 1] = b->hc1s[m - 1] * clhsn[m - 1] - b->hc1n[m - 1] * clhss[m - 1];",
                             "c2[m - 1] = b->hc1n[m - 1] * clhsn[m - 1] - b->hc1s[m - 1] * clhss[m - 1];"),
 ("cyclic_homog row nyp-1", "clhsn[m - 1] = n + cr_int_P_dx(b, &E3(inh, 1, NYP - 1, m)) / b->P.dy;", "clhsn[m - 1] = n + cr_int_P_dx(b, &E3(inh, 1, NYP - 2, m)) / b->P.dy;"),
 ("homsol pch rdm2 term", "pch1ys = pch1ys + 0.5 * dy * b->rdm2[m - 1] * h1[0] * xl;", "pch1ys = pch1ys - 0.5 * dy * b->rdm2[m - 1] * h1[0] * xl;"),
 ("hbsi inverted", "b->hbsi = b->yl / b->xl;", "b->hbsi = b->xl / b->yl;"),
 ("del2_T south flux sign", "else dyl = south_flux ? TT(t, i, 1) - sval : 0.0;", "else dyl = south_flux ? sval - TT(t, i, 1) : 0.0;"),
 ("del2_T x difference", "if (i >= 2) dxl = TT(t, i, j) - TT(t, i - 1, j);", "if (i >= 2) dxl = TT(t, i - 1, j) - TT(t, i, j);"),
 ("vek sign", "cv = -P->tau_sign / (P->fnot * hm);", "cv = P->tau_sign / (P->fnot * hm);"),
 ("wekt x term", "(1.0 / (P->dx * P->fnot)) * (tyr - tyl)", "(1.0 / (P->dx * P->fnot)) * (tyl - tyr)"),
 ("wall velocity kept", "if (!P->cyclic && (I == 1 || I == nxp)) vu = 0.0;", "if (0) vu = 0.0;"),
 ("meridional flux sign", "(-1.0 / P->dy) * (vt[1] - vt[0]);", "(1.0 / P->dy) * (vt[1] - vt[0]);"),
 ("heat flux without hm", "(1.0 / (o->rho_cp * o->hm)) * fnet[t]", "(1.0 / o->rho_cp) * fnet[t]"),
 ("convection sign", "dtc[t] = fmax(0.0, o->T1_rel - snew[t]);", "dtc[t] = fmin(0.0, o->T1_rel - snew[t]);"),
 ("entrainment mean kept", "for (size_t t = 0; t < nt; t++) xfo[t] = xfo[t] - avg;", "for (size_t t = 0; t < nt; t++) xfo[t] = xfo[t];"),
 ("map_T_to_P corner", "PP(out, 1, 1) = TT(t, 1, 1);", "PP(out, 1, 1) = TT(t, 2, 1);"),
 ("load_grid ny1 offset", "g->ny1 = 1 + (nyta - g->nyaooc) / 2;", "g->ny1 = 1 + (nyta - g->nyaooc + 1) / 2;"),
 ("Hermite slope basis", "{0, 1, -2, 1}, {0, 0, -1, 1}};", "{0, 1, -2, 1}, {0, 0, 1, -1}};"),
 ("bicubic corner order", "d[4 * t + ip + 2 * jq] =", "d[4 * t + 2 * ip + jq] ="),
 ("dPdx_2bc wall alpha", "PP(out, 1, j) = (PP(p, 2, j) - PP(p, 1, j)) / (P->dx * (0.5 * alpha + 1.0));",
                         "PP(out, 1, j) = (PP(p, 2, j) - PP(p, 1, j)) / (P->dx * (alpha + 1.0));"),
 ("dP2dy2_bc south sign", "PP(out, i, 1) = fac * alpha / (0.5 * alpha + 1.0) * (PP(p, i, 2) - PP(p, i, 1));",
                          "PP(out, i, 1) = fac * alpha / (0.5 * alpha + 1.0) * (PP(p, i, 1) - PP(p, i, 2));"),
 ("stress turning sign", "tauya[t] = cdochi * (v1 + cuohf * u1);", "tauya[t] = cdochi * (v1 - cuohf * u1);"),
 ("udiff ocean sign", "u1o[t] = u1o[t] + oy[t] / fo;", "u1o[t] = u1o[t] - oy[t] / fo;"),
 ("udiff depth ratio", "(g->cdat / fa) * (1.0 / g->hma + raoro / g->hmo)", "(g->cdat / fa) * (1.0 / g->hma + 1.0 / g->hmo)"),
 ("ux scale", "ux[c] = -(dxa / fa) * dxy[a];", "ux[c] = -(dya / fa) * dyy[a];"),
 ("bilint y weight", "const double wpy = dyainv * (yto - yam), wmy = 1.0 - wpy;",
                     "const double wmy = dyainv * (yto - yam), wpy = 1.0 - wmy;"),
 ("ocean flux not gathered", "TT(fnetat, ia, ja) = TT(fnetat, ia, ja) + ocfrac * (up[t] + sl[t]);",
                             "TT(fnetat, ia, ja) = TT(fnetat, ia, ja) + ocfrac * (up[t] - sl[t]);"),
 ("fsprim ocean rows", "const double fsp = fsprim(Pa, a->fspco, ytrel(Po, jo));",
                       "const double fsp = fsprim(Pa, a->fspco, ytrel(Pa, jo));"),
 ("eta term factor", "const double me = (fmafac * 0.25)", "const double me = (fmatop * 0.25)"),
 ("aml diabatic sign", "double h = hmixam[t] + tdta * (dh[t] - diab);", "double h = hmixam[t] + tdta * (dh[t] + diab);"),
 ("aml convection sign", "dtc[t] = fmin(0.0, a->T1_rel - an);", "dtc[t] = fmax(0.0, a->T1_rel - an);"),
 ("aml hmixa walls", "cr_oml_params fh = {hm, a->T1_rel, a->dT_12, a->rho_cp, a->ahmd, 0.0, 1.0, hm, hm, 1, 1};",
                     "cr_oml_params fh = {hm, a->T1_rel, a->dT_12, a->rho_cp, a->ahmd, 0.0, 1.0, hm, hm, 0, 0};"),
 ("aml entrainment current h", "(a->xcexp / (tdta * a->dT_12)) * hmixa[t] * dtc[t];",
                               "(a->xcexp / (tdta * a->dT_12)) * hmixam[t] * dtc[t];"),
 ("aml topography term", "ent1[t] = ent1[t] + a->bface * dtop[t];", "ent1[t] = ent1[t] + a->dface * dtop[t];"),
]


def _caught(case, tmp):
    name, old, new = case
    text = open(SRC).read()
    assert text.count(old) == 1, f"{name}: the oracle line to mutate is not unique"
    d = os.path.join(tmp, name.replace(" ", "_"))
    os.makedirs(d)
    c = os.path.join(d, "cpuref.c")
    open(c, "w").write(text.replace(old, new))
    with open(os.path.join(ROOT, "oracle", "cpuref.h")) as h:
        open(os.path.join(d, "cpuref.h"), "w").write(h.read())
    so = os.path.join(d, "libcpuref.so")
    subprocess.run(["gcc", *CFLAGS, "-o", so, c, "-lm"], check=True, capture_output=True)
    env = dict(os.environ, OQG_CPUREF_LIB=so, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_pins.py"), os.path.join(ROOT, "tests", "test_oracle.py"),
                        os.path.join(ROOT, "tests", "test_oml_oracle.py"), os.path.join(ROOT, "tests", "test_cpl_oracle.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True)
    return name, r.returncode != 0


def test_every_planted_mistake_fails_the_pins(tmp_path):
    with ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        res = list(ex.map(lambda c: _caught(c, str(tmp_path)), M))
    missed = [n for n, ok in res if not ok]
    assert not missed, f"planted mistakes not caught: {missed}"
