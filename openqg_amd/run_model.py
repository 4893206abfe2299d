"""Run an OpenQG experiment on the GPU: a Python 3 counterpart of the reference's
run_model.py (run_model.py:1-300).

The flow is the reference's: stage the experiment's files into `stage-n`
(setup_from_config / setup_from_dir, run_model.py:95-118), run the model into
`out-n` (run_single, :142-151), then build `in-(n+1)` from the stage and the
run's `*_lastday.nc` files with the clock advanced by one run (update_clock,
:154-175) and the basins' state names pointing at the lastday files
(update_basin, :177-204; create_next_input_dir, :206-218), `-r` times.

What differs: the model runs in this process through the HIP library
(openqg_amd.experiment.Experiment) instead of an executable, and the CDL
configuration files are read as text instead of being turned into netCDF by
ncgen (absent here, as is Python 2).  `-x` and `-n` are accepted for
command-line compatibility and ignored.  The predefined experiments (`-e`) take
their files from the reference's data directory, given by `--data` (default
$OPENQG_DATA, else ./data).

    python -m openqg_amd.run_model -e dg_fast -o out --data /path/to/openqg/data
"""
from __future__ import annotations

import argparse
import os
import re
import shutil
import sys


class Config:
    glam = None
    ocn = None
    atm = None
    wind = None
    clock = None
    data = None


class DoubleGyre(Config):
    glam = "glam-dg.cdl"
    ocn = "ocn_dg.cdl"
    atm = "atm_dg.cdl"
    wind = "windstress-1.3.cdl"
    clock = "clock-0-5yr-6.cdl"


class DoubleGyreFast(DoubleGyre):
    clock = "clock-0-10dy-6.cdl"


class SouthernOceanOnly(Config):
    glam = "glam-so-only.cdl"
    ocn = "ocn_so_only.cdl"
    wind = "windstress-avges.cdl"
    clock = "clock-0-5yr-3.cdl"
    data = ["soctopog.26deg.10km.new.nc", "avges.nc"]


class SouthernOceanOnlyFast(SouthernOceanOnly):
    clock = "clock-0-10dy-3.cdl"


class SouthernOcean(Config):
    glam = "glam-so.cdl"
    ocn = "ocn_so.cdl"
    atm = "atm_so.cdl"
    wind = "windstress-1.5.cdl"
    clock = "clock-100-5yr-3.cdl"
    data = ["soctopog.26deg.10km.new.nc", "lastday_so_ocn.nc", "lastday_so_atm.nc", "lastday_so_aml.nc",
            "lastday_so_oml.nc"]


class SouthernOceanTau(SouthernOcean):
    wind = "windstress-1.5-udiff.cdl"


class SouthernOceanFast(SouthernOcean):
    clock = "clock-100yr-10dy-3.cdl"


class SouthernOceanTauFast(SouthernOceanTau):
    clock = "clock-100yr-10dy-3.cdl"


# the reference's experiment names (run_model.py:78-87)
CONFIGS = {"dg": DoubleGyre, "dg_fast": DoubleGyreFast, "so_only": SouthernOceanOnly,
           "so_only_fast": SouthernOceanOnlyFast, "so": SouthernOcean, "so_fast": SouthernOceanFast,
           "so_tau": SouthernOceanTau, "so_tau_fast": SouthernOceanTauFast}


def setup_from_config(config, setup_dir: str, data_dir: str):
    """Copy a predefined experiment's files into setup_dir under the names the model reads."""
    os.makedirs(setup_dir, exist_ok=True)
    shutil.copy(os.path.join(data_dir, "glam", config.glam), os.path.join(setup_dir, "glam.cdl"))
    shutil.copy(os.path.join(data_dir, "basin", config.ocn), os.path.join(setup_dir, "ocn_basin.cdl"))
    if config.atm:
        shutil.copy(os.path.join(data_dir, "basin", config.atm), os.path.join(setup_dir, "atm_basin.cdl"))
    shutil.copy(os.path.join(data_dir, "windstress", config.wind), os.path.join(setup_dir, "windstress.cdl"))
    shutil.copy(os.path.join(data_dir, "clocks", config.clock), os.path.join(setup_dir, "clock.cdl"))
    for fn in config.data or []:
        src = os.path.join(data_dir, "nc_files", fn)
        if not os.path.isfile(src):
            raise FileNotFoundError(f"{src}: the experiment's data file is not available")
        shutil.copy(src, setup_dir)


def setup_from_dir(input_dir: str, setup_dir: str):
    """Copy every .nc and .cdl file of input_dir into setup_dir."""
    os.makedirs(setup_dir, exist_ok=True)
    for fn in sorted(os.listdir(input_dir)):
        if fn.endswith((".nc", ".cdl")):
            shutil.copy(os.path.join(input_dir, fn), setup_dir)


class _Tee:
    def __init__(self, *streams):
        self.streams = streams

    def write(self, s):
        for st in self.streams:
            st.write(s)

    def flush(self):
        for st in self.streams:
            st.flush()


def run_model(stage_dir: str, out_dir: str) -> int:
    """Run the staged experiment into out_dir, its output also in out_dir/stdout.txt."""
    from .experiment import Experiment

    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, "stdout.txt"), "w") as log:
        tee = _Tee(sys.stdout, log)
        exp = Experiment(stage_dir)
        clk = exp.clk
        print(f"stage {stage_dir} -> {out_dir}: steps {clk.nsteps0 + 1}..{clk.nsteps} of {clk.dta:g} s "
              f"(ocean every {clk.nstr})", file=tee, flush=True)
        try:
            nt = exp.run(out_dir)
        finally:
            exp.close()
        print(f"done at step {nt}, day {nt * clk.dta / 86400.0:.3f}", file=tee, flush=True)
    return nt


def run_single(input_dir: str, stage_dir: str, out_dir: str) -> int:
    setup_from_dir(input_dir, stage_dir)
    return run_model(stage_dir, out_dir)


def update_clock(filename: str):
    """Advance the clock file by one run: tini += trun."""
    text = open(filename).read()
    tini = float(re.search(r"^\s*tini\s*=\s*([^;]+);", text, re.M).group(1))
    trun = float(re.search(r"^\s*trun\s*=\s*([^;]+);", text, re.M).group(1))
    text = re.sub(r"^(\s*)tini\s*=\s*[^;]+;", lambda m: f"{m.group(1)}tini = {tini + trun:f};", text, flags=re.M)
    text = re.sub(r"^(\s*)trun\s*=\s*[^;]+;", lambda m: f"{m.group(1)}trun = {trun:f};", text, flags=re.M)
    with open(filename + ".tmp", "w") as f:
        f.write(text)
    os.replace(filename + ".tmp", filename)


def update_basin(filename: str):
    """Point a basin's state entries (qg_state, oml_state, aml_state) at the
    `<name>_<qg|ml>_lastday.nc` files of the previous run."""
    if not os.path.exists(filename):
        return
    text = open(filename).read()
    name = re.search(r'^\s*name\s*=\s*"([^"]*)"', text, re.M).group(1).strip()

    def repl(m):
        kind = m.group(1).replace("oml", "ml").replace("aml", "ml")
        return f'{m.group(1)}_state = "{name}_{kind}_lastday.nc"'

    text = re.sub(r'\b([A-Za-z0-9]+)_state\s*=\s*"[^"]*"', repl, text)
    with open(filename + ".tmp", "w") as f:
        f.write(text)
    os.replace(filename + ".tmp", filename)


def create_next_input_dir(stage_dir: str, out_dir: str, next_in: str):
    os.makedirs(next_in, exist_ok=True)
    for fn in os.listdir(stage_dir):
        shutil.copy(os.path.join(stage_dir, fn), next_in)
    for fn in os.listdir(out_dir):
        if fn.endswith("lastday.nc"):
            shutil.copy(os.path.join(out_dir, fn), next_in)
    update_basin(os.path.join(next_in, "ocn_basin.cdl"))
    update_basin(os.path.join(next_in, "atm_basin.cdl"))
    update_clock(os.path.join(next_in, "clock.cdl"))


def main(output_dir: str, repeats: int = 1, input_dir: str | None = None, exp: str | None = None,
         data_dir: str | None = None):
    os.makedirs(output_dir, exist_ok=True)
    setup_dir = os.path.join(output_dir, "in-orig")
    if input_dir is not None:
        setup_from_dir(input_dir, setup_dir)
    elif exp is not None:
        setup_from_config(CONFIGS[exp], setup_dir, data_dir or os.environ.get("OPENQG_DATA", "data"))
    else:
        raise ValueError("one of input_dir or exp is needed")
    shutil.copytree(setup_dir, os.path.join(output_dir, "in-0"), dirs_exist_ok=True)
    for r in range(repeats):
        in_dir = os.path.join(output_dir, f"in-{r}")
        stage_dir = os.path.join(output_dir, f"stage-{r}")
        out_dir = os.path.join(output_dir, f"out-{r}")
        run_single(in_dir, stage_dir, out_dir)
        create_next_input_dir(stage_dir, out_dir, os.path.join(output_dir, f"in-{r + 1}"))


def parse_args(argv=None):
    p = argparse.ArgumentParser(description="Run an OpenQG experiment on the GPU.",
                                epilog="One of the -e or -i options must be provided.")
    p.add_argument("-x", "--exe", dest="exe", type=str, default=None, help="accepted and ignored (runs in-process)")
    p.add_argument("-o", "--output", dest="output_dir", type=str, required=True, help="Output directory")
    p.add_argument("-n", "--num_threads", dest="num_threads", type=int, default=4, help="accepted and ignored")
    p.add_argument("-r", "--runs", dest="repeats", type=int, default=1, help="The number of sequential runs")
    p.add_argument("-i", "--input", dest="input_dir", type=str, help="Directory containing input files.")
    p.add_argument("-e", "--exp", dest="exp", type=str, choices=CONFIGS, help="A predefined experiment.")
    p.add_argument("--data", dest="data_dir", type=str, default=None,
                   help="the reference's data directory for -e (default $OPENQG_DATA or ./data)")
    a = p.parse_args(argv)
    if a.exp is None and a.input_dir is None:
        p.print_usage()
        print(f"{sys.argv[0]}: error: One of the -e or -i options must be provided.", file=sys.stderr)
        sys.exit(1)
    strip = lambda d: d[:-1] if d and d.endswith("/") else d
    return strip(a.output_dir), a.repeats, strip(a.input_dir), a.exp, a.data_dir


if __name__ == "__main__":
    main(*parse_args())
