"""A small coupled double-gyre experiment written as CDL files, for the harness
tests: the double-gyre physics on a 16 x 8 mesh at 120 km (atmosphere channel
17 x 9, ocean box 33 x 33 at 15 km starting at mesh row 3)."""
import os

GLAM = """netcdf glam {{
variables:
\tinteger nxt; integer nyt; double dx; double dy; double x0; double y0;
\tinteger cyclic; double lat; double x1; integer use_atmos; integer use_ocean;
data:
\tnxt = 16; nyt = 8; dx = 120.0e3; dy = 120.0e3; x0 = 0.0; y0 = 0.0;
\tcyclic = 0; lat = 43.29; x1 = 0.0;
\tuse_atmos = 1;
\tuse_ocean = 1;
}}
"""

OCN = """netcdf ocn_basin {{
dimensions:
\tstrlen = 64 ; nl = 3 ;
variables:
\tchar name(strlen); integer nx; integer ny; integer ndx; integer ndy; integer nx1; integer ny1;
\tdouble hm; double h(nl); char topog(strlen); double rho;
\tinteger use_oml; integer flux_n; integer flux_s; double st2d, st4d, ycexp; double cp;
\tinteger use_rad_temp; double T1_abs; double T2_abs; char oml_state(strlen);
\tinteger use_qg; double bcco; double delek; double ah2(nl); double ah4(nl); double gp(nl);
\tchar qg_state(strlen); integer check_valids; double dtvalid; double max_p; double max_q;
data:
\tname = "small_ocean";
\tnx = 4; ny = 4; ndx = 8; ndy = 8; nx1 = 1; ny1 = 3;
\thm = 100.0; h = 300.0, 1100.0, 2600.0;
\ttopog = "flat"; rho = 1.0e3;
\tuse_oml = 1; flux_n = 0; flux_s = 1; st2d = 380.0; st4d = 4.0e10; ycexp = 1.0; cp = 4.0e3;
\tuse_rad_temp = 1; T1_abs = 278.0; T2_abs = 268.0; oml_state = "{oml}";
\tuse_qg = 1; bcco = 0.5; delek = 1.0; ah2 = 0.0, 0.0, 0.0; ah4 = 1.0e10, 1.0e10, 1.0e10; gp = 5.0e-2, 2.5e-2;
\tqg_state = "{qg}";
\tcheck_valids = 1; dtvalid = 0.1; max_p = 1.0e6; max_q = 1.0;
}}
"""

ATM = """netcdf atm_basin {{
dimensions:
\tstrlen = 64 ; nl = 3 ;
variables:
\tchar name(strlen); integer nx; integer ny; integer ndx; integer ndy; integer nx1; integer ny1;
\tdouble hm; double h(nl); char topog(strlen); double rho;
\tinteger use_aml; double st2d, st4d, ahmd; double cp; double hmamin; double hmadmp; double xcexp;
\tdouble T_abs(nl); double fsbar; double fspamp; double zopt_m; double zopt(nl); double gamma; double xlamda;
\tchar aml_state(strlen); integer use_qg; double bcco; double delek; double ah2(nl); double ah4(nl); double gp(nl);
\tchar qg_state(strlen); integer check_valids; double dtvalid; double max_p; double max_q;
data:
\tname = "small_atmos";
\tnx = 16; ny = 8; ndx = 1; ndy = 1; nx1 = 1; ny1 = 1;
\thm = 1000.0; h = 2000.0, 3000.0, 4000.0;
\ttopog = "flat"; rho = 1.0;
\tuse_aml = 1; st2d = 5.0e4; st4d = 3.0e14; ahmd = 2.0e5; cp = 1.0e3; hmamin = 100.0; hmadmp = 0.15; xcexp = 1.0;
\tT_abs = 330.0, 340.0, 350.0; fsbar = -210.0; fspamp = 80.0; zopt_m = 200.0; zopt = 3.0e4, 3.0e4, 4.0e4;
\tgamma = 1.0e-2; xlamda = 35.0; aml_state = "{aml}";
\tuse_qg = 1; bcco = 1.0; delek = 0.0; ah2 = 0.0, 0.0, 0.0; ah4 = 2.0e13, 2.0e13, 2.0e13; gp = 1.2, 0.4;
\tqg_state = "{qg}";
\tcheck_valids = 1; dtvalid = 0.1; max_p = 1.0e7; max_q = 0.05;
}}
"""

WIND = """netcdf windstress {
variables:
\tinteger coupled; double cdat; integer tau_udiff;
data:
\tcoupled = 1;
\tcdat = 1.3e-3; // drag coefficient
\ttau_udiff = 0;
}
"""

CLOCK = """netcdf clock {{
variables:
\tdouble dta; double tini; double trun; integer nstr; double resday;
data:
\tdta = 300.0;
\ttini = {tini};
\ttrun = {trun};
\tnstr = 6;
\tresday = {resday};
}}
"""


def write_experiment(d, days=0.5, resday=0.25, tini=0.0, qg_o="zero", qg_a="rbal", oml="rbal", aml="rbal"):
    os.makedirs(d, exist_ok=True)
    files = {"glam.cdl": GLAM.format(), "ocn_basin.cdl": OCN.format(oml=oml, qg=qg_o),
             "atm_basin.cdl": ATM.format(aml=aml, qg=qg_a), "windstress.cdl": WIND,
             "clock.cdl": CLOCK.format(tini=repr(tini), trun=repr(days / 365.0), resday=repr(resday))}
    for n, t in files.items():
        with open(os.path.join(d, n), "w") as f:
            f.write(t)
    return d
