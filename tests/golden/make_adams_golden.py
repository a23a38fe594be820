"""Golden vectors of the reference's Adams PECE solver (libs/ode.cpp).

oracle/Makefile compiles the reference's ode.cpp from where it lies together with
oracle/ref_ode.cpp (target _ref/ref_ode), which drives it like
CEPCell_HACKATHON/modules/cipa_drug_shago_bench.cpp:119-146 (ode() to tout = dt, 2 dt, ...).
This script runs it on the cases below and stores inputs and outputs in adams.npz.  Run from
the repository root:

    make -C oracle _ref/ref_ode && python tests/golden/make_adams_golden.py
"""
from __future__ import annotations

import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = os.path.join(ROOT, "oracle", "_ref")
OUT = os.path.dirname(os.path.abspath(__file__))
Q = np.loadtxt(os.path.join(OUT, "quinidine_IC50_samples10.csv"), delimiter=",", skiprows=1)

# name, mode, relerr, abserr, t_end, dt_out, every [, conc, bcl, ic50[14]]
CASES = {
    "linear": ("linear", [1e-6, 1e-8, 20.0, 0.1, 1]),
    "vdp": ("vdp", [1e-6, 1e-6, 20.0, 0.1, 1]),
    # the hackathon's tolerances (rel_err = 10E-7, abs_err = 10E-5) and output step 0.1 ms
    "ord11_ctl": ("ord11", [1e-6, 1e-4, 400.0, 0.1, 20, 0.0, 1000.0] + [0.0] * 14),
    "ord11_quin": ("ord11", [1e-6, 1e-4, 400.0, 0.1, 20, 9711.0, 2000.0] + list(Q[0])),
}


def main():
    out = {}
    for key, (mode, inp) in CASES.items():
        a = np.asarray(inp, dtype="<f8")
        fin, fout = os.path.join(REF, f"ode_{key}.in"), os.path.join(REF, f"ode_{key}.out")
        a.tofile(fin)
        subprocess.run([os.path.join(REF, "ref_ode"), mode, fin, fout], check=True)
        out[f"{key}_in"] = a
        out[f"{key}_out"] = np.fromfile(fout, dtype="<f8")
    np.savez_compressed(os.path.join(OUT, "adams.npz"), **out)


if __name__ == "__main__":
    main()
