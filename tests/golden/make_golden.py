"""Generate the golden vectors under tests/golden/ from the reference's own model code.

Runs the binaries that oracle/Makefile builds from the reference's CellML model sources
(target ``ref``, outputs in oracle/_ref/) on inputs drawn here with numpy, and stores
inputs + outputs as small .npz fixtures.  Run from the repository root:

    make -C oracle ref && python tests/golden/make_golden.py          (every fixture)
    make -C oracle ref && python tests/golden/make_golden.py ord11    (the ORd 2011 ones only)

The fixtures pin the C oracle (oracle/*.c) and the HIP kernels; nothing at run time on
the GPU reads the reference tree.
"""
from __future__ import annotations

import importlib
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = os.path.join(ROOT, "oracle", "_ref")
OUT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
workload = importlib.import_module("2ddrugsimulation_amd.workload")

# ALGEBRAIC slots exported by the oracle / kernels for the ORd-CiPA model
ORD_ASLOTS = np.array([0, 12, 36, 42, 44, 46, 48, 50, 53, 54, 58, 60, 66, 83, 84, 87, 90, 96, 98,
                       130, 160, 179, 181, 185, 189, 190, 191, 193, 198, 199])

# reference IC50 samples (first data rows of the reference's fixtures)
BEPRIDIL_ROW = [2704, 0.6954, 0, 0, 50490, 0.6277, 2371, 1.984, 1947, 1.473, 12460, 2.885, 53.1, 1.075]
QUINIDINE_ROW = [47970, 0.6301, 5.751e09, 4.739, 4910, 1.415, 15950, 1.135, 8589, 1.527, 3489, 1.429, 980.9, 0.8176]


def run(binary: str, mode: str, inp: np.ndarray, rec_out: int) -> np.ndarray:
    fin = os.path.join(REF, f"{mode}.in")
    fout = os.path.join(REF, f"{mode}.out")
    np.ascontiguousarray(inp, dtype="<f8").tofile(fin)
    subprocess.run([os.path.join(REF, binary), mode, fin, fout], check=True)
    out = np.fromfile(fout, dtype="<f8")
    return out.reshape(-1, rec_out)


def drug_sets(rng):
    """(conc, herg[19]) tuples: control, seeded drugs, NA entries, huge IC50, reference rows."""
    sets = []
    h1, cmax1 = workload.draw_drug(rng)
    h2, cmax2 = workload.draw_drug(rng)
    sets.append((0.0, h1))                       # control path through log(0)
    sets.append((2.0 * cmax1, h1))
    sets.append((4.0 * cmax2, h2))
    hna = h2.copy(); hna[2:4] = 0.0; hna[8] = 0.0    # "NA" IK1 and INaL
    sets.append((3.0 * cmax2, hna))
    hq = np.zeros(19); hq[:14] = QUINIDINE_ROW; hq[14:] = h1[14:]
    sets.append((6474.0, hq))                    # huge IK1 IC50 (5.751e9 nM)
    hb = np.zeros(19); hb[:14] = BEPRIDIL_ROW; hb[14:] = h2[14:]
    sets.append((33.0, hb))
    return sets


# ORd 2011 ALGEBRAIC slots of the fixtures: the rate-law intermediates and currents the oracle
# stores (gate steady states and time constants 0-48, INa 62, ICaL/ICaNa/ICaK 81-85, Jrel 86-93
# and 191, IKr 96, IKs 98, IK1 100, INaCa 132/162, INaK 181, background and pump 183-188, Jup 196)
# plus assp 34, tiFp 54, tiSp 55, INaL 64 and Ito 70, which the patch-clamp copy never completes
ORD11_ASLOTS = np.array([0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 19, 20, 21, 22,
                         23, 24, 25, 26, 27, 28, 30, 31, 32, 34, 35, 36, 37, 38, 40, 41, 42, 44, 46,
                         48, 54, 55, 62, 64, 70, 81, 82, 85, 86, 87, 92, 93, 96, 98, 100, 132, 162,
                         181, 183, 184, 186, 188, 191, 196])
# DrugSimulation's values of the two constants the patch-clamp copy's ___initConsts leaves unset
# (DrugSimulation/cellmodels/Ohara_Rudy_2011.cpp:538 PKNa, :600 KmCaAct)
ORD11_PKNA, ORD11_KMCAACT = (30, 0.01833), (63, 150e-6)


def ord11_golden():
    """ORd 2011 from the reference's patch-clamp copy of the model (ref_o11): its initConsts
    with the Hill block (endocardial, no Dutta scaling, which that copy applies to other slots
    than DrugSimulation's), and computeRates on perturbed states of the shipped steady state,
    with and without the stimulus."""
    rng = np.random.default_rng(20251205)
    q = np.loadtxt(os.path.join(OUT, "quinidine_IC50_samples10.csv"), delimiter=",", skiprows=1)
    steady = np.loadtxt(os.path.join(OUT, "ord2011_output_orudy.dat"))
    drugs = [(0.0, np.zeros(14))] + [(6474.0, q[k]) for k in range(len(q))]
    drugs.append((9711.0, np.where(np.arange(14) >= 12, 0.0, q[8])))     # "NA" hERG pair
    inp = np.array([np.concatenate([[0.0, c], h, [0.0]]) for c, h in drugs])
    out = run("ref_o11", "o11_init", inp, 142 + 44)
    np.savez_compressed(os.path.join(OUT, "ord11_init.npz"), inp=inp, C=out[:, :142])
    recs, meta = [], []
    times = (5.0, 10.2, 500.0, 1010.3)                 # stimulus off, on, off, on (second beat)
    for d, (c, h) in enumerate(drugs):
        Cc = out[d, :142].copy()
        for k, v in (ORD11_PKNA, ORD11_KMCAACT):
            Cc[k] = v
        for j in range(6):
            st = steady.copy()
            if j:
                st[0] = rng.uniform(-90.0, 40.0)
                st[1:10] *= rng.uniform(0.8, 1.2, 9)
                st[10:39] = rng.uniform(0.0, 1.0, 29)
                st[33] = rng.uniform(0.0, 0.01)
                st[39:41] = rng.uniform(0.0, 0.01, 2)
            t = times[(d + j) % 4]
            recs.append(np.concatenate([[t], Cc, st]))
            meta.append(d)
    recs = np.array(recs)
    out = run("ref_o11", "o11_rhs", recs, 41 + 198)
    np.savez_compressed(os.path.join(OUT, "ord11_rhs.npz"), t=recs[:, 0], C=recs[:, 1:143], S=recs[:, 143:],
                        drug=np.array(meta), conc=np.array([drugs[d][0] for d in meta]),
                        hill=np.array([drugs[d][1] for d in meta]), R=out[:, :41],
                        A=out[:, 41:][:, ORD11_ASLOTS], aslots=ORD11_ASLOTS)


def tt06_init_golden():
    """TT06 endo / epi / M constants and initial states of the compiled reference models."""
    out = run("ref_tt", "tt_init", np.array([0.0, 1.0, 2.0]), 53 + 19)
    np.savez_compressed(os.path.join(OUT, "tt06_init.npz"), C=out[:, :53], S=out[:, 53:])


def main():
    rng = np.random.default_rng(20251203)
    sets = drug_sets(rng)
    # ---------------- ORd-CiPA init + drug block
    inp = []
    for ct in (0, 1, 2):
        for conc, h in sets:
            inp.append(np.concatenate([[ct, conc], h]))
    inp = np.array(inp)
    out = run("ref_ord", "ord_init", inp, 206 + 52)
    np.savez_compressed(os.path.join(OUT, "ord_init.npz"), inp=inp, C=out[:, :206], S=out[:, 206:])

    # ---------------- ORd-CiPA reference trajectories by RK4 (small fixed step)
    traj = {}
    for name, ct, (conc, h) in (("endo_ctl", 0, sets[0]), ("epi_drug", 1, sets[1]), ("m_drug", 2, sets[2])):
        rec = np.concatenate([[ct, conc], h, [1000.0, 0.0005, 1000]])   # 1 beat, h=0.5us, sample 0.5 ms
        y = run("ref_ord", "ord_rk4", rec, 52)
        traj[name] = (ct, conc, h, y)
    np.savez_compressed(os.path.join(OUT, "ord_rk4.npz"),
                        **{f"{k}_inp": np.concatenate([[v[0], v[1]], v[2]]) for k, v in traj.items()},
                        **{f"{k}_y": v[3][::20] for k, v in traj.items()},         # every 10 ms
                        **{f"{k}_vm": v[3][:, 0] for k, v in traj.items()})       # Vm every 0.5 ms

    # ---------------- ORd-CiPA RHS on realistic + perturbed states
    times_ms = [0.0, 3.2, 3.4, 6.0, 30.0, 200.0, 300.0, 700.0]
    recs = []
    for name, (ct0, conc0, h0, y) in traj.items():
        for tm in times_ms:
            s = y[int(round(tm / 0.5))].copy()
            for ct in (0, 1, 2):
                for conc, h in (sets[0], sets[3] if ct != 1 else sets[4], (conc0, h0)):
                    st = s.copy(); st[43] = conc
                    recs.append(np.concatenate([[ct, conc], h, [tm], st]))
                    # perturbed copy: gates stay in (0,1), concentrations positive
                    sp = st.copy()
                    sp[:43] *= np.exp(rng.normal(0.0, 0.02, 43))
                    sp[44:49] *= np.exp(rng.normal(0.0, 0.02, 5))
                    sp[0] = st[0] + rng.normal(0.0, 2.0)
                    recs.append(np.concatenate([[ct, conc], h, [tm + 0.25], sp]))
    recs = np.array(recs)
    out = run("ref_ord", "ord_rhs", recs, 52 + 200)
    np.savez_compressed(os.path.join(OUT, "ord_rhs.npz"), inp=recs, R=out[:, :52],
                        A=out[:, 52:][:, ORD_ASLOTS], aslots=ORD_ASLOTS)

    # ---------------- TT06 Rush-Larsen trajectories (1 beat, dt 0.02) + RHS
    rl = {}
    for ct in (0, 1, 2):
        y = run("ref_tt", "tt_rl", np.array([ct, 1000.0, 0.02, 50]), 19)      # sample every 1 ms
        rl[ct] = y
    np.savez_compressed(os.path.join(OUT, "tt06_rl.npz"), **{f"ct{ct}": y[::5] for ct, y in rl.items()},
                        **{f"ct{ct}_vm": y[:, 0] for ct, y in rl.items()})
    recs = []
    for ct in (0, 1, 2):
        for tm in (0, 5, 10, 10.5, 11, 12, 15, 20, 50, 100, 200, 300, 400, 500, 800):
            s = rl[ct][int(tm)].copy()
            recs.append(np.concatenate([[ct, float(tm)], s]))
            sp = s.copy()
            sp[1:] *= np.exp(rng.normal(0.0, 0.02, 18))
            sp[0] += rng.normal(0.0, 2.0)
            recs.append(np.concatenate([[ct, float(tm) + 0.3], sp]))
        # the GHK singular point V = 15 mV of the TT06 ICaL expression
        s = rl[ct][300].copy(); s[0] = 15.0
        recs.append(np.concatenate([[ct, 300.0], s]))
    recs = np.array(recs)
    out = run("ref_tt", "tt_rhs", recs, 19 + 70)
    np.savez_compressed(os.path.join(OUT, "tt06_rhs.npz"), inp=recs, R=out[:, :19], A=out[:, 19:])
    tt06_init_golden()
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)))


if __name__ == "__main__":
    if sys.argv[1:] == ["ord11"]:
        ord11_golden()
    elif sys.argv[1:] == ["tt06_init"]:
        tt06_init_golden()
    else:
        main()
        ord11_golden()
