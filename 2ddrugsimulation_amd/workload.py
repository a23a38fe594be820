"""Seeded synthetic workloads for the drug-population path.

Parameter conventions follow the reference drivers:
  * ``herg[0..13]`` are (IC50 nM, Hill) pairs read from the IC50 sample CSV in the order
    ICaL, IK1, IKs, INa, INaL, Ito, hERG (CEP7Cell/bin/drugs/*/IC50_samples.csv header;
    mapping applied in CEP5Cell/cellmodels/ohara_rudy_cipa_v1_2017.cpp:826-837; the
    hERG pair is not used by the Markov IKr model).  An IC50 or Hill of 0 ("NA") means
    no block.
  * ``herg[14..18]`` are the dynamic-hERG parameters Kmax, Ku, n, halfmax, Vhalf
    (ohara_rudy_cipa_v1_2017.cpp:838-842); the trapping rate Kt stays 0 as in every
    reference run.
  * concentrations are in nM and are multiples of Cmax (CEP5Cell/modules/drugs_herg.cpp:149-156).
The random draws use ``numpy.random.default_rng(seed)`` (PCG64); the device code never
re-implements the generator, arrays are drawn here and uploaded.
"""
from __future__ import annotations

import numpy as np

ORD_NSTATE = 52
TT_NSTATE = 19

# population scale slots (order of the gscale arrays)
ORD_GSCALE_NAMES = ("GNa", "GNaL", "PCa", "GKr", "GKs", "GK1", "Gto")
TT_GSCALE_NAMES = ("GNa", "GK1", "GKr", "GKs", "Gto", "GCaL", "GbNa", "GbCa", "GpCa", "GpK")


def draw_drug(rng: np.random.Generator) -> tuple[np.ndarray, float]:
    """One synthetic drug: herg[19] and its Cmax (nM)."""
    herg = np.zeros(19)
    for k in range(7):
        herg[2 * k] = 10.0 ** rng.uniform(2.0, 5.0)      # IC50 0.1-100 uM, in nM
        herg[2 * k + 1] = rng.uniform(0.5, 2.0)          # Hill
    n = rng.uniform(0.5, 1.5)
    herg[14] = 10.0 ** rng.uniform(0.0, 8.0)             # Kmax
    herg[15] = 10.0 ** rng.uniform(-5.0, 0.0)            # Ku (1/ms)
    herg[16] = n                                         # n
    herg[17] = (10.0 ** rng.uniform(2.0, 5.0)) ** n      # halfmax (nM^n)
    herg[18] = rng.uniform(-200.0, -1.0)                 # Vhalf (mV)
    cmax = 10.0 ** rng.uniform(0.0, 3.0)
    return herg, cmax


def ord_population(seed: int, ndrugs: int, nconc: int, nvariants: int,
                   bcl_choices=(500.0, 1000.0, 2000.0), sigma: float = 0.2,
                   celltype: int = 0) -> dict:
    """drugs x conc x variants ORd-CiPA cells (10,000 = 50 x 4 x 50 for the batched config)."""
    rng = np.random.default_rng(seed)
    drugs = [draw_drug(rng) for _ in range(ndrugs)]
    n = ndrugs * nconc * nvariants
    herg = np.zeros((n, 19))
    conc = np.zeros(n)
    gscale = np.exp(rng.normal(0.0, sigma, size=(ndrugs, nvariants, 7)))
    bcl = rng.choice(np.asarray(bcl_choices, dtype=float), size=n)
    k = 0
    for d, (h, cmax) in enumerate(drugs):
        for c in range(nconc):
            for v in range(nvariants):
                herg[k] = h
                conc[k] = (c + 1) * cmax
                k += 1
    gs = np.repeat(gscale[:, None, :, :], nconc, axis=1).reshape(n, 7)
    return {"celltype": np.full(n, float(celltype)), "conc": conc, "herg": herg,
            "gscale": gs, "bcl": bcl}


def ord_single_drug(seed: int = 20251203, celltype: int = 0) -> dict:
    """One drug, control + 4 concentrations at 1-4 x Cmax (the CPU-runnable config)."""
    rng = np.random.default_rng(seed)
    h, cmax = draw_drug(rng)
    concs = np.array([0.0, cmax, 2 * cmax, 3 * cmax, 4 * cmax])
    n = len(concs)
    return {"celltype": np.full(n, float(celltype)), "conc": concs,
            "herg": np.repeat(h[None, :], n, axis=0), "gscale": np.ones((n, 7)),
            "bcl": np.full(n, 1000.0), "cmax": cmax}


def tt_population(seed: int, ncell: int, sigma: float = 0.25, bcl: float = 1000.0) -> dict:
    """TT06 cells: celltype uniform over endo/epi/M, Hill block of IKr/ICaL/INa at
    1-4 x Cmax, ten lognormal conductance scales."""
    rng = np.random.default_rng(seed)
    celltype = rng.integers(0, 3, size=ncell).astype(float)
    ic50 = 10.0 ** rng.uniform(2.0, 5.0, size=(ncell, 3))
    hill = rng.uniform(0.5, 2.0, size=(ncell, 3))
    cmax = 10.0 ** rng.uniform(0.0, 3.0, size=ncell)
    conc = cmax * rng.integers(1, 5, size=ncell)
    gscale = np.exp(rng.normal(0.0, sigma, size=(ncell, 10)))
    return {"celltype": celltype, "conc": conc, "ic50": ic50, "hill": hill,
            "gscale": gscale, "bcl": np.full(ncell, bcl)}


def group_slots(bcl: np.ndarray, per_group: dict | None = None, group: int = 16) -> np.ndarray:
    """Slot layout of an ORd-CiPA population for the BDF kernel, whose wavefronts hold `group`
    cells each: cells ordered by BCL (longest first), each BCL class filling groups of `group`
    slots with `per_group[bcl]` cells (default: all `group`), the rest of a group padding.
    Returns the cell of each slot, -1 for padding (to be parked with ``park_cells``).

    A wavefront's pass costs nearly the same with fewer cells, but a group of long beats with
    fewer cells meets fewer cells' rare events (factorisations, Jacobians) per pass, and the
    long beats set a launch's time; padding the long-beat groups trades slots for time while the
    groups still all start at once."""
    per_group = per_group or {}
    slots = []
    for b in sorted(set(np.asarray(bcl, dtype=float).tolist()), reverse=True):
        idx = np.nonzero(bcl == b)[0]
        c = int(per_group.get(b, group))
        if not 1 <= c <= group:
            raise ValueError(f"cells per group for BCL {b}: {c}")
        for g0 in range(0, len(idx), c):
            part = list(idx[g0:g0 + c])
            slots += part + [-1] * (group - len(part))
    return np.array(slots, dtype=np.int64)
