"""Pin the C oracle against the reference's compiled model (golden vectors), closed forms
and textbook problems.  CPU only."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O


def _load(golden_dir, name):
    return np.load(f"{golden_dir}/{name}")


def test_ord_init_consts_match_reference(golden_dir):
    g = _load(golden_dir, "ord_init.npz")
    for inp, Cg, Sg in zip(g["inp"], g["C"], g["S"]):
        Cc, S = O.ord_init(inp[0], inp[1], inp[2:])
        np.testing.assert_array_equal(Cc, Cg)
        np.testing.assert_array_equal(S, Sg)


def test_ord_hill_closed_form():
    # initConsts applies g *= 1/(1+(D/IC50)^h); "NA" (0) leaves g unchanged
    herg = np.zeros(19); herg[6], herg[7] = 1000.0, 1.5   # INa
    herg[16] = 1.0
    C0, _ = O.ord_init(0)
    C1, S1 = O.ord_init(0, 2000.0, herg)
    assert C1[45] == pytest.approx(C0[45] / (1 + 2.0 ** 1.5), rel=1e-15)
    assert C1[156] == C0[156] and S1[43] == 2000.0


def test_ord_rhs_match_reference(golden_dir):
    g = _load(golden_dir, "ord_rhs.npz")
    for inp, Rg, Ag in zip(g["inp"], g["R"], g["A"]):
        Cc, _ = O.ord_init(inp[0], inp[1], inp[2:21])
        R, A = O.ord_rates(inp[21], Cc, inp[22:])
        np.testing.assert_array_equal(R, Rg)
        np.testing.assert_array_equal(A[g["aslots"]], Ag)


def test_ord_markov_occupancy_conserved(golden_dir):
    # the 9 non-drug Markov states move probability around: sum of their rates is 0
    g = _load(golden_dir, "ord_rhs.npz")
    for inp in g["inp"][:60]:
        Cc, _ = O.ord_init(inp[0], inp[1], inp[2:21])
        R, _ = O.ord_rates(inp[21], Cc, inp[22:])
        scale = np.abs(R[34:43]).max() + 1e-30
        assert abs(R[34:43].sum()) <= 1e-12 * scale + 1e-18
        assert R[43] == 0.0


def test_tt06_rhs_match_reference(golden_dir):
    g = _load(golden_dir, "tt06_rhs.npz")
    for inp, Rg, Ag in zip(g["inp"], g["R"], g["A"]):
        Cc, _ = O.tt_init(int(inp[0]))
        R, A = O.tt_rates(inp[1], Cc, inp[2:], int(inp[0]))
        np.testing.assert_array_equal(R, Rg)       # includes the reference's NaN at V = 15 mV
        np.testing.assert_array_equal(A, Ag)


@pytest.mark.parametrize("ct", [0, 1, 2])
def test_tt06_rush_larsen_beat_matches_reference(golden_dir, ct):
    g = _load(golden_dir, "tt06_rl.npz")
    cfg = O.make_tt_cfg(celltype=ct, nbeats=1, dt=0.02)
    S, res, vm = O.tt_run(cfg, trace=True)
    np.testing.assert_array_equal(vm[49::50], g[f"ct{ct}_vm"][1:])
    np.testing.assert_array_equal(S, g[f"ct{ct}"][-1])


def test_rush_larsen_exact_for_linear_gate():
    # RL is exact for dx/dt = (xinf - x)/tau with constant xinf, tau: check one gate update
    Cc, S = O.tt_init(0)
    R, A = O.tt_rates(0.0, Cc, S, 0)
    S2 = S.copy(); A2 = A.copy()
    O.lib().tt_rush_larsen(0.02, O.ptr(Cc), O.ptr(R), O.ptr(S2), O.ptr(A2))
    xinf, tau = A[7], A[20]
    assert S2[12] == pytest.approx(xinf + (S[12] - xinf) * np.exp(-0.02 / tau), rel=1e-14)


@pytest.mark.parametrize("name", ["endo_ctl", "epi_drug", "m_drug"])
def test_bdf_beat_matches_reference_rk4(golden_dir, name):
    """Own CVODE-style BDF on the ORd-CiPA model vs the reference model integrated by RK4
    at h = 0.5 us (one beat from the reference initial state)."""
    g = _load(golden_dir, "ord_rk4.npz")
    inp = g[f"{name}_inp"]
    cfg = O.make_ord_cfg(celltype=inp[0], conc=inp[1], herg=inp[2:], nbeats=1, rtol=1e-6, atol=1e-8)
    rc, S, res, vm, st = O.ord_run(cfg, trace=True)
    assert rc == 0
    d = np.abs(vm[4::5] - g[f"{name}_vm"][1:])
    assert d.max() < 0.2          # mV, worst point is in the upstroke
    np.testing.assert_allclose(S[:49], g[f"{name}_y"][-1][:49], rtol=1e-4, atol=1e-6)


def _bdf_solve(f, y0, t_out, rtol, atol, hmax=0.0, maxord=5):
    L = O.lib()
    n = len(y0)

    @O.BDF_RHS
    def cb(t, y, yd, user):
        yy = np.ctypeslib.as_array(y, shape=(n,))
        out = np.ctypeslib.as_array(yd, shape=(n,))
        out[:] = f(t, yy)

    opts = O.BdfOpts(rtol, atol, hmax, 0.0, maxord, 100000)
    m = L.bdf_create(n, cb, None, C.byref(opts))
    y = np.array(y0, dtype=np.float64)
    L.bdf_init(m, 0.0, O.ptr(y))
    outs = []
    for t in t_out:
        rc = L.bdf_advance(m, float(t), O.ptr(y))
        assert rc == 0
        outs.append(y.copy())
    st = O.BdfStats(); L.bdf_get_stats(m, C.byref(st))
    L.bdf_free(m)
    return np.array(outs), st


def test_bdf_linear_closed_form():
    # stiff linear system with eigenvalues -1 and -1000: y = exp(At) y0
    A = np.array([[-1.0, 0.0], [999.0, -1000.0]])
    w, V = np.linalg.eig(A)
    y0 = np.array([1.0, 2.0])
    ts = np.linspace(0.1, 5.0, 50)
    ys, st = _bdf_solve(lambda t, y: A @ y, y0, ts, 1e-8, 1e-10)
    exact = np.array([(V @ np.diag(np.exp(w * t)) @ np.linalg.solve(V, y0)).real for t in ts])
    np.testing.assert_allclose(ys, exact, rtol=2e-6, atol=1e-8)
    assert st.nje < st.nst     # Jacobian reuse


def test_bdf_robertson():
    # Robertson chemical kinetics (textbook stiff problem); conservation y1+y2+y3 = 1 and
    # the well-known value y1(40) = 0.7158
    def f(t, y):
        return np.array([-0.04 * y[0] + 1e4 * y[1] * y[2],
                         0.04 * y[0] - 1e4 * y[1] * y[2] - 3e7 * y[1] ** 2,
                         3e7 * y[1] ** 2])
    ys, _ = _bdf_solve(f, [1.0, 0.0, 0.0], [0.4, 4.0, 40.0], 1e-8, 1e-12)
    assert ys[-1][0] == pytest.approx(0.7158271, abs=2e-6)
    assert ys[-1].sum() == pytest.approx(1.0, abs=1e-7)


def test_bdf_tolerance_convergence():
    # error shrinks with tolerance (order behaviour) on y' = -y + sin(t)
    f = lambda t, y: np.array([-y[0] + np.sin(t)])
    exact = lambda t: 0.5 * (np.sin(t) - np.cos(t)) + 1.5 * np.exp(-t)
    errs = []
    for tol in (1e-4, 1e-6, 1e-8):
        ys, _ = _bdf_solve(f, [1.0], [10.0], tol, tol * 1e-2)
        errs.append(abs(ys[-1][0] - exact(10.0)))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-6


def test_bdf_hmax_respected():
    # hmax caps the step: with hmax = dt the number of steps >= span / dt
    ys, st = _bdf_solve(lambda t, y: -0.001 * y, [1.0], [10.0], 1e-6, 1e-8, hmax=0.1)
    assert st.nst >= 100
