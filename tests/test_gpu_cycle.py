"""GPU tests of the benchmark cycle configurations at their full sizes.

c1 (the CPU-runnable configuration: planar 100 x 40 km fault at 2 km
patches, 1000 patches x 2 slip components = 2000 DOF, 300 yr, rk5dp rtol
1e-8): the device cycle in FP64, FP32 and the bf16 near-field format against
the FP64 oracle's cycle, scored with the reference's trajectory metrics
(test_hmatrix/rsf_accuracy.py:65-99): onsets, recurrence, peak slip rates
and the log10 max|v| trace.

c3 (0.5 km patches, 16000 patches x 2 = 32000 DOF, an 8.2 GB FP64 operator
whose row offsets pass 2^31 bytes): sampled entries against the oracle,
products on row blocks against host products of the same rows, along-strike
translation invariance, FP32 entries equal to the rounded FP64 entries, the
bf16 near-field format's product error, and a cycle segment through the
first event in FP64 and bf16nf.
"""
import numpy as np
import pytest

import oracle
from helpers import SEC_PER_YEAR, vector_operator

pytestmark = pytest.mark.gpu


def _planar(ds, store):
    from interact_amd import faults, lib
    from interact_amd.model import FaultSystem, initial_state
    g = faults.planar_grid(ds)
    f = faults.cycle_fields(g)
    fs = FaultSystem(g)
    fs.assemble(lib.SLIP_VECTOR, shear_modulus=f["shear_modulus"], store=store)
    fs.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"])
    fs.set_loading(f["vpl"])
    X0 = initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=2)
    return g, f, fs, X0


# ------------------------------------------------------------------ c1
def test_c1_cycle_trajectories_match_the_oracle():
    """300 yr of c1 on the device (FP64, FP32, bf16nf) against the oracle's
    FP64 cycle: the same events, every onset within 0.05 yr (FP64: 1e-4 yr),
    peak slip rates within 1% (FP64: 1e-4), the log10 max|v| trace of
    the FP64 device run within 0.01 RMS of the oracle's, and the FP64
    run's catalog rows (ruptured cells and area exactly, the rest 1e-4)."""
    from interact_amd import accuracy, lib
    store = lib.PREC_FP32 | lib.PREC_BF16NF
    g, f, fs, X0 = _planar(2.0, store)
    n = len(g["l"])
    T = oracle.traction_tensor(g, 0, 16)
    K = vector_operator(T, f["shear_modulus"] / 1e4)
    # the oracle's operator is the device's FP64 operator
    assert np.abs(fs.dense() - K).max() <= 1e-12 * np.abs(K).max()
    xpl = np.zeros(2 * n)
    xpl[0::2] = -f["vpl"]
    P = oracle.RSFProblem(K, f["a"], f["b"], f["dc"], K @ xpl, ncomp=2, eta=f["eta"],
                          v0=f["v0"], f0=f["f0"], nthreads=16)
    t_end = 300 * SEC_PER_YEAR
    ro = P.solve(X0, t_end, rtol=1e-8, cat_area=4.0 * g["l"] * g["w"],
                 cat_mu=f["shear_modulus"])
    assert ro["status"] == 0
    ref_ev, ref_mon = ro["events"], ro["monitor"]
    assert len(accuracy.onsets(ref_ev)) == 2
    scores = {}
    for name, prec in (("fp64", lib.PREC_FP64), ("fp32", lib.PREC_FP32),
                       ("bf16nf", lib.PREC_BF16NF)):
        fs.set_state(X0)
        st = fs.solve(t_end, rtol=1e-8, prec=prec, steps_per_launch=64)
        assert st["status"] == 0 and st["events_dropped"] == 0
        s = accuracy.score(ref_mon, ref_ev, fs.monitor(), fs.events())
        scores[name] = s
        if name == "fp64":
            # catalog rows (rsf_finalize_event): the same ruptured cells
            cd, co = fs.catalog(), ro["catalog"]
            assert len(cd) == len(co) == 2
            assert np.array_equal(cd[:, 2:4], co[:, 2:4])
            assert np.abs(cd[:, :2] - co[:, :2]).max() / SEC_PER_YEAR < 1e-4
            for k in (4, 5, 6, 7, 8, 9):
                assert np.abs(cd[:, k] / co[:, k] - 1.0).max() < 1e-4, k
        assert s["nev_run"] == s["nev_ref"], (name, s)
        tol_on, tol_pk = (1e-4, 1e-4) if name == "fp64" else (0.05, 0.01)
        assert max(abs(d) for d in s["donset"]) < tol_on, (name, s)
        assert max(abs(d) for d in s["dpeak"]) < tol_pk, (name, s)
    assert scores["fp64"]["rmsdev"] < 0.01, scores["fp64"]
    print({k: {q: v[q] for q in ("maxdev", "rmsdev", "maxdev_uniform", "rmsdev_uniform", "dt1",
                                 "drec", "donset", "dpeak")}
           for k, v in scores.items()})
    fs.close()


# ------------------------------------------------------------------ c3
@pytest.fixture(scope="module")
def c3():
    from interact_amd import lib
    g, f, fs, X0 = _planar(0.5, lib.PREC_FP32 | lib.PREC_BF16NF)
    yield g, f, fs, X0
    fs.close()


def test_c3_operator_entries_and_products(c3):
    """Sampled (receiver, source) tractions against the oracle; products on
    row blocks (some past the 2^31-byte row offset of the FP64 and FP32
    arrays) against host products of the device's entries; columns of the
    FP32 product equal the rounded FP64 entries; along-strike translation
    invariance of the planar mesh."""
    from interact_amd import lib
    g, f, fs, X0 = c3
    n = len(g["l"])
    D = 2 * n
    rows, cols, ld = fs.dims()
    assert rows == cols == D == 32000
    assert 8 * ld * (D - 1) > 2 ** 31 and 4 * ld * (D - 1) > 2 ** 31
    rng = np.random.default_rng(3)
    # oracle tractions on a sub-mesh of sampled patches (all pairs among them)
    pick = np.sort(rng.choice(n, 96, replace=False))
    sub = {k: g[k][pick] for k in ("xyz", "strike", "dip", "l", "w")}
    Ks = vector_operator(oracle.traction_tensor(sub, 0, 16), f["shear_modulus"] / 1e4)
    dof = np.stack([2 * pick, 2 * pick + 1], 1).ravel()
    Kd = fs.entries(dof, dof)
    assert np.abs(Kd - Ks).max() <= 1e-12 * np.abs(Ks).max()
    # row blocks: start, past 2^31 bytes in FP64, past 2^31 in FP32, the end
    x = np.sin(np.arange(D) + 1.0) + 0.3 * np.cos(3.0 * np.arange(D))
    y64 = fs.matvec(x, lib.PREC_FP64)
    y32 = fs.matvec(x, lib.PREC_FP32)
    for r0 in (0, 8400, 16800, D - 64):
        rb = np.arange(r0, r0 + 64)
        Kr = fs.entries(rb, np.arange(D))
        yr = Kr @ x
        assert np.abs(y64[rb] - yr).max() <= 1e-12 * np.abs(yr).max(), r0
        y32r = Kr.astype(np.float32).astype(np.float64) @ x.astype(np.float32)
        assert np.abs(y32[rb] - y32r).max() <= 1e-6 * np.abs(yr).max(), r0
    # FP32 entries = FP64 entries rounded to nearest: a unit input returns
    # one column exactly (a single nonzero product per row)
    for j in (5, 16001, D - 1):
        e = np.zeros(D)
        e[j] = 1.0
        col64 = fs.entries(np.arange(D), [j])[:, 0]
        assert np.array_equal(fs.matvec(e, lib.PREC_FP32),
                              col64.astype(np.float32).astype(np.float64)), j
    # translation invariance: patch k = i nz + j; shifting receiver and
    # source along strike by s columns leaves the entry unchanged
    nz = g["nz"]
    for (i, j, di, dj) in ((10, 5, 3, 7), (50, 70, -20, 0), (120, 0, 1, 79)):
        for s in (1, 17, 40):
            a = fs.entries([2 * (i * nz + j), 2 * (i * nz + j) + 1],
                           [2 * ((i + di) * nz + dj), 2 * ((i + di) * nz + dj) + 1])
            b = fs.entries([2 * ((i + s) * nz + j), 2 * ((i + s) * nz + j) + 1],
                           [2 * ((i + s + di) * nz + dj), 2 * ((i + s + di) * nz + dj) + 1])
            assert np.abs(a - b).max() <= 1e-12 * np.abs(Kd).max()


def test_c3_bf16nf_product_error(c3):
    """The near-field bf16 format's product error on the reference's test
    vector x_i = sin(i+1) + 0.3 cos(3i) (src/compress_interaction_matrix.c:
    68-77) and on the plate-rate deviation pattern: below 2e-5 relative,
    with at most 2.2 bytes per entry."""
    from interact_amd import lib
    g, f, fs, X0 = c3
    D = 2 * len(g["l"])
    i = np.arange(D)
    for x in (np.sin(i + 1.0) + 0.3 * np.cos(3.0 * i),
              np.where(i % 2 == 0, 1e-3 * np.random.default_rng(1).standard_normal(D), 0.0)):
        y64 = fs.matvec(x, lib.PREC_FP64)
        ynf = fs.matvec(x, lib.PREC_BF16NF)
        assert np.linalg.norm(ynf - y64) <= 2e-5 * np.linalg.norm(y64)
    assert fs.operator_bytes(lib.PREC_BF16NF) <= 2.2 * D * D


def test_c3_cycle_segment_fp64_and_bf16nf(c3):
    """The first 34 yr of c3 (through the first event near 32.9 yr) in FP64
    and in bf16nf: one event each, onsets within 0.05 yr, peak slip rates
    within 1%, the bf16nf run chopped into continued segments."""
    from interact_amd import accuracy, lib
    g, f, fs, X0 = c3
    t_end = 34.0 * SEC_PER_YEAR
    fs.set_state(X0)
    st = fs.solve(t_end, rtol=1e-8, steps_per_launch=64)
    e64 = fs.events()
    fs.set_state(X0)
    fs.solve(t_end / 2, rtol=1e-8, prec=lib.PREC_BF16NF, steps_per_launch=64)
    st2 = fs.solve_continue(t_end, rtol=1e-8, prec=lib.PREC_BF16NF, steps_per_launch=64)
    enf = fs.events()
    assert st["status"] == 0 and st2["status"] == 0
    o64, onf = accuracy.onsets(e64), accuracy.onsets(enf)
    assert len(o64) == len(onf) == 1 and 32.0 < o64[0] < 34.0
    assert abs(onf[0] - o64[0]) < 0.05
    p64, pnf = accuracy.event_peaks(e64), accuracy.event_peaks(enf)
    assert abs(pnf[0, 1] / p64[0, 1] - 1.0) < 0.01
