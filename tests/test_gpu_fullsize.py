"""GPU tests of the batched ensemble and the adjoint at their configured
sizes.

Ensemble (four planar outlines 40 x 20, 60 x 24, 80 x 32 and 100 x 40 km at
1 km patches: 1600 to 8000 DOF, 32 seeded friction draws each, 128 members
in one ensemble): every member runs, and one sampled member per geometry
agrees with the same member run alone through the single-fault FP32
solver; on the two-byte operator every member keeps its events, with first
onsets near the FP32 ensemble's.

Adjoint (the 1 km, 8000-DOF fault, 64 seeded receivers): the device
gradient of the log slip-rate misfit with respect to a, b and D_c of all
4000 patches against the oracle's FP64 reference adjoint (oracle/rsfo_adj.c)
on 16 accepted steps, and against central differences of the device replay
on random directions.
"""
import numpy as np
import pytest

from helpers import SEC_PER_YEAR, initial_state

pytestmark = pytest.mark.gpu

OUTLINES = [(40, 20), (60, 24), (80, 32), (100, 40)]
MEMBERS = 32


def test_ensemble_at_configured_size_matches_single_runs():
    from interact_amd import accuracy, faults, lib
    from interact_amd.model import Ensemble, FaultSystem
    geoms, gdat = [], []
    for (L, W) in OUTLINES:
        g = faults.planar_grid(1.0, L, W)
        fs = FaultSystem(g)
        fs.assemble(lib.SLIP_VECTOR, shear_modulus=32e9, store=lib.PREC_FP32)
        geoms.append(fs)
        gdat.append((g, L, W))
    assert [2 * len(g["l"]) for g, _, _ in gdat] == [1600, 2880, 5120, 8000]
    ens = Ensemble(geoms, MEMBERS)
    fields = []
    for gi, (g, L, W) in enumerate(gdat):
        for k in range(MEMBERS):
            f = faults.ensemble_fields(g, L, W, k, gi)
            m = ens.members[gi * MEMBERS + k]
            m.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"])
            m.set_loading(f["vpl"])
            X0 = initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=2)
            m.set_state(X0)
            fields.append((f, X0))
    t_end = 25.0 * SEC_PER_YEAR
    st = ens.solve(t_end, rtol=1e-8, steps_per_launch=64)
    assert len(st) == 128
    assert all(s["status"] == 0 and s["t"] >= t_end and s["steps"] > 0 for s in st)
    n_events = 0
    for gi in range(4):
        i = gi * MEMBERS + 7 * gi + 3          # one sampled member per geometry
        f, X0 = fields[i]
        single = geoms[gi].member()
        single.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"])
        single.set_loading(f["vpl"])
        single.set_state(X0)
        s1 = single.solve(t_end, rtol=1e-8, prec=lib.PREC_FP32, steps_per_launch=64)
        assert abs(s1["steps"] - st[i]["steps"]) <= 0.02 * s1["steps"] + 2
        o1 = accuracy.onsets(single.events())
        o2 = accuracy.onsets(ens.members[i].events())
        assert len(o1) == len(o2), (i, o1, o2)
        if len(o1):
            assert np.abs(o1 - o2).max() < 1e-3
        n_events += len(o1)
        # the runs end at different times (each steps over t_end on its own
        # steps), so their log10 max|v| monitor traces are compared on a
        # uniform grid of common times instead of the final states
        maxdev, rmsdev = accuracy.trace_deviation_uniform(single.monitor(),
                                                          ens.members[i].monitor())
        assert rmsdev < (0.05 if len(o1) else 1e-4), (i, maxdev, rmsdev)
        single.close()
    # the same members on the two-byte operator (bf16 hi + near-field lo):
    # every member has the FP32 ensemble's event count, and the first onsets
    # differ by a median below 1e-3 yr and a 90th percentile below 2e-2 yr
    # (single members can move further: one of the 128 moves its first event
    # by 2.5 yr over 100 yr, as the single-fault bf16nf solver moves it)
    on32 = [accuracy.onsets(m.events()) for m in ens.members]
    ens.set_operator(lib.PREC_BF16NF)
    for m, (f, X0) in zip(ens.members, fields):
        m.set_state(X0)
    st = ens.solve(t_end, rtol=1e-8, steps_per_launch=64)
    assert all(s["status"] == 0 and s["t"] >= t_end for s in st)
    first = []
    for i, m in enumerate(ens.members):
        oh = accuracy.onsets(m.events())
        assert len(oh) == len(on32[i]), (i, oh, on32[i])
        if len(oh):
            first.append(abs(oh[0] - on32[i][0]))
    assert len(first) >= 32
    assert np.median(first) < 1e-3 and np.percentile(first, 90) < 2e-2, sorted(first)[-5:]
    ens.close()
    for fs in geoms:
        fs.close()


def test_adjoint_at_configured_size_matches_the_oracle():
    import oracle
    from interact_amd import faults, lib
    from interact_amd.autograd import MisfitProblem
    from interact_amd.model import FaultSystem
    g = faults.planar_grid(1.0)
    f = faults.cycle_fields(g)
    n = len(g["l"])
    fs = FaultSystem(g)
    fs.assemble(lib.SLIP_VECTOR, shear_modulus=f["shear_modulus"], store=lib.PREC_FP32)
    fs.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"])
    fs.set_loading(f["vpl"])
    X0 = initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=2)
    fs.set_state(X0)
    fs.set_step_log(100000)
    fs.solve(2.0 * SEC_PER_YEAR, rtol=1e-8, prec=lib.PREC_FP64)
    steps = fs.step_log()[:16]
    assert len(steps) == 16
    fs.set_state(X0)
    recv = np.sort(np.random.default_rng(20261202).choice(n, 64, replace=False)).astype(np.int32)
    trace = fs.replay(steps, recv)["trace"]
    target = trace + 0.01 * np.random.default_rng(3).standard_normal(trace.shape)
    weight = np.full(len(steps) + 1, 1.0 / (len(steps) + 1))
    r = fs.adjoint_gradient(steps, recv, target, weight, prec=lib.PREC_FP64, checkpoint_every=8)
    K = fs.dense()
    xpl = np.zeros(2 * n)
    xpl[0::2] = -f["vpl"]
    P = oracle.RSFProblem(K, f["a"], f["b"], f["dc"], K @ xpl, ncomp=2, eta=f["eta"],
                          v0=f["v0"], f0=f["f0"], nthreads=16)
    Jo, gao, gbo, gdo = P.adjoint(X0, steps, recv, target, weight, 0)
    assert abs(r["J"] - Jo) <= 1e-11 * Jo
    for gd, go in ((r["grad_a"], gao), (r["grad_b"], gbo), (r["grad_dc"], gdo)):
        assert np.linalg.norm(gd - go) <= 1e-8 * np.linalg.norm(go)
    # central differences (eps 1e-5, truncation error O(eps^2)) of the device
    # replay on random directions in (a - b at fixed a, D_c), scaled by the
    # parameters
    pb = MisfitProblem(fs, X0, 0.0, f["a"], steps, recv, target, weight, v0=f["v0"],
                       f0=f["f0"], eta=f["eta"], prec=lib.PREC_FP64)
    amb0 = f["a"] - f["b"]
    J, g_amb, g_dc = pb.value_and_grad(amb0, f["dc"])
    rng = np.random.default_rng(21)
    eps = 1e-5
    for k in range(4):
        d1 = rng.standard_normal(n) * np.abs(amb0)
        d2 = rng.standard_normal(n) * f["dc"]
        fdv = (pb.value(amb0 + eps * d1, f["dc"] + eps * d2) -
               pb.value(amb0 - eps * d1, f["dc"] - eps * d2)) / (2 * eps)
        ad = g_amb @ d1 + g_dc @ d2
        assert abs(fdv - ad) <= 5e-6 * abs(ad), (k, fdv, ad)
    fs.close()
