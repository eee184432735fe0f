"""GPU tests of the batched ensemble (rsfg_ensemble_*): member handles,
every member against the FP64 oracle, the tile-shape tuner and the
argument checks.  The ensemble has no counterpart in the reference (which
runs one fault per process, src/rsf_solve.c); its members are checked
against the oracle's restatement of that single run."""
import numpy as np
import pytest

import oracle
from helpers import SEC_PER_YEAR, initial_state, strike_operator, vector_operator

pytestmark = pytest.mark.gpu

MU = 32e9
OUTLINES = [(40, 20), (60, 24)]     # km, two uneven geometries


def _geoms(ds, mode, store=None, outlines=OUTLINES):
    from interact_amd import faults, lib
    from interact_amd.model import FaultSystem
    store = lib.PREC_FP32 if store is None else store
    out = []
    for (L, W) in outlines:
        g = faults.planar_grid(ds, L, W)
        fs = FaultSystem(g)
        fs.assemble(mode, shear_modulus=MU, store=store)
        out.append((fs, g, L, W))
    return out


def _configure(m, f, ncomp):
    m.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"])
    m.set_loading(f["vpl"])
    X0 = initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=ncomp)
    m.set_state(X0)
    return X0


def _onsets(ev):
    return ev[ev[:, 1] > 0, 0] / SEC_PER_YEAR if len(ev) else np.zeros(0)


def _oracle_run(g, f, ncomp, X0, t_end):
    T = oracle.traction_tensor(g, 0, 16)
    K = vector_operator(T, MU / 1e4) if ncomp == 2 else strike_operator(T, MU / 1e4)
    n = len(f["a"])
    xpl = np.zeros(ncomp * n)
    xpl[0::ncomp] = -f["vpl"]
    P = oracle.RSFProblem(K, f["a"], f["b"], f["dc"], K @ xpl, ncomp=ncomp, eta=f["eta"],
                          v0=f["v0"], f0=f["f0"], nthreads=16)
    return P.solve(X0, t_end, rtol=1e-6)


def test_member_handle_matches_a_fresh_handle():
    """A member borrows its geometry's operator and is otherwise a handle of
    its own: the same inputs give bitwise the same run as a fresh handle
    with its own assembled operator."""
    from interact_amd import faults, lib
    from interact_amd.model import FaultSystem
    (fs, g, L, W), = _geoms(2.0, lib.SLIP_VECTOR, outlines=[(40, 20)])
    f = faults.ensemble_fields(g, L, W, 2)
    m = fs.member()
    other = FaultSystem(g)
    other.assemble(lib.SLIP_VECTOR, shear_modulus=MU, store=lib.PREC_FP32)
    _configure(m, f, 2)
    _configure(other, f, 2)
    t_end = 13.0 * SEC_PER_YEAR
    s1 = m.solve(t_end, rtol=1e-6, prec=lib.PREC_FP32)
    s2 = other.solve(t_end, rtol=1e-6, prec=lib.PREC_FP32)
    assert s1["steps"] == s2["steps"] and s1["rejects"] == s2["rejects"]
    X1, t1 = m.get_state()
    X2, t2 = other.get_state()
    assert t1 == t2 and np.array_equal(X1, X2)
    assert np.array_equal(m.events(), other.events()) and len(_onsets(m.events())) > 0
    # assembling through a member is refused (it does not own the operator)
    with pytest.raises(lib.RsfgError):
        m.assemble(lib.SLIP_VECTOR, shear_modulus=MU)
    m.close()
    other.close()
    fs.close()


@pytest.mark.parametrize("ncomp", [2, 1])
def test_ensemble_members_match_the_oracle(ncomp):
    """Two uneven geometries at 2 km (200 and 360 patches) x 16 friction
    draws, 40 yr at rtol 1e-6: each member's event onsets match the FP64
    oracle run of the same member within 2e-3 yr (the FP32 operator and the
    matrix-core summation order are the only differences) and its step
    count within 5 %."""
    from interact_amd import faults, lib
    from interact_amd.model import Ensemble
    mode = lib.SLIP_VECTOR if ncomp == 2 else lib.SLIP_STRIKE
    geoms = _geoms(2.0, mode)
    nm = 16
    ens = Ensemble([q[0] for q in geoms], nm)
    runs = []
    for gi, (fs, g, L, W) in enumerate(geoms):
        for k in range(nm):
            f = faults.ensemble_fields(g, L, W, k, gi)
            X0 = _configure(ens.members[gi * nm + k], f, ncomp)
            runs.append((g, f, X0))
    t_end = 40.0 * SEC_PER_YEAR
    st = ens.solve(t_end, rtol=1e-6)
    n_events = 0
    for i, (g, f, X0) in enumerate(runs):
        assert st[i]["status"] == 0 and st[i]["t"] >= t_end
        ro = _oracle_run(g, f, ncomp, X0, t_end)
        on_o = _onsets(ro["events"])
        on_e = _onsets(ens.members[i].events())
        assert len(on_e) == len(on_o), (i, on_e, on_o)
        if len(on_o):
            assert np.abs(on_e - on_o).max() < 2e-3, (i, on_e, on_o)
        n_events += len(on_o)
        assert abs(st[i]["steps"] - ro["steps"]) <= 0.05 * ro["steps"] + 5, \
            (i, st[i]["steps"], ro["steps"])
    assert n_events >= 8
    ens.close()
    for q in geoms:
        q[0].close()


def test_ensemble_matches_single_fp32_runs_and_tuner():
    """The tuner times every candidate tile shape and keeps the fastest;
    after tuning, every member of a 32-member ensemble on two geometries
    reproduces the single-fault FP32 solver (K v on the vector units) over
    a window with events: same number of onsets, within 1e-4 yr, and step
    counts within 3 %."""
    from interact_amd import faults, lib
    from interact_amd.model import Ensemble
    geoms = _geoms(2.0, lib.SLIP_VECTOR)
    nm = 32
    ens = Ensemble([q[0] for q in geoms], nm)
    runs = []
    for gi, (fs, g, L, W) in enumerate(geoms):
        for k in range(nm):
            f = faults.ensemble_fields(g, L, W, k, gi)
            X0 = _configure(ens.members[gi * nm + k], f, 2)
            runs.append((gi, f, X0))
    choice, ms, shapes = ens.tune(rtol=1e-6)
    assert len(shapes) == len(ms) and all(len(s) == 7 for s in shapes)
    timed = [t for t in ms if t > 0]
    assert len(timed) >= 4 and ms[choice] == min(timed)
    assert ens.config() == choice
    pm, pe = ens.bench(5, rtol=1e-6)
    assert 0 < pm < ms[choice] and 0 < pe < ms[choice]
    for m, (gi, f, X0) in zip(ens.members, runs):
        m.set_state(X0)
    t_end = 20.0 * SEC_PER_YEAR
    st = ens.solve(t_end, rtol=1e-6, steps_per_launch=16)
    for i in range(0, len(runs), 5):
        gi, f, X0 = runs[i]
        s = geoms[gi][0].member()
        _configure(s, f, 2)
        s1 = s.solve(t_end, rtol=1e-6, prec=lib.PREC_FP32)
        on1, on2 = _onsets(s.events()), _onsets(ens.members[i].events())
        assert len(on1) == len(on2)
        if len(on1):
            assert np.abs(on1 - on2).max() < 1e-4
        assert abs(s1["steps"] - st[i]["steps"]) <= 0.03 * s1["steps"] + 3
        s.close()
    ens.close()
    for q in geoms:
        q[0].close()


@pytest.mark.parametrize("nm,outlines", [(32, OUTLINES), (64, OUTLINES[:1])])
def test_every_unit_shape_gives_the_same_run(nm, outlines):
    """Every candidate product shape that fits the member count, selected
    without tuning, on two geometries x 32 members (and one geometry x 64)
    over 15 yr.  Shapes with 1024-column units sum each operator row in the
    same order (FP32 over 16-column chunks from the unit's start, flushed to
    FP64 every 256 columns, blocks added in order), whichever rows a wave
    holds, how many chunks it loads at once, how deep its load ring is and
    whether the unit's waves share the slip rates through LDS, so their runs
    are bitwise the same; the other column-block widths and the stream-K
    shapes (equal shares of the product per workgroup, cut at row-group
    boundaries) change only the FP64 partial sums (onsets within 1e-4 yr)."""
    from interact_amd import faults, lib
    from interact_amd.model import Ensemble
    geoms = _geoms(2.0, lib.SLIP_VECTOR, outlines=outlines)
    ens = Ensemble([q[0] for q in geoms], nm)
    X0s = []
    for gi, (fs, g, L, W) in enumerate(geoms):
        for k in range(nm):
            X0s.append(_configure(ens.members[gi * nm + k], faults.ensemble_fields(g, L, W, k, gi), 2))
    shapes = ens.shapes()
    assert any(s[5] == 1 for s in shapes) and any(s[4] >= 4 for s in shapes)
    assert any(s[6] > 0 and s[3] == 0 for s in shapes)
    t_end = 15.0 * SEC_PER_YEAR
    ref = None
    ntried = 0
    for cfg, sh in enumerate(shapes):
        try:
            ens.select(cfg)
        except lib.RsfgError as err:
            assert "does not fit" in str(err)   # register budget of the member count
            continue
        ntried += 1
        assert ens.config() == cfg
        for m, X0 in zip(ens.members, X0s):
            m.set_state(X0)
        st = ens.solve(t_end, rtol=1e-6)
        out = [(m.get_state()[0], m.events(), q["steps"]) for m, q in zip(ens.members, st)]
        if ref is None:
            assert sh[3] == 1024
            ref = out
            assert sum(len(_onsets(ev)) for _, ev, _ in out) >= 8
            continue
        for i, ((X, ev, n), (Xr, evr, nr)) in enumerate(zip(out, ref)):
            if sh[3] == 1024:
                assert n == nr and np.array_equal(X, Xr) and np.array_equal(ev, evr), (cfg, sh, i)
            else:
                on, onr = _onsets(ev), _onsets(evr)
                assert len(on) == len(onr), (cfg, sh, i)
                if len(on):
                    assert np.abs(on - onr).max() < 1e-4, (cfg, sh, i)
    assert ntried >= (8 if nm == 32 else 3)
    with pytest.raises(lib.RsfgError, match="out of range"):
        ens.select(len(shapes))
    ens.close()
    for q in geoms:
        q[0].close()


def test_ensemble_of_64_members():
    """64 members on one geometry (four 16-member MFMA blocks), after
    tuning: every eighth member reproduces the single-fault FP32 solver's
    onsets within 1e-4 yr and its step count within 3 %."""
    from interact_amd import faults, lib
    from interact_amd.model import Ensemble
    geoms = _geoms(2.0, lib.SLIP_VECTOR, outlines=[(40, 20)])
    fs, g, L, W = geoms[0]
    nm = 64
    ens = Ensemble([fs], nm)
    runs = []
    for k in range(nm):
        f = faults.ensemble_fields(g, L, W, k)
        runs.append((f, _configure(ens.members[k], f, 2)))
    choice, ms, shapes = ens.tune(rtol=1e-6)
    assert ms[choice] > 0
    for m, (f, X0) in zip(ens.members, runs):
        m.set_state(X0)
    t_end = 20.0 * SEC_PER_YEAR
    st = ens.solve(t_end, rtol=1e-6)
    assert all(q["status"] == 0 for q in st)
    for i in range(0, nm, 8):
        f, X0 = runs[i]
        s = fs.member()
        _configure(s, f, 2)
        s1 = s.solve(t_end, rtol=1e-6, prec=lib.PREC_FP32)
        on1, on2 = _onsets(s.events()), _onsets(ens.members[i].events())
        assert len(on1) == len(on2)
        if len(on1):
            assert np.abs(on1 - on2).max() < 1e-4
        assert abs(s1["steps"] - st[i]["steps"]) <= 0.03 * s1["steps"] + 3
        s.close()
    ens.close()
    fs.close()


@pytest.mark.parametrize("shape", [0, 18])
def test_direct_launches_match_graph_launches(monkeypatch, shape):
    """RSFG_NO_GRAPH=1 launches each batch's kernels directly instead of
    replaying the captured graph (for tracing tools): the same kernels in the
    same order, so bitwise the same states, events and step counts, with a
    unit shape (0) and a stream-K shape (18), whose work split is rebuilt
    when a geometry's members have all halted."""
    from interact_amd import faults, lib
    from interact_amd.model import Ensemble
    geoms = _geoms(2.0, lib.SLIP_VECTOR)
    nm = 16
    ens = Ensemble([q[0] for q in geoms], nm)
    X0s = []
    for gi, (fs, g, L, W) in enumerate(geoms):
        for k in range(nm):
            X0s.append(_configure(ens.members[gi * nm + k], faults.ensemble_fields(g, L, W, k, gi), 2))
    ens.select(shape)
    out = {}
    for mode in ("graph", "direct"):
        if mode == "direct":
            monkeypatch.setenv("RSFG_NO_GRAPH", "1")
        for m, X0 in zip(ens.members, X0s):
            m.set_state(X0)
        st = ens.solve(8.0 * SEC_PER_YEAR, rtol=1e-6)
        out[mode] = [(m.get_state()[0], m.events(), q["steps"]) for m, q in zip(ens.members, st)]
    monkeypatch.delenv("RSFG_NO_GRAPH")
    assert sum(len(_onsets(ev)) for _, ev, _ in out["graph"]) >= 4
    for (X, ev, n), (Xd, evd, nd) in zip(out["graph"], out["direct"]):
        assert n == nd and np.array_equal(X, Xd) and np.array_equal(ev, evd)
    ens.close()
    for q in geoms:
        q[0].close()


def test_ensemble_rk3bs_matches_single_runs():
    """The rk3bs pair in the batched stages (four-stage tableau, its own
    controller exponent): 16 members on one geometry over 20 yr reproduce
    the single-fault FP32 rk3bs runs' onsets within 1e-4 yr and their step
    counts within 3 %."""
    from interact_amd import faults, lib
    from interact_amd.model import Ensemble
    geoms = _geoms(2.0, lib.SLIP_VECTOR, outlines=[(40, 20)])
    fs, g, L, W = geoms[0]
    nm = 16
    ens = Ensemble([fs], nm)
    runs = []
    for k in range(nm):
        f = faults.ensemble_fields(g, L, W, k)
        runs.append((f, _configure(ens.members[k], f, 2)))
    t_end = 20.0 * SEC_PER_YEAR
    st = ens.solve(t_end, rtol=1e-6, method=lib.RK3BS)
    assert all(q["status"] == 0 and q["t"] >= t_end for q in st)
    n_on = 0
    for i in range(0, nm, 3):
        f, X0 = runs[i]
        s = fs.member()
        _configure(s, f, 2)
        s1 = s.solve(t_end, rtol=1e-6, prec=lib.PREC_FP32, method=lib.RK3BS)
        on1, on2 = _onsets(s.events()), _onsets(ens.members[i].events())
        assert len(on1) == len(on2), (i, on1, on2)
        if len(on1):
            assert np.abs(on1 - on2).max() < 1e-4, (i, on1, on2)
        n_on += len(on1)
        assert abs(s1["steps"] - st[i]["steps"]) <= 0.03 * s1["steps"] + 3, \
            (i, s1["steps"], st[i]["steps"])
        s.close()
    assert n_on >= 3
    ens.close()
    fs.close()


def test_ensemble_argument_checks():
    from interact_amd import faults, lib
    from interact_amd.model import Ensemble
    geoms = _geoms(2.0, lib.SLIP_VECTOR, outlines=[(40, 20)])
    fs, g, L, W = geoms[0]
    with pytest.raises(lib.RsfgError, match="16, 32 or 64"):
        Ensemble([fs], 24)
    only64 = _geoms(2.0, lib.SLIP_VECTOR, store=lib.PREC_FP64, outlines=[(40, 20)])[0][0]
    with pytest.raises(lib.RsfgError, match="FP32 operator"):
        Ensemble([only64], 16)
    ens = Ensemble([fs], 16)
    # members without friction, loading and state
    with pytest.raises(lib.RsfgError, match="needs friction"):
        ens.solve(SEC_PER_YEAR, rtol=1e-6)
    for k in range(16):
        _configure(ens.members[k], faults.ensemble_fields(g, L, W, k), 2)
    with pytest.raises(lib.RsfgError, match="prec must name"):
        ens.solve(SEC_PER_YEAR, rtol=1e-6, prec=lib.PREC_FP64)
    with pytest.raises(lib.RsfgError, match="rk5dp or rk3bs"):
        ens.solve(SEC_PER_YEAR, rtol=1e-6, method=lib.ARKIMEX_L2)
    # still usable after the refusals
    st = ens.solve(0.5 * SEC_PER_YEAR, rtol=1e-6)
    assert all(s["status"] == 0 for s in st)
    ens.close()
    only64.close()
    fs.close()
