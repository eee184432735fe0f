"""GPU parity of the physics branches beyond the default aging-law path,
through the C ABI against the oracle: the state laws 2-5 (src/rsf_engine.c:
44-132), the normal-traction "In" rows (src/rsf_solve.c:245-247) against the
reference-built golden tractions, the normal-stress evolution with its
limiter (src/rsf_engine.c:264-275), the monitor rows (src/rsf_output.c:
779-864), the event/monitor log capacities, and the continuation of a solve.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN
from helpers import SEC_PER_YEAR, geom_from_text, initial_state, row_rel

pytestmark = pytest.mark.gpu

G_BP5 = 3.204e10


def _L():
    from interact_amd import lib
    return lib


def _bp5(ds, law=1, prestress="make_bp5"):
    from interact_amd import faults
    from interact_amd.model import FaultSystem
    L = _L()
    g = faults.planar_grid(ds)
    f = faults.bp5_fields(g, prestress=prestress)
    fs = FaultSystem(g)
    fs.assemble(L.SLIP_STRIKE, shear_modulus=G_BP5)
    fs.set_friction(f["a"], f["b"], f["dc"], v0=1e-6, f0=0.6, eta=f["eta"], law=law)
    fs.set_loading(1e-9)
    X0 = initial_state(f["tau"], 25e6, f["vel"], f["a"], 1e-6)
    return g, f, fs, X0


def _oracle(g, f, law=1, **kw):
    K = oracle.traction_tensor(g, 0, 16)[:, :, 0, 0] * (G_BP5 / 1e4)
    return oracle.RSFProblem(K, f["a"], f["b"], f["dc"], K @ np.full(len(f["a"]), -1e-9),
                             eta=f["eta"], law=law, nthreads=16, **kw), K


# ------------------------------------------------------------------ operator
def test_normal_rows_match_reference_tractions():
    """The In rows equal -scale times the reference's normal tractions on
    the golden grid, scalar and two-component layouts (row-relative 1e-13)."""
    L = _L()
    from interact_amd.model import FaultSystem
    z = np.load(os.path.join(GOLDEN, "ktensor_grid.npz"))
    g = geom_from_text(str(z["geom"]))
    T = z["T"]
    n = T.shape[0]
    sc = G_BP5 / 1e4
    fs = FaultSystem(g)
    fs.assemble(L.SLIP_STRIKE, shear_modulus=G_BP5, with_normal=True)
    assert row_rel(fs.dense_normal(), -sc * T[:, :, 0, 2]) < 1e-13
    assert row_rel(fs.dense(), sc * T[:, :, 0, 0]) < 1e-13
    fs.assemble(L.SLIP_VECTOR, shear_modulus=G_BP5, with_normal=True)
    In = np.zeros((n, 2 * n))
    In[:, 0::2] = -sc * T[:, :, 0, 2]
    In[:, 1::2] = -sc * T[:, :, 1, 2]
    assert row_rel(fs.dense_normal(), In) < 1e-13
    fs.close()


# ------------------------------------------------------------------ state laws
@pytest.mark.parametrize("law", [2, 3, 4, 5])
def test_rhs_state_laws_match_oracle(law):
    """Device RHS under each state law vs the oracle (BP5 2 km fields, state
    off steady state, slip rates over seven decades)."""
    g, f, fs, X0 = _bp5(2.0, law=law)
    P, K = _oracle(g, f, law=law)
    rng = np.random.default_rng(law)
    X = X0.reshape(-1, 4).copy()
    X[:, 0] += rng.normal(0, 0.02, len(X))
    X = X.ravel()
    Fg, ok = fs.rhs(X)
    Fo = P.rhs(X)
    assert ok
    Fg, Fo = Fg.reshape(-1, 4), Fo.reshape(-1, 4)
    for k in range(4):
        assert np.abs(Fg[:, k] - Fo[:, k]).max() <= 1e-10 * np.abs(Fo[:, k]).max(), (law, k)
    fs.close()


@pytest.mark.parametrize("law", [2, 3, 4, 5])
def test_solve_state_laws_match_oracle(law):
    """rk5dp at rtol 1e-8 over the creep phase of the 2 km cycle fields
    (25 yr; 10 yr under Kato-Tullis, whose shifted steady state starts slip
    transients earlier): the device takes the oracle's steps and rejections
    and ends at its state (to the step-size rounding).  Through fast slip,
    where one accept/reject decision can flip with the product's summation
    order, the slider test below compares events instead."""
    from interact_amd import faults
    from interact_amd.model import FaultSystem
    L = _L()
    g = faults.planar_grid(2.0)
    f = faults.cycle_fields(g, shear_modulus=G_BP5)
    fs = FaultSystem(g)
    fs.assemble(L.SLIP_STRIKE, shear_modulus=G_BP5)
    fs.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"], law=law)
    fs.set_loading(f["vpl"])
    P, K = _oracle(g, f, law=law)
    X0 = initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"])
    t_end = (10.0 if law == 5 else 25.0) * SEC_PER_YEAR
    fs.set_state(X0)
    st = fs.solve(t_end, rtol=1e-8)
    ro = P.solve(X0, t_end, rtol=1e-8)
    assert st["status"] == 0 and ro["status"] == 0
    assert (st["steps"], st["rejects"]) == (ro["steps"], ro["rejects"])
    # the error norms differ in the last bits (the product's summation
    # order), and with them the step sizes, by ~1e-8 relative
    X, t = fs.get_state()
    assert abs(t - ro["t_final"]) <= 1e-6 * t
    Xr = X.reshape(-1, 4)
    Xo = ro["X"].reshape(-1, 4)
    for k in range(4):
        assert np.abs(Xr[:, k] - Xo[:, k]).max() <= 1e-7 * np.abs(Xo[:, k]).max(), k
    fs.close()


def _slider(law):
    from interact_amd.model import FaultSystem
    L = _L()
    g = oracle.read_patches(os.path.join(GOLDEN, "ref_results", "geom_slider10.in"))
    fs = FaultSystem(g)
    fs.assemble(L.SLIP_STRIKE, shear_modulus=G_BP5)
    eta = G_BP5 / (2 * 3464.0)
    a, b, dc = np.array([0.01]), np.array([0.015]), np.array([0.008])
    fs.set_friction(a, b, dc, eta=eta, law=law)
    fs.set_loading(1e-9)
    sig = 50e6
    X0 = initial_state(0.6 * sig + eta * 1e-9, sig, 1e-9, a, 1e-6)
    K = oracle.traction_tensor(g, 0, 1)[:, :, 0, 0] * (G_BP5 / 1e4)
    P = oracle.RSFProblem(K, a, b, dc, K @ np.full(1, -1e-9), eta=eta, law=law, nthreads=1)
    return fs, P, X0


@pytest.mark.parametrize("law", [2, 3, 4, 5])
def test_slider_cycles_under_each_law_match_oracle(law):
    """The 10 km slider with the lab parameters (slider/run_slider) over
    450 yr at rtol 1e-9 under each law: the device's events against the
    oracle's, onsets within 1e-3 yr and peak slip rates within 1e-3."""
    from interact_amd import accuracy
    fs, P, X0 = _slider(law)
    kw = dict(rtol=1e-9, rdx_monitor=1e-5, rupture_vth=0.1)
    fs.set_state(X0)
    st = fs.solve(450 * SEC_PER_YEAR, steps_per_launch=256, **kw)
    ro = P.solve(X0, 450 * SEC_PER_YEAR, **kw)
    assert st["status"] == 0 and ro["status"] == 0
    s = accuracy.score(ro["monitor"], ro["events"], fs.monitor(), fs.events())
    assert s["nev_run"] == s["nev_ref"] >= 7, s
    assert max(abs(d) for d in s["donset"]) < 1e-3, s
    assert max(abs(d) for d in s["dpeak"]) < 1e-3, s
    fs.close()


def test_catalog_rows_match_the_oracle(tmp_path):
    """rsfg_catalog on the slider (aging law, 450 yr): the device's
    catalog rows (rsf_finalize_event, src/rsf_output.c:1078-1180) against
    the oracle's, and the three text files written from the device logs
    read back with the rows they came from."""
    from interact_amd import output
    fs, P, X0 = _slider(1)
    area = [1e8]   # 4 l w of the 10 km square slider
    kw = dict(rtol=1e-9, rdx_monitor=1e-5, rupture_vth=0.1)
    fs.set_state(X0)
    st = fs.solve(450 * SEC_PER_YEAR, steps_per_launch=256, **kw)
    ro = P.solve(X0, 450 * SEC_PER_YEAR, cat_area=area, cat_mu=G_BP5, **kw)
    cd, co = fs.catalog(), ro["catalog"]
    assert st["ncatalog"] == len(cd) == len(co) == 7 and st["catalog_dropped"] == 0
    assert np.abs(cd[:, :2] - co[:, :2]).max() / SEC_PER_YEAR < 1e-3
    assert np.array_equal(cd[:, 2:4], co[:, 2:4])
    for k in (4, 5, 6, 7, 8, 9):
        assert np.abs(cd[:, k] / co[:, k] - 1.0).max() < 1e-3, k
    assert np.abs(cd[:, 10] - co[:, 10]).max() < 1e-3
    output.write_outputs(fs, str(tmp_path), vel_event=1e-3, rupture_vth=0.1)
    mon = np.loadtxt(str(tmp_path / output.MONITOR_FILE))
    evf = np.loadtxt(str(tmp_path / output.EVENTS_FILE))
    catf = np.loadtxt(str(tmp_path / output.CATALOG_FILE))
    m, e = fs.monitor(), fs.events()
    assert mon.shape == (len(m), 10) and np.allclose(mon[:, 1], m[:, 1], rtol=1e-8)
    assert evf.shape == (len(e), 6) and np.array_equal(evf[:, 2], e[:, 1])
    assert catf.shape == (7, 13) and np.allclose(catf[:, 11], cd[:, 9], rtol=1e-6)
    fs.close()


# ------------------------------------------------------------------ normal stress
def _dipping_grid(ds_km, nx=40, nz=12, dip=30.0, top_km=1.0):
    """planar fault of nx x nz square patches, strike 0 (+y), dipping dip
    degrees, top edge top_km deep: strike slip on it changes the normal
    stress on the fault (zero on a vertical strike-slip fault)."""
    d = np.radians(dip)
    i, j = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    i, j = i.ravel(), j.ravel()
    ds = ds_km * 1e3
    s = (j + 0.5) * ds                           # down-dip distance
    xyz = np.stack([s * np.cos(d), (i + 0.5 - nx / 2) * ds, -top_km * 1e3 - s * np.sin(d)], 1)
    n = nx * nz
    return dict(xyz=xyz, strike=np.zeros(n, np.float32), dip=np.full(n, dip, np.float32),
                l=np.full(n, ds / 2), w=np.full(n, ds / 2),
                along=(i + 0.5 - nx / 2) * ds_km, depth=top_km + s / 1e3 * np.sin(d))


def _sigma_setup(limit):
    from interact_amd import faults
    from interact_amd.model import FaultSystem
    L = _L()
    g = _dipping_grid(2.0)
    f = faults.bp5_fields(g)
    n = len(g["l"])
    sc = G_BP5 / 1e4
    T = oracle.traction_tensor(g, 0, 16)
    K, Kn = sc * T[:, :, 0, 0], -sc * T[:, :, 0, 2]
    assert np.abs(Kn).max() > 0.05 * np.abs(K).max()
    fs = FaultSystem(g)
    fs.assemble(L.SLIP_STRIKE, shear_modulus=G_BP5, with_normal=True)
    assert row_rel(fs.dense_normal(), Kn) < 1e-12
    kw = dict(calc_sigma_dot=True, limit_sigma=limit, min_sigma=24.9e6, max_sigma=25.1e6)
    fs.set_friction(f["a"], f["b"], f["dc"], v0=1e-6, f0=0.6, eta=f["eta"], **kw)
    fs.set_loading(1e-9)
    vpl = np.full(n, -1e-9)
    P = oracle.RSFProblem(K, f["a"], f["b"], f["dc"], K @ vpl, Kn=Kn, sinc_n=Kn @ vpl,
                          eta=f["eta"], nthreads=16, **kw)
    X0 = initial_state(f["tau"], 25e6, f["vel"], f["a"], 1e-6)
    return g, f, fs, P, X0


@pytest.mark.parametrize("limit", [False, True])
def test_sigma_evolution_matches_oracle(limit):
    """calc_sigma_dot (sigma_dot = K_n v + sinc_n, with and without the
    limiter) on a 30-degree dipping fault: the device RHS against the oracle,
    then rk5dp over the seeded nucleation with the oracle's steps, sigma
    changing and, with the limiter, held where it left [24.9, 25.1] MPa."""
    g, f, fs, P, X0 = _sigma_setup(limit)
    rng = np.random.default_rng(9)
    X = X0.reshape(-1, 4).copy()
    X[:, 0] += rng.normal(0, 0.01, len(X))
    X[:, 2] *= 1 + rng.normal(0, 0.01, len(X))
    X = X.ravel()
    Fg, ok = fs.rhs(X)
    Fo = P.rhs(X)
    assert ok
    Fg, Fo = Fg.reshape(-1, 4), Fo.reshape(-1, 4)
    for k in range(4):
        assert np.abs(Fg[:, k] - Fo[:, k]).max() <= 1e-10 * np.abs(Fo[:, k]).max(), k
    assert np.abs(Fo[:, 2]).max() > 0
    if limit:
        sig = X.reshape(-1, 4)[:, 2]
        out = (sig < 24.9e6) | (sig > 25.1e6)
        assert out.any() and (Fg[out, 2] == 0).all()
    t_end = 0.02 * SEC_PER_YEAR
    fs.set_state(X0)
    st = fs.solve(t_end, rtol=1e-6)
    ro = P.solve(X0, t_end, rtol=1e-6)
    assert (st["steps"], st["rejects"]) == (ro["steps"], ro["rejects"])
    Xg, t = fs.get_state()
    assert np.abs(Xg - ro["X"]).max() / np.abs(ro["X"]).max() < 1e-8
    dsig = np.abs(Xg.reshape(-1, 4)[:, 2] - 25e6).max()
    assert dsig > 1e3
    fs.close()


def test_per_patch_plate_rate_and_additive_rates():
    """-rsf_vpl_file and -rsf_stress_rate_file (src/rsf_solve.c:268-363):
    a per-patch plate rate in the backslip products and per-patch additive
    shear and normal rates, on the dipping fault with the normal-stress
    evolution on.  The device RHS against the oracle's with the same
    loading (FP64 exact to 1e-10; the FP32 operator, which folds the plate
    rate into K (v - vpl), to 1e-5), then rk5dp through the seeded
    nucleation with the oracle's steps."""
    from interact_amd import faults
    from interact_amd.model import FaultSystem
    L = _L()
    g = _dipping_grid(2.0)
    f = faults.bp5_fields(g)
    n = len(g["l"])
    k = np.arange(n)
    vpl = 1e-9 * (1.0 + 0.5 * np.sin(0.37 * k))
    tdot = 2e-3 * np.cos(0.11 * k)
    sdot = 1e-3 * np.sin(0.23 * k)
    sc = G_BP5 / 1e4
    T = oracle.traction_tensor(g, 0, 16)
    K, Kn = sc * T[:, :, 0, 0], -sc * T[:, :, 0, 2]
    fs = FaultSystem(g)
    fs.assemble(L.SLIP_STRIKE, shear_modulus=G_BP5, with_normal=True,
                store=L.PREC_FP32)
    kw = dict(calc_sigma_dot=True)
    fs.set_friction(f["a"], f["b"], f["dc"], v0=1e-6, f0=0.6, eta=f["eta"], **kw)
    fs.set_loading(1e-9, vpl_patch=vpl, taudot_add=tdot, sigdot_add=sdot)
    P = oracle.RSFProblem(K, f["a"], f["b"], f["dc"], K @ (-vpl) + tdot, Kn=Kn,
                          sinc_n=Kn @ (-vpl) + sdot, eta=f["eta"], nthreads=16, **kw)
    X0 = initial_state(f["tau"], 25e6, f["vel"], f["a"], 1e-6)
    rng = np.random.default_rng(12)
    X = X0.reshape(-1, 4).copy()
    X[:, 0] += rng.normal(0, 0.01, len(X))
    X = X.ravel()
    Fo = P.rhs(X).reshape(-1, 4)
    for prec, tol in ((L.PREC_FP64, 1e-10), (L.PREC_FP32, 1e-5)):
        Fg, ok = fs.rhs(X, prec)
        assert ok
        Fg = Fg.reshape(-1, 4)
        for c in range(4):
            assert np.abs(Fg[:, c] - Fo[:, c]).max() <= tol * np.abs(Fo[:, c]).max(), (prec, c)
    # the per-patch terms move the rates: without them the RHS differs
    P0 = oracle.RSFProblem(K, f["a"], f["b"], f["dc"], K @ np.full(n, -1e-9), Kn=Kn,
                           sinc_n=Kn @ np.full(n, -1e-9), eta=f["eta"], nthreads=16, **kw)
    F0 = P0.rhs(X).reshape(-1, 4)
    assert np.abs(F0[:, 2] - Fo[:, 2]).max() > 1e-4
    t_end = 0.02 * SEC_PER_YEAR
    fs.set_state(X0)
    st = fs.solve(t_end, rtol=1e-6)
    ro = P.solve(X0, t_end, rtol=1e-6)
    assert (st["steps"], st["rejects"]) == (ro["steps"], ro["rejects"])
    Xg, t = fs.get_state()
    assert np.abs(Xg - ro["X"]).max() / np.abs(ro["X"]).max() < 1e-8
    # the FP32 solve folds each patch's plate rate into K (v - vpl_p); a
    # uniform rate in its place would move tau by ~3e-4 here
    fs.set_state(X0)
    st = fs.solve(t_end, rtol=1e-6, prec=L.PREC_FP32)
    X32, _ = fs.get_state()
    d = np.abs(X32 - ro["X"]).reshape(-1, 4).max(0) / np.abs(ro["X"]).reshape(-1, 4).max(0)
    print("FP32 fold vs oracle, per component:", d)
    assert st["status"] == 0 and d[1] < 2e-5 and d[2] < 2e-5, d
    fs.close()


# ------------------------------------------------------------------ monitor
def test_monitor_rows_match_oracle():
    """rsf_monitor rows (step, t, t [yr], dt, log10 max|v|, mean slip, mean
    mu, sigma max, sigma min, log10 mean|v|) of the device controller against
    the oracle's, BP5 2 km through the seeded event at rtol 1e-6."""
    g, f, fs, X0 = _bp5(2.0)
    P, K = _oracle(g, f)
    t_end = 0.05 * SEC_PER_YEAR
    fs.set_state(X0)
    st = fs.solve(t_end, rtol=1e-6)
    ro = P.solve(X0, t_end, rtol=1e-6)
    mg, mo = fs.monitor(), ro["monitor"]
    assert st["nmonitor"] == len(mg) == len(mo) > 10
    assert np.array_equal(mg[:, 0], mo[:, 0])           # same step numbers
    for c in range(1, 10):
        scale = np.abs(mo[:, c]).max()
        assert np.abs(mg[:, c] - mo[:, c]).max() <= 1e-9 * scale + 1e-300, c
    fs.close()


def test_event_and_monitor_logs_report_overflow():
    """Logs capped through the solver options keep their first rows and
    count the rest as dropped (no silent truncation); the slider's seven
    cycles give 14 event rows."""
    fs, P, X0 = _slider(1)
    kw = dict(rtol=1e-9, rdx_monitor=1e-5, rupture_vth=0.1, steps_per_launch=256)
    t_end = 450 * SEC_PER_YEAR
    fs.set_state(X0)
    full = fs.solve(t_end, **kw)
    ev, mon = fs.events(), fs.monitor()
    assert full["events_dropped"] == 0 and full["monitor_dropped"] == 0
    assert len(ev) == 14 and len(mon) > 100
    fs.set_state(X0)
    st = fs.solve(t_end, max_events=3, max_monitor=5, **kw)
    assert st["nevent"] == 3 and st["events_dropped"] == len(ev) - 3
    assert st["nmonitor"] == 5 and st["monitor_dropped"] == len(mon) - 5
    assert np.array_equal(fs.events(), ev[:3])
    assert np.array_equal(fs.monitor(), mon[:5])
    # a later solve with the default capacity stores every row again
    fs.set_state(X0)
    again = fs.solve(t_end, **kw)
    assert again["events_dropped"] == 0 and np.array_equal(fs.events(), ev)
    fs.close()


# ------------------------------------------------------------------ continuation
def test_solve_continue_equals_one_solve():
    """A solve chopped into three continued segments takes exactly the steps
    of one solve to the final time: same state bits, events and monitor rows
    (BP5 2 km through the seeded event, rk5dp rtol 1e-6)."""
    L = _L()
    g, f, fs, X0 = _bp5(2.0)
    T = 0.05 * SEC_PER_YEAR
    fs.set_state(X0)
    one = fs.solve(T, rtol=1e-6)
    X1, t1 = fs.get_state()
    ev1, mon1 = fs.events(), fs.monitor()
    fs.set_state(X0)
    fs.solve(T / 3, rtol=1e-6)
    fs.solve_continue(2 * T / 3, rtol=1e-6)
    seg = fs.solve_continue(T, rtol=1e-6)
    X2, t2 = fs.get_state()
    assert (seg["steps"], seg["rejects"], seg["rhs_evals"]) == \
        (one["steps"], one["rejects"], one["rhs_evals"])
    assert t1 == t2 and np.array_equal(X1, X2)
    assert np.array_equal(fs.events(), ev1) and np.array_equal(fs.monitor(), mon1)
    # continuing to a time already reached does nothing
    same = fs.solve_continue(T / 2, rtol=1e-6)
    assert same["steps"] == one["steps"] and np.array_equal(fs.get_state()[0], X1)
    # a changed state, or another operator precision, cannot be continued
    with pytest.raises(L.RsfgError, match="precision"):
        fs.solve_continue(2 * T, rtol=1e-6, prec=L.PREC_FP32)
    fs.set_state(X0)
    with pytest.raises(L.RsfgError, match="no solve to continue"):
        fs.solve_continue(2 * T, rtol=1e-6)
    fs.close()
