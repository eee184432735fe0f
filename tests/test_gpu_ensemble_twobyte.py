"""GPU tests of the batched ensemble's two-byte operator
(rsfg_ensemble_set_operator with RSFG_PREC_BF16NF): bf16 hi tiles of every
entry on v_mfma_f32_16x16x32_bf16, bf16 lo tiles in each row's near-field
256-column segments, member slip rates as three bf16 terms whose sum is
their FP32 value.  Checked against the same ensemble on its FP32 operator,
whose members the other ensemble tests check against the FP64 oracle and
the single-fault solver.  No counterpart in the reference."""
import numpy as np
import pytest

from helpers import SEC_PER_YEAR, initial_state

pytestmark = pytest.mark.gpu

MU = 32e9


def _ensemble(outlines, nm, ds=1.0, ncomp=2):
    from interact_amd import faults, lib
    from interact_amd.model import Ensemble, FaultSystem
    geoms, X0s = [], []
    for (L, W) in outlines:
        g = faults.planar_grid(ds, L, W)
        fs = FaultSystem(g)
        fs.assemble(lib.SLIP_VECTOR if ncomp == 2 else lib.SLIP_STRIKE, shear_modulus=MU,
                    store=lib.PREC_FP32)
        geoms.append((fs, g, L, W))
    ens = Ensemble([q[0] for q in geoms], nm)
    for gi, (fs, g, L, W) in enumerate(geoms):
        for k in range(nm):
            f = faults.ensemble_fields(g, L, W, k, gi)
            m = ens.members[gi * nm + k]
            m.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"])
            m.set_loading(f["vpl"])
            X0 = initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=ncomp)
            m.set_state(X0)
            X0s.append(X0)
    return ens, geoms, X0s


def _close(ens, geoms):
    ens.close()
    for q in geoms:
        q[0].close()


def _onsets(ev):
    return ev[ev[:, 1] > 0, 0] / SEC_PER_YEAR if len(ev) else np.zeros(0)


def _first_step(ens, X0s, prec):
    """each member's state increment over one accepted step of the same size"""
    from interact_amd import lib
    for m, X0 in zip(ens.members, X0s):
        m.set_state(X0)
    ens.solve(1e3 * SEC_PER_YEAR, rtol=1e-6, prec=prec, max_steps=1, dt_max=1e6,
              dt_init=1e6)
    return [m.get_state()[0] - X0 for m, X0 in zip(ens.members, X0s)]


@pytest.mark.parametrize("nm,ncomp", [(16, 2), (32, 2), (32, 1)])
def test_two_byte_product_against_the_fp32_operator(nm, ncomp):
    """One step of 1e6 s from the initial states on a 1 km (1600 DOF, seven
    256-column segments; 800 in the one-component layout) and a 1 km
    60 x 24 km geometry, both slip layouts: each member's state increment
    on the two-byte operator (every stream-K shape) agrees with the FP32
    operator's to 2e-5 relative; with tolerance 0.5 (lo only over each row's
    largest segments) the disagreement is larger (the lo tiles carry the
    near field), and the operator streams more than half and less than the
    FP32 bytes (on these small geometries the lo hulls span most segments)."""
    from interact_amd import lib
    ens, geoms, X0s = _ensemble([(40, 20), (60, 24)], nm, ds=1.0, ncomp=ncomp)
    ref = _first_step(ens, X0s, lib.PREC_FP32)
    b32 = ens.set_operator(lib.PREC_FP32)
    scale = [np.linalg.norm(d) for d in ref]
    assert min(scale) > 0

    def worst(out):
        return max(np.linalg.norm(d - r) / s for d, r, s in zip(out, ref, scale))

    bh = ens.set_operator(lib.PREC_BF16NF)
    assert 0.5 * b32 < bh < b32, (bh, b32)
    shapes = ens.shapes()
    tried = []
    for cfg, sh in enumerate(shapes):
        try:
            ens.select(cfg)
        except lib.RsfgError as err:
            assert "does not fit" in str(err) and sh[6] == 0, (cfg, sh)
            continue
        tried.append((cfg, worst(_first_step(ens, X0s, lib.PREC_BF16NF))))
    assert len(tried) == 8, tried
    assert max(e for _, e in tried) < 2e-5, tried
    b_coarse = ens.set_operator(lib.PREC_BF16NF, nf_tol=0.5)
    assert b_coarse < bh
    e_coarse = worst(_first_step(ens, X0s, lib.PREC_BF16NF))
    assert e_coarse > 3 * max(e for _, e in tried), (e_coarse, tried)
    _close(ens, geoms)


def test_two_byte_ensemble_run_matches_the_fp32_ensemble():
    """Two geometries at 2 km x 32 members over 20 yr: every member on the
    two-byte operator (tuned shape) has the FP32-operator run's events,
    onsets within 5e-3 yr, and its step count within 5 %; switching back to
    the FP32 operator reproduces the FP32 run bit for bit."""
    from interact_amd import lib
    ens, geoms, X0s = _ensemble([(40, 20), (60, 24)], 32, ds=2.0)
    t_end = 20.0 * SEC_PER_YEAR

    def run(prec):
        for m, X0 in zip(ens.members, X0s):
            m.set_state(X0)
        st = ens.solve(t_end, rtol=1e-6, prec=prec)
        assert all(q["status"] == 0 and q["t"] >= t_end for q in st)
        return st, [(m.get_state()[0], m.events()) for m in ens.members]

    st32, out32 = run(lib.PREC_FP32)
    ens.set_operator(lib.PREC_BF16NF)
    choice, ms, shapes = ens.tune(rtol=1e-6)
    assert shapes[choice][6] > 0 and all(ms[c] < 0 for c, s in enumerate(shapes) if s[6] == 0)
    sth, outh = run(lib.PREC_BF16NF)
    n_on = 0
    for i, ((X, ev), (Xr, evr)) in enumerate(zip(outh, out32)):
        on, onr = _onsets(ev), _onsets(evr)
        assert len(on) == len(onr), (i, on, onr)
        if len(on):
            assert np.abs(on - onr).max() < 5e-3, (i, on, onr)
        n_on += len(on)
        assert abs(sth[i]["steps"] - st32[i]["steps"]) <= 0.05 * st32[i]["steps"] + 5
    assert n_on >= 16
    ens.set_operator(lib.PREC_FP32)
    ens.select(0)   # the shape of the first FP32 run
    st2, out2 = run(lib.PREC_FP32)
    _close(ens, geoms)
    for (X, ev), (Xr, evr) in zip(out2, out32):
        assert np.array_equal(X, Xr) and np.array_equal(ev, evr)


def test_two_byte_operator_argument_checks():
    from interact_amd import lib
    ens, geoms, X0s = _ensemble([(40, 20)], 64, ds=2.0)
    with pytest.raises(lib.RsfgError, match="16 or 32 members"):
        ens.set_operator(lib.PREC_BF16NF)
    with pytest.raises(lib.RsfgError, match="RSFG_PREC_FP32 or RSFG_PREC_BF16NF"):
        ens.set_operator(lib.PREC_BF16)
    _close(ens, geoms)
    ens, geoms, X0s = _ensemble([(40, 20)], 16, ds=2.0)
    with pytest.raises(lib.RsfgError, match="below 1"):
        ens.set_operator(lib.PREC_BF16NF, nf_tol=1.5)
    ens.set_operator(lib.PREC_BF16NF)
    with pytest.raises(lib.RsfgError, match="prec must name"):
        ens.solve(SEC_PER_YEAR, rtol=1e-6, prec=lib.PREC_FP32)
    with pytest.raises(lib.RsfgError, match="does not fit"):
        ens.select(0)
    st = ens.solve(0.5 * SEC_PER_YEAR, rtol=1e-6)
    assert all(s["status"] == 0 for s in st)
    _close(ens, geoms)
