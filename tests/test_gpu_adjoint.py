"""GPU tests of the sensitivities (rsfg_set_step_log, rsfg_replay,
rsfg_adjoint_gradient) against the oracle's FP64 reference adjoint
(oracle/rsfo_adj.c, pinned on the CPU against central differences) and
against central differences of the device replay itself."""
import numpy as np
import pytest

from helpers import SEC_PER_YEAR, AdjointCase

pytestmark = pytest.mark.gpu


def _device(case, prec=None):
    from interact_amd import lib
    from interact_amd.model import FaultSystem
    fs = FaultSystem(case.g)
    mode = lib.SLIP_VECTOR if case.ncomp == 2 else lib.SLIP_STRIKE
    fs.assemble(mode, shear_modulus=32e9, store=lib.PREC_FP32 if prec is None else prec)
    f = case.f
    fs.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"])
    fs.set_loading(f["vpl"])
    fs.set_state(case.X_start, case.t_start)
    return fs


@pytest.fixture(scope="module")
def case2():
    return AdjointCase(ncomp=2, nthreads=16)


@pytest.fixture(scope="module")
def case1():
    return AdjointCase(ncomp=1, window_yr=20.0, nthreads=16)


def test_step_log_replay_reproduces_the_device_solve(case2):
    """The accepted steps of a device solve, replayed from the same state,
    land on the solve's final state (the replay's kernels round differently
    from the fused stage kernel: agreement 1e-10 of each field's scale), and
    the oracle replays the same steps onto the same state."""
    from interact_amd import lib
    fs = _device(case2)
    fs.set_step_log(100000)
    st = fs.solve(case2.t_start + 40.0 * SEC_PER_YEAR, rtol=1e-6, prec=lib.PREC_FP64)
    steps = fs.step_log()
    assert len(steps) == st["steps"] > 50
    X1, t1 = fs.get_state()
    assert abs(case2.t_start + steps.sum() - t1) <= 1e-9 * t1
    fs.set_state(case2.X_start, case2.t_start)
    r = fs.replay(steps)
    P = case2.problem(case2.f["a"], case2.f["b"], case2.f["dc"], 16)
    _, Xo = P.replay(case2.X_start, steps, [], 0)
    d = 6
    for k in range(d):
        sc = np.abs(X1[k::d]).max()
        assert np.abs(r["X_end"][k::d] - X1[k::d]).max() <= 1e-10 * sc
        assert np.abs(r["X_end"][k::d] - Xo[k::d]).max() <= 1e-10 * sc
    # the handle's state is unchanged by the replay
    X2, t2 = fs.get_state()
    assert t2 == case2.t_start and np.array_equal(X2, case2.X_start)
    fs.close()


@pytest.mark.parametrize("which,method", [(2, 0), (1, 1)])
def test_adjoint_matches_the_oracle(case2, case1, which, method):
    """J and dJ/d(a, b, D_c) from the device adjoint (FP64 operator) against
    the oracle's FP64 reference adjoint on the same steps: J to 1e-11 and
    each gradient to 1e-8 of its norm; two slip components with rk5dp and
    one with rk3bs."""
    from interact_amd import lib
    c = case2 if which == 2 else case1
    fs = _device(c)
    P = c.problem(c.f["a"], c.f["b"], c.f["dc"], 16)
    Jo, gao, gbo, gdo = P.adjoint(c.X_start, c.steps, c.recv, c.target, c.weight, method)
    r = fs.adjoint_gradient(c.steps, c.recv, c.target, c.weight, method=method,
                            prec=lib.PREC_FP64, checkpoint_every=32)
    assert abs(r["J"] - Jo) <= 1e-11 * Jo
    for g, go in ((r["grad_a"], gao), (r["grad_b"], gbo), (r["grad_dc"], gdo)):
        assert np.linalg.norm(g - go) <= 1e-8 * np.linalg.norm(go)
    # the replay objective is the same number
    rr = fs.replay(c.steps, c.recv, c.target, c.weight, method=method)
    assert rr["J"] == r["J"]
    fs.close()


def test_adjoint_matches_central_differences_of_the_device_replay(case2):
    """8 random directions in (a - b at fixed a, D_c), scaled by the
    parameters: <grad, d> against central differences of the device replay
    objective at eps = 1e-5 within 1e-6."""
    from interact_amd import lib
    from interact_amd.autograd import MisfitProblem
    c = case2
    fs = _device(c)
    f = c.f
    pb = MisfitProblem(fs, c.X_start, c.t_start, f["a"], c.steps, c.recv, c.target, c.weight,
                       v0=f["v0"], f0=f["f0"], eta=f["eta"], prec=lib.PREC_FP64)
    amb0 = f["a"] - f["b"]
    J, g_amb, g_dc = pb.value_and_grad(amb0, f["dc"])
    rng = np.random.default_rng(21)
    eps = 1e-5
    for k in range(8):
        d1 = rng.standard_normal(c.n) * np.abs(amb0)
        d2 = rng.standard_normal(c.n) * f["dc"]
        Jp = pb.value(amb0 + eps * d1, f["dc"] + eps * d2)
        Jm = pb.value(amb0 - eps * d1, f["dc"] - eps * d2)
        fdv = (Jp - Jm) / (2 * eps)
        ad = g_amb @ d1 + g_dc @ d2
        assert abs(fdv - ad) <= 1e-6 * abs(ad), (k, fdv, ad)
    fs.close()


def test_checkpoint_interval_does_not_change_the_gradient(case2):
    """Recomputation from checkpoints replays bitwise the same stages: the
    gradient is identical for checkpoints every 1, 7 and 256 steps."""
    from interact_amd import lib
    c = case2
    fs = _device(c)
    outs = [fs.adjoint_gradient(c.steps, c.recv, c.target, c.weight, prec=lib.PREC_FP64,
                                checkpoint_every=k) for k in (1, 7, 256)]
    for o in outs[1:]:
        assert o["J"] == outs[0]["J"]
        for key in ("grad_a", "grad_b", "grad_dc"):
            assert np.array_equal(o[key], outs[0][key])
    fs.close()


def test_fp32_operator_adjoint_is_close(case2):
    """The FP32 operator (K v and K^T w from 4-byte entries, FP64 state)
    changes the gradient by less than 1e-4 of its norm."""
    from interact_amd import lib
    c = case2
    fs = _device(c)
    r64 = fs.adjoint_gradient(c.steps, c.recv, c.target, c.weight, prec=lib.PREC_FP64)
    r32 = fs.adjoint_gradient(c.steps, c.recv, c.target, c.weight, prec=lib.PREC_FP32)
    assert abs(r32["J"] - r64["J"]) <= 1e-5 * r64["J"]
    for key in ("grad_a", "grad_b", "grad_dc"):
        assert np.linalg.norm(r32[key] - r64[key]) <= 1e-4 * np.linalg.norm(r64[key])
    fs.close()


def test_torch_autograd_wrapper(case2):
    """cycle_misfit's backward returns the adjoint gradient (dJ/d(a - b) at
    fixed a = -dJ/db) and its forward the replay objective."""
    import torch
    from interact_amd import lib
    from interact_amd.autograd import MisfitProblem, cycle_misfit
    c = case2
    fs = _device(c)
    f = c.f
    pb = MisfitProblem(fs, c.X_start, c.t_start, f["a"], c.steps, c.recv, c.target, c.weight,
                       v0=f["v0"], f0=f["f0"], eta=f["eta"], prec=lib.PREC_FP64)
    amb = torch.tensor(f["a"] - f["b"], dtype=torch.float64, requires_grad=True)
    dc = torch.tensor(f["dc"], dtype=torch.float64, requires_grad=True)
    J = cycle_misfit(amb, dc, pb)
    J.backward()
    J1, g_amb, g_dc = pb.value_and_grad(amb.detach().numpy(), dc.detach().numpy())
    assert float(J.detach()) == J1
    assert np.array_equal(amb.grad.numpy(), g_amb)
    assert np.array_equal(dc.grad.numpy(), g_dc)
    # dJ/d(a - b) at fixed a is -dJ/db
    r = fs.adjoint_gradient(c.steps, c.recv, c.target, c.weight, prec=lib.PREC_FP64)
    assert np.array_equal(g_amb, -r["grad_b"])
    with torch.no_grad():
        J0 = cycle_misfit(amb, dc, pb)
    assert float(J0) == J1
    fs.close()


def test_adjoint_argument_checks(case2):
    from interact_amd import lib
    c = case2
    fs = _device(c)
    with pytest.raises(lib.RsfgError, match="rk5dp or rk3bs"):
        fs.adjoint_gradient(c.steps, c.recv, c.target, c.weight, method=lib.ARKIMEX_L2)
    with pytest.raises(lib.RsfgError, match="positive and finite"):
        bad = c.steps.copy()
        bad[3] = -1.0
        fs.replay(bad)
    with pytest.raises(lib.RsfgError, match="out of range"):
        fs.replay(c.steps, [0, c.n])
    with pytest.raises(lib.RsfgError, match="FP64 or FP32"):
        fs.replay(c.steps, method=lib.RK5DP, prec=lib.PREC_BF16)
    fs.close()
    from interact_amd.model import FaultSystem
    fn = FaultSystem(c.g)
    fn.assemble(lib.SLIP_VECTOR, shear_modulus=32e9, with_normal=True)
    f = c.f
    fn.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"])
    fn.set_loading(f["vpl"])
    fn.set_state(c.X_start, c.t_start)
    with pytest.raises(lib.RsfgError, match="normal rows"):
        fn.replay(c.steps)
    fn.close()
