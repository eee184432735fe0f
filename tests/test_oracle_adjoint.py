"""The oracle's discrete adjoint (oracle/rsfo_adj.c), pinned on the CPU:
its complex restatement of the patch law against the pinned right-hand
side (rsfo_rhs, src/rsf_engine.c:186-302), its complex-step Jacobians
against central differences of that right-hand side, its replay against
the adaptive solver it reuses, and its gradient against central differences
of its own objective.  (The reference has no adjoint: what fixes this one
is the mathematics of the replayed map.)"""
import numpy as np
import pytest

import oracle
from helpers import AdjointCase, initial_state


def _patch_problem(ncomp, law, rng):
    n = 6
    D = ncomp * n
    K = rng.normal(0, 1e3, (D, D))
    a = rng.uniform(0.008, 0.012, n)
    b = rng.uniform(0.012, 0.016, n)
    dc = rng.uniform(0.01, 0.05, n)
    P = oracle.RSFProblem(K, a, b, dc, rng.normal(0, 1e-2, D), ncomp=ncomp, eta=4.6e6,
                          law=law, nthreads=1)
    vel = rng.uniform(1e-10, 1e-6, n)
    tau = rng.uniform(0.5, 0.7, n) * 25e6
    X = initial_state(tau, 25e6, vel, a, 1e-6, ncomp=ncomp)
    X = X.reshape(n, -1)
    X[:, 0] += rng.normal(0, 0.02, n)          # off steady state
    if ncomp == 2:
        X[:, 2] = 0.3 * X[:, 1]                # a dip component of the shear stress
    return P, X


@pytest.mark.parametrize("ncomp", [1, 2])
@pytest.mark.parametrize("law", [1, 2, 3, 4, 5])
def test_complex_restatement_and_jacobians(ncomp, law):
    """f and v from the complex restatement equal rsfo_rhs patch by patch
    (the stressing rates are those rsfo_rhs forms); every complex-step
    Jacobian column equals a central difference of rsfo_rhs (or of the slip
    rate, or ln|v|) within 1e-5 of its scale at the best of four steps
    (differences of the law reach about 1e-6 at best where its terms
    cancel)."""
    rng = np.random.default_rng(10 * law + ncomp)
    P, X = _patch_problem(ncomp, law, rng)
    n, dim = X.shape
    F = P.rhs(X.ravel()).reshape(n, dim)
    Kd = P._keep[0].reshape(ncomp * n, ncomp * n)
    v = np.zeros(ncomp * n)
    for i in range(n):
        _, v[ncomp * i:ncomp * i + ncomp] = P.patch_rhs_c(i, X[i], np.zeros(ncomp))
    td = Kd @ v
    coef = {"a": P._keep[1], "b": P._keep[2], "dc": P._keep[3]}
    for i in range(n):
        t = td[ncomp * i:ncomp * i + ncomp]
        f, vi = P.patch_rhs_c(i, X[i], t)
        np.testing.assert_allclose(f, F[i], rtol=1e-13, atol=1e-13 * np.abs(F[i]).max())
        jac = P.patch_jacobian(i, X[i], t)

        def fd_err(fun, x0, k, scale, ref):
            """smallest relative distance of ref to central differences of
            fun at relative steps 1e-4 .. 1e-7 (truncation error falls and
            cancellation error grows with the step; a wrong term or sign is
            off by O(1) at every step)"""
            best = np.inf
            for rel in (1e-4, 1e-5, 1e-6, 1e-7):
                h = rel * scale
                xp, xm = x0.copy(), x0.copy()
                xp[k] += h
                xm[k] -= h
                col = np.atleast_1d((fun(xp) - fun(xm)) / (2 * h))
                r = np.atleast_1d(ref)
                best = min(best, np.abs(col - r).max() / max(np.abs(col).max(), 1e-300))
            return best

        def lnv(vv):
            return np.log(np.sqrt(np.sum(vv ** 2)))

        for q in range(dim):
            if q >= ncomp + 2:
                continue   # slip does not enter the law
            sc = max(abs(X[i, q]), 1e-3)
            assert fd_err(lambda x: P.patch_rhs_c(i, x, t)[0], X[i], q, sc, jac["Jx"][:, q]) < 1e-5
            assert fd_err(lambda x: P.patch_rhs_c(i, x, t)[1], X[i], q, sc, jac["Vx"][:, q]) < 1e-5
            assert fd_err(lambda x: lnv(P.patch_rhs_c(i, x, t)[1]), X[i], q, sc, jac["Lx"][q]) < 1e-5
        for c in range(ncomp):
            sc = max(abs(t[c]), 1.0)
            assert fd_err(lambda tt: P.patch_rhs_c(i, X[i], tt)[0], t, c, sc, jac["Jtd"][:, c]) < 1e-7
        # parameters: perturb the problem's own coefficient arrays
        for k, name in enumerate(("a", "b", "dc")):
            arr = coef[name]
            x0 = arr[i]

            def at(pk, out):
                arr[i] = pk[0]
                r = P.patch_rhs_c(i, X[i], t)
                arr[i] = x0
                return r[0] if out == 0 else (r[1] if out == 1 else lnv(r[1]))
            p0 = np.array([x0])
            assert fd_err(lambda pk: at(pk, 0), p0, 0, x0, jac["Jp"][:, k]) < 1e-5
            if name == "a":
                assert fd_err(lambda pk: at(pk, 1), p0, 0, x0, jac["Vp"]) < 1e-5
                assert fd_err(lambda pk: at(pk, 2), p0, 0, x0, jac["La"]) < 1e-5


@pytest.fixture(scope="module")
def case():
    return AdjointCase(ncomp=2)


def test_oracle_replay_reproduces_the_adaptive_solve(case):
    """The replay of the logged accepted steps is the solve: bitwise the same
    final state (both step with rsfo_rk_attempt's stage arithmetic)."""
    P = case.problem(case.f["a"], case.f["b"], case.f["dc"])
    _, Xe = P.replay(case.X_start, case.steps, [], 0)
    assert len(case.steps) > 50
    assert np.array_equal(Xe, case.X_solve)


def test_oracle_adjoint_matches_central_differences(case):
    """dJ/d(a, b, D_c) of the oracle adjoint against central differences of
    the replayed objective on 6 random directions (separately in a, b and
    D_c, each scaled by the parameter): agreement to 1e-7 at eps = 1e-5
    (the difference error is O(eps^2))."""
    f = case.f
    P = case.problem(f["a"], f["b"], f["dc"])
    J, ga, gb, gd = P.adjoint(case.X_start, case.steps, case.recv, case.target, case.weight, 0)
    assert abs(J - case.J_oracle(f["a"], f["b"], f["dc"])) <= 1e-12 * J
    rng = np.random.default_rng(3)
    eps = 1e-5
    for k in range(6):
        d = rng.standard_normal(case.n)
        da = d * f["a"] if k % 3 == 0 else 0 * d
        db = d * f["b"] if k % 3 == 1 else 0 * d
        dd = d * f["dc"] if k % 3 == 2 else 0 * d
        Jp = case.J_oracle(f["a"] + eps * da, f["b"] + eps * db, f["dc"] + eps * dd)
        Jm = case.J_oracle(f["a"] - eps * da, f["b"] - eps * db, f["dc"] - eps * dd)
        fdv = (Jp - Jm) / (2 * eps)
        ad = ga @ da + gb @ db + gd @ dd
        assert abs(fdv - ad) <= 1e-6 * abs(ad), (k, fdv, ad)


def test_oracle_adjoint_rk3bs_and_scalar_layout():
    """The same check for the rk3bs tableau on the one-component layout."""
    c = AdjointCase(ncomp=1, window_yr=20.0)
    f = c.f
    P = c.problem(f["a"], f["b"], f["dc"])
    J, ga, gb, gd = P.adjoint(c.X_start, c.steps, c.recv, c.target, c.weight, 1)
    assert abs(J - c.J_oracle(f["a"], f["b"], f["dc"], method=1)) <= 1e-12 * J
    rng = np.random.default_rng(4)
    eps = 1e-5
    for k in range(3):
        d = rng.standard_normal(c.n)
        da, db, dd = d * f["a"] * (k == 0), d * f["b"] * (k == 1), d * f["dc"] * (k == 2)
        Jp = c.J_oracle(f["a"] + eps * da, f["b"] + eps * db, f["dc"] + eps * dd, method=1)
        Jm = c.J_oracle(f["a"] - eps * da, f["b"] - eps * db, f["dc"] - eps * dd, method=1)
        fdv = (Jp - Jm) / (2 * eps)
        ad = ga @ da + gb @ db + gd @ dd
        assert abs(fdv - ad) <= 1e-6 * abs(ad), (k, fdv, ad)
