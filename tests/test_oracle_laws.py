"""CPU tests that pin the oracle's five state-evolution laws and its normal
stress branch from the mathematics, independently of the restatement.

The reference carries the state in psi = f0 + b ln(v0 theta / D_c) and
evaluates d psi/dt in closed psi form (src/rsf_engine.c:44-132).  The laws
are published in theta form (rsf_solve.md:183-216):

  aging          theta' = 1 - Omega
  slip           theta' = -Omega ln Omega
  PRZ            theta' = (1 - Omega^2) / 2
  Sato-type      theta' = exp(-Omega / beta) - Omega ln Omega
  Kato-Tullis    theta' = exp(-|v| / Vc) - Omega ln Omega
  with Omega = |v| theta / D_c, so d psi/dt = (b / theta) theta'.

The checks below use the theta forms, closed-form solutions at fixed slip
rate, the gate limits of the composite laws and the Kato-Tullis low-speed
steady state Omega = 1 / W(1) (Lambert W), so that a dropped term, a wrong
sign or an exchanged exponent in the psi forms fails one of them.  The
normal-stress branch (src/rsf_engine.c:264-275) is checked against
sigma_dot = K_n v + sinc_n computed with numpy and the limiter invariant.
"""
import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.optimize import brentq
from scipy.special import lambertw

import oracle

V0, F0 = 1e-6, 0.6
BETA = 1e-2          # Sato gate (src/rsf_init.c: sato_beta)
VC = 1e-2 * V0       # Kato-Tullis Vc = v0 / 100


def _problem(law, n=4, b=None, dc=None, vmin=1e-16):
    rng = np.random.default_rng(11)
    a = rng.uniform(0.008, 0.02, n)
    b = rng.uniform(0.01, 0.03, n) if b is None else np.full(n, b)
    dc = rng.uniform(0.01, 0.1, n) if dc is None else np.full(n, dc)
    K = -np.eye(n) * 3e6
    return oracle.RSFProblem(K, a, b, dc, np.zeros(n), law=law, v0=V0, f0=F0, eta=0.0,
                             vmin_state=vmin, sato_beta=BETA, kt_vc=VC, nthreads=1), a, b, dc


def _theta_rate(law, theta, v, dc, vmin=1e-16):
    """published theta' of each law (floor of |v| for laws 2, 4, 5)"""
    va = abs(v)
    if law in (2, 4, 5):
        va = max(va, vmin)
    om = va * theta / dc
    if law == 1:
        return 1.0 - om
    if law == 2:
        return -om * np.log(om)
    if law == 3:
        return 0.5 * (1.0 - om * om)
    if law == 4:
        return np.exp(-om / BETA) - om * np.log(om)
    return np.exp(-va / VC) - om * np.log(om)


@pytest.mark.parametrize("law", [1, 2, 3, 4, 5])
def test_psi_form_equals_published_theta_form(law):
    """d psi/dt of the oracle equals (b / theta) theta' of the published law
    at random states and slip rates of both signs (the laws use |v|)."""
    P, a, b, dc = _problem(law)
    rng = np.random.default_rng(100 + law)
    for _ in range(200):
        i = int(rng.integers(len(b)))
        v = 10 ** rng.uniform(-12, 0) * rng.choice([-1.0, 1.0])
        om = 10 ** rng.uniform(-3, 2)          # Omega across both sides of 1
        theta = om * dc[i] / abs(v)
        psi = F0 + b[i] * np.log(V0 * theta / dc[i])
        S = P.state_rate(i, psi, v)[0]
        ref = b[i] / theta * _theta_rate(law, theta, v, dc[i])
        # scale: the largest term of theta' times b / theta
        terms = b[i] / theta * (1.0 + om * (1.0 + abs(np.log(om))) + om * om)
        assert abs(S - ref) <= 1e-12 * terms, (law, v, om, S, ref)


def test_slip_law_relaxes_exponentially_at_fixed_rate():
    """slip law at fixed v: psi(t) = psi_ss + (psi0 - psi_ss) exp(-v t / D_c),
    psi_ss = f0 - b ln(v / v0); the oracle's rate integrated by scipy."""
    P, a, b, dc = _problem(2)
    i, v = 1, 3e-6
    pss = F0 - b[i] * np.log(v / V0)
    for psi0 in (pss + 0.05, pss - 0.08):
        tau = dc[i] / v
        ts = np.array([0.3, 1.0, 3.0]) * tau
        sol = solve_ivp(lambda t, y: [P.state_rate(i, y[0], v)[0]], (0, ts[-1]), [psi0],
                        t_eval=ts, rtol=1e-12, atol=1e-15, method="DOP853")
        exact = pss + (psi0 - pss) * np.exp(-v * ts / dc[i])
        assert np.abs(sol.y[0] - exact).max() < 1e-9


def test_prz_law_closed_form_at_fixed_rate():
    """PRZ at fixed v: Omega' = (v / 2 D_c)(1 - Omega^2), so
    Omega(t) = tanh(v t / 2 D_c + atanh(Omega0)) below steady state and
    coth(v t / 2 D_c + acoth(Omega0)) above it."""
    P, a, b, dc = _problem(3)
    i, v = 2, 2e-7
    s = v / (2 * dc[i])
    psi_of = lambda om: F0 + b[i] * np.log(V0 * om / v)   # theta = Omega D_c / v
    for om0 in (0.2, 3.0):
        ts = np.array([0.5, 2.0, 6.0]) / s
        sol = solve_ivp(lambda t, y: [P.state_rate(i, y[0], v)[0]], (0, ts[-1]),
                        [psi_of(om0)], t_eval=ts, rtol=1e-12, atol=1e-15, method="DOP853")
        if om0 < 1:
            om = np.tanh(s * ts + np.arctanh(om0))
        else:
            om = 1.0 / np.tanh(s * ts + np.arctanh(1.0 / om0))
        assert np.abs(sol.y[0] - psi_of(om)).max() < 1e-9


def test_laws_one_to_four_share_the_steady_state():
    """psi_ss = f0 - b ln(|v| / v0) zeroes the aging, slip and PRZ rates
    exactly and the Sato-type rate up to its closed gate exp(-1/beta)."""
    for law in (1, 2, 3, 4):
        P, a, b, dc = _problem(law)
        for v in (1e-9, 3e-6, -2e-3):
            pss = F0 - b[0] * np.log(abs(v) / V0)
            S = P.state_rate(0, pss, v)[0]
            scale = b[0] * abs(v) / dc[0]
            bound = 1e-14 + (2 * np.exp(-1.0 / BETA) if law == 4 else 0.0)
            assert abs(S) <= bound * scale, (law, v, S)


def test_composite_laws_reduce_to_the_slip_law_when_the_gate_closes():
    """Sato-type with Omega >> beta and Kato-Tullis with |v| >> Vc: the gate
    exp(-100) is below rounding, so both equal the slip law."""
    Ps, _, b, dc = _problem(2)
    P4 = _problem(4)[0]
    P5 = _problem(5)[0]
    for psi_off, v in ((0.02, 1e-5), (0.05, 3e-3), (-0.03, 3e-3), (0.0, 5e-6)):
        psi = F0 - b[0] * np.log(v / V0) + psi_off
        s2 = Ps.state_rate(0, psi, v)[0]
        if psi_off > 0:
            # Omega = exp(psi_off / b) > 1.9 >> beta: Sato gate below exp(-190)
            assert P4.state_rate(0, psi, v)[0] == pytest.approx(s2, rel=1e-12)
        # |v| >= 5e-6 = 500 Vc: Kato-Tullis gate exp(-500)
        # (at psi_ss the slip rate is zero; ln Omega then rounds to ~1e-16)
        assert P5.state_rate(0, psi, v)[0] == pytest.approx(s2, rel=1e-12,
                                                             abs=1e-14 * v / dc[0])


def test_kato_tullis_low_speed_steady_state():
    """Kato-Tullis steady state at fixed |v|: theta' = g - Omega ln Omega = 0
    with g = exp(-|v| / Vc) gives Omega_ss = g / W(g) (Lambert W); for
    |v| << Vc that is 1 / W(1) = 1.7632, psi_ss shifted up by
    b ln(1 / W(1)) = b W(1) (the offset src/rsf_engine.c:250-253 states)."""
    P, a, b, dc = _problem(5)
    for v in (1e-14, 1e-12, 1e-9, 3e-8):
        g = np.exp(-v / VC)
        om_ss = float(np.real(g / lambertw(g)))
        base = F0 - b[0] * np.log(v / V0)          # Omega = 1
        root = brentq(lambda psi: P.state_rate(0, psi, v)[0], base - b[0], base + 2 * b[0],
                      xtol=1e-15, rtol=1e-15)
        om = np.exp((root - base) / b[0])
        assert om == pytest.approx(om_ss, rel=1e-10), v
    assert float(np.real(1.0 / lambertw(1.0))) == pytest.approx(1.7632228343518968, rel=1e-15)


def test_velocity_floor_and_zero_b():
    """Laws 2, 4, 5 floor |v| at vmin_state (rate and dS/d|v| = 0 below it);
    b = 0 cells have a zero rate for every law."""
    for law in (2, 4, 5):
        P, a, b, dc = _problem(law, vmin=1e-16)
        psi = 0.7
        S_floor, _, dv_floor = P.state_rate(0, psi, 1e-16)
        for v in (1e-20, -3e-18, 0.0):
            S, _, dv = P.state_rate(0, psi, v)
            assert S == S_floor and dv == 0.0
    for law in (1, 2, 3, 4, 5):
        P, a, b, dc = _problem(law, b=0.0)
        assert P.state_rate(0, 0.7, 1e-6) == (0.0, 0.0, 0.0)


# ----------------------------------------------------------- normal stress branch
def _sigma_problem(calc, limit, law=1):
    rng = np.random.default_rng(21)
    n = 8
    K = -np.eye(n) * 3e6 + rng.normal(0, 2e5, (n, n))
    Kn = rng.normal(0, 4e5, (n, n))
    a = rng.uniform(0.008, 0.02, n)
    b = rng.uniform(0.01, 0.02, n)
    dc = rng.uniform(0.01, 0.05, n)
    vpl = 1e-9
    sinc = K @ np.full(n, -vpl)
    sinc_n = Kn @ np.full(n, -vpl)
    P = oracle.RSFProblem(K, a, b, dc, sinc, Kn=Kn, sinc_n=sinc_n, eta=4.6e6, law=law,
                          calc_sigma_dot=calc, limit_sigma=limit, min_sigma=20e6,
                          max_sigma=40e6, nthreads=1)
    sig = rng.uniform(15e6, 45e6, n)
    v = 10 ** rng.uniform(-10, -5, n)
    psi = F0 + b * np.log(V0 / v) + rng.normal(0, 0.01, n)
    tau = sig * a * np.arcsinh(v / (2 * V0) * np.exp(psi / a))
    X = np.stack([psi, tau, sig, np.zeros(n)], 1).ravel()
    return P, K, Kn, sinc, sinc_n, a, X


def test_sigma_dot_is_normal_product_plus_loading():
    """calc_sigma_dot: sigma_dot = K_n v(X) + sinc_n, v from vel_from_rsf
    (numpy product); off: sigma_dot = 0."""
    P, K, Kn, sinc, sinc_n, a, X = _sigma_problem(True, False)
    Y = X.reshape(-1, 4)
    v = 2 * V0 * np.exp(-Y[:, 0] / a) * np.sinh(Y[:, 1] / (Y[:, 2] * a))
    F = P.rhs(X).reshape(-1, 4)
    ref = Kn @ v + sinc_n
    assert np.abs(F[:, 2] - ref).max() <= 1e-13 * np.abs(ref).max()
    P0, *_ = _sigma_problem(False, False)
    assert np.abs(P0.rhs(X).reshape(-1, 4)[:, 2]).max() == 0.0


def test_sigma_limiter_invariant():
    """limit_sigma zeroes sigma_dot exactly where sigma is outside
    [min_sigma, max_sigma] and changes nothing else there; elsewhere the
    whole row equals the unlimited one."""
    P, K, Kn, sinc, sinc_n, a, X = _sigma_problem(True, True)
    Pu, *_ = _sigma_problem(True, False)
    F = P.rhs(X).reshape(-1, 4)
    Fu = Pu.rhs(X).reshape(-1, 4)
    sig = X.reshape(-1, 4)[:, 2]
    out = (sig < 20e6) | (sig > 40e6)
    assert out.any() and (~out).any()
    assert (F[out, 2] == 0.0).all()
    assert np.array_equal(F[~out], Fu[~out])
    assert np.array_equal(F[:, [0, 3]], Fu[:, [0, 3]])


def test_quasi_dynamic_balance_with_sigma_rate():
    """With sigma evolving, tau_dot still solves tau_dot = K v + sinc -
    eta dv/dt, dv/dt now including dv/dsigma sigma_dot (finite differences
    of v along the flow, Richardson extrapolated)."""
    P, K, Kn, sinc, sinc_n, a, X = _sigma_problem(True, False)
    F = P.rhs(X)
    assert np.abs(F.reshape(-1, 4)[:, 2]).max() > 0
    n = len(a)

    def vel(Y):
        Y = Y.reshape(-1, 4)
        return 2 * V0 * np.exp(-Y[:, 0] / a) * np.sinh(Y[:, 1] / (Y[:, 2] * a))

    Fr = F.reshape(-1, 4)
    h = 1e-3 * np.min(np.abs(X.reshape(-1, 4)[:, :3]) / (np.abs(Fr[:, :3]) + 1e-300))
    cd = lambda hh: (vel(X + hh * F) - vel(X - hh * F)) / (2 * hh)
    dvdt = (4.0 * cd(h / 2) - cd(h)) / 3.0
    rhs = K @ vel(X) + sinc - P.p.eta * dvdt
    assert np.abs(Fr[:, 1] - rhs).max() <= 1e-6 * np.abs(Fr[:, 1]).max()
    assert n == 8
