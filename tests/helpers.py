"""Shared helpers for the parity tests."""
import gzip
import io
import os
import tempfile

import numpy as np

from conftest import GOLDEN

SEC_PER_YEAR = 31557600.0


def geom_from_text(txt):
    """Parse patch text with the oracle's reference-format reader."""
    import oracle
    with tempfile.NamedTemporaryFile("w", suffix=".in", delete=False) as f:
        f.write(txt)
        name = f.name
    try:
        return oracle.read_patches(name)
    finally:
        os.unlink(name)


def mesh_from_text(txt):
    """Parse a patch file with rectangles and triangles (l, w < 0 followed
    by nine vertex coordinates, src/read_geometry.c:148-175) into a mesh
    dict: xyz, strike, dip (float32), l, w, vert (zeros for rectangles)."""
    rows = [line.split() for line in str(txt).strip().split("\n") if line.strip()]
    return dict(
        xyz=np.array([[float(r[0]), float(r[1]), float(r[2])] for r in rows]),
        strike=np.array([np.float32(r[3]) for r in rows]),
        dip=np.array([np.float32(r[4]) for r in rows]),
        l=np.array([float(r[5]) for r in rows]), w=np.array([float(r[6]) for r in rows]),
        vert=np.array([[float(v) for v in r[8:17]] if len(r) >= 17 else [0.0] * 9
                       for r in rows]))


def tri_golden(name):
    """(mesh, T, iret[, T_full, iret_full]) of a triangle fixture set."""
    d = np.load(os.path.join(GOLDEN, "ktensor_tri.npz"))
    out = [mesh_from_text(d[name + "_geom"]), d[name + "_T"], d[name + "_iret"]]
    if name + "_Tfull" in d.files:
        out += [d[name + "_Tfull"], d[name + "_iretfull"]]
    return out


def fvh_geometry():
    with gzip.open(os.path.join(GOLDEN, "geom_fvh.in.gz"), "rt") as f:
        return geom_from_text(f.read())


def bp5_2km():
    import oracle
    d = os.path.join(GOLDEN, "bp5_2km")
    g = oracle.read_patches(os.path.join(d, "geom_bp5_2km.in"))
    ab = np.loadtxt(os.path.join(d, "rsf_bp5_2km.dat"))
    ic = np.loadtxt(os.path.join(d, "ic_bp5_2km.in"))
    dc = np.loadtxt(os.path.join(d, "dc_bp5_2km.in"))
    return g, dict(a=ab[:, 0], b=ab[:, 1], tau=ic[:, 0], vel=ic[:, 1], dc=dc)


def strike_operator(T, scale):
    return T[:, :, 0, 0] * scale


def vector_operator(T, scale):
    n = T.shape[0]
    K = np.zeros((2 * n, 2 * n))
    for t in range(2):
        for k in range(2):
            K[t::2, k::2] = T[:, :, k, t] * scale
    return K


def row_rel(A, B):
    """max_ij |A - B| / max_j |B_ij| (row scale of the operator)."""
    A = np.asarray(A).reshape(B.shape[0], -1)
    B = np.asarray(B).reshape(B.shape[0], -1)
    return float((np.abs(A - B) / np.maximum(np.abs(B).max(1), 1e-300)[:, None]).max())


def initial_state(tau, sigma, vel, a, v0, ncomp=1):
    tau, sigma, vel, a = (np.broadcast_to(np.asarray(v, np.float64), np.shape(a)) for v in (tau, sigma, vel, a))
    psi = a * np.log(2.0 * v0 / vel * np.sinh(tau / sigma / a))
    n = len(a)
    X = np.zeros((n, 4 if ncomp == 1 else 6))
    X[:, 0], X[:, 1] = psi, tau
    X[:, ncomp + 1] = sigma
    return X.ravel()


def first_recurrence(events, t_min_yr=1.0):
    """Onset time [yr] of the first event after t_min (skips a seeded one)."""
    on = events[events[:, 1] > 0, 0] / SEC_PER_YEAR
    on = on[on > t_min_yr]
    return float(on[0]) if len(on) else float("nan")


class AdjointCase:
    """The small sensitivity problem shared by the oracle and device adjoint
    tests: the 40 x 20 km outline at 2 km (200 patches), friction draw 2 of
    the ensemble generator, started just after its first event (13 yr, from
    an oracle solve) and run over an interseismic window; 16 seeded
    receivers; the target is the ln|v| trace of a run with perturbed b and
    D_c (its own adaptive steps), interpolated in time onto the states of
    the model's steps; weights are the step lengths over the window."""

    def __init__(self, ncomp=2, window_yr=40.0, nthreads=8, seed=5):
        import oracle
        from interact_amd import faults
        g = faults.planar_grid(2.0, 40, 20)
        T = oracle.traction_tensor(g, 0, nthreads)
        self.ncomp = ncomp
        self.K = vector_operator(T, 32e9 / 1e4) if ncomp == 2 else strike_operator(T, 32e9 / 1e4)
        self.g = g
        self.n = n = len(g["l"])
        self.f = f = faults.ensemble_fields(g, 40, 20, 2, 0)
        X0 = initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=ncomp)
        P = self.problem(f["a"], f["b"], f["dc"], nthreads)
        r0 = P.solve(X0, 13.0 * SEC_PER_YEAR, rtol=1e-6)
        self.X_start, self.t_start = r0["X"], r0["t_final"]
        ro = P.solve(self.X_start, self.t_start + window_yr * SEC_PER_YEAR, t0=self.t_start,
                     rtol=1e-6, log_steps=100000)
        self.steps, self.X_solve = ro["hlog"], ro["X"]
        rng = np.random.default_rng(seed)
        self.recv = np.sort(rng.choice(n, 16, replace=False)).astype(np.int32)
        bt = f["b"] * (1 + 0.02 * rng.standard_normal(n))
        dt = f["dc"] * (1 + 0.05 * rng.standard_normal(n))
        Pt = self.problem(f["a"], bt, dt, nthreads)
        rt = Pt.solve(self.X_start, self.t_start + 1.01 * window_yr * SEC_PER_YEAR,
                      t0=self.t_start, rtol=1e-6, log_steps=100000)
        trt, _ = Pt.replay(self.X_start, rt["hlog"], self.recv, 0)
        tt = np.concatenate([[0.0], np.cumsum(rt["hlog"])])
        tm = np.concatenate([[0.0], np.cumsum(self.steps)])
        self.target = np.stack([np.interp(tm, tt, trt[:, r]) for r in range(len(self.recv))], 1)
        self.weight = np.zeros(len(self.steps) + 1)
        self.weight[1:] = self.steps / self.steps.sum()

    def problem(self, a, b, dc, nthreads=8):
        import oracle
        f, n, nc = self.f, self.n, self.ncomp
        xpl = np.zeros(nc * n)
        xpl[0::nc] = -f["vpl"]
        return oracle.RSFProblem(self.K, a, b, dc, self.K @ xpl, ncomp=nc, eta=f["eta"],
                                 v0=f["v0"], f0=f["f0"], nthreads=nthreads)

    def J_oracle(self, a, b, dc, steps=None, method=0):
        tr, _ = self.problem(a, b, dc).replay(self.X_start,
                                              self.steps if steps is None else steps,
                                              self.recv, method)
        return 0.5 * np.sum(self.weight[:, None] * (tr - self.target) ** 2)
