"""CPU oracle for the rsf_solve dense hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker / the timed
CPU baseline.  The product (``interact_amd``) never imports it.

Python bindings (ctypes) for ``librsfo.so`` built from ``rsfo_okada.c`` and
``rsfo_rsf.c`` in this directory (``make -C oracle``).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SEC_PER_YEAR = 31557600.0


def build(force=False):
    """Compile librsfo.so (gcc, OpenMP)."""
    so = os.path.join(_HERE, "librsfo.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "librsfo.so")
        if not os.path.exists(so):
            build()
        _LIB = C.CDLL(so)
        _setup(_LIB)
    return _LIB


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Problem(C.Structure):
    _fields_ = [("n", C.c_int), ("ncomp", C.c_int),
                ("K", C.c_void_p), ("Kn", C.c_void_p),
                ("a", C.c_void_p), ("b", C.c_void_p), ("dc", C.c_void_p),
                ("sinc", C.c_void_p), ("sinc_n", C.c_void_p),
                ("v0", C.c_double), ("f0", C.c_double), ("eta", C.c_double),
                ("vmin_state", C.c_double), ("sato_beta", C.c_double),
                ("kt_vc", C.c_double), ("law", C.c_int),
                ("calc_sigma_dot", C.c_int), ("limit_sigma", C.c_int),
                ("min_sigma", C.c_double), ("max_sigma", C.c_double),
                ("nthreads", C.c_int)]


class Options(C.Structure):
    _fields_ = [("method", C.c_int), ("rtol", C.c_double),
                ("atol", C.c_double * 6), ("t0", C.c_double),
                ("t_end", C.c_double), ("dt_init", C.c_double),
                ("dt_max", C.c_double), ("dt_monitor", C.c_double),
                ("rdx_monitor", C.c_double), ("adx_monitor", C.c_double),
                ("vel_event", C.c_double), ("vel_event_hyst", C.c_double),
                ("rupture_vth", C.c_double), ("max_steps", C.c_long),
                ("max_domain_reject", C.c_int), ("max_events", C.c_int),
                ("max_monitor", C.c_int)]


class Result(C.Structure):
    _fields_ = [("steps", C.c_long), ("rejects", C.c_long),
                ("domain_rejects", C.c_long), ("rhs_evals", C.c_long),
                ("t_final", C.c_double), ("dt_final", C.c_double),
                ("status", C.c_int), ("nevent", C.c_int),
                ("ev", C.c_void_p), ("nmon", C.c_int), ("mon", C.c_void_p),
                ("hlog", C.c_void_p), ("hlog_cap", C.c_long),
                ("ncat", C.c_int), ("cat", C.c_void_p), ("cat_area", C.c_void_p),
                ("cat_mu", C.c_double)]


def _setup(L):
    L.rsfo_dc3d.restype = C.c_int
    L.rsfo_dc3d.argtypes = [C.c_double] * 12 + [C.c_int, _dp]
    L.rsfo_patch_frame.restype = None
    L.rsfo_patch_frame.argtypes = [C.c_float, C.c_float, _dp]
    L.rsfo_traction_tensor.restype = None
    L.rsfo_traction_tensor.argtypes = [C.c_int, _dp, _fp, _fp, _dp, _dp,
                                       C.c_int, C.c_int, _dp, C.c_void_p]
    L.rsfo_tri_frame.restype = None
    L.rsfo_tri_frame.argtypes = [_dp] * 7
    L.rsfo_tri_stress.restype = C.c_int
    L.rsfo_tri_stress.argtypes = [_dp] * 4 + [C.c_double] * 6 + [_dp]
    L.rsfo_set_tri_vertical_tol.restype = None
    L.rsfo_set_tri_vertical_tol.argtypes = [C.c_double]
    L.rsfo_mesh_traction_pairs.restype = None
    L.rsfo_mesh_traction_pairs.argtypes = [C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, C.c_int,
                                           C.c_int, C.c_longlong, _ip, _ip, _dp]
    L.rsfo_mesh_traction_tensor.restype = None
    L.rsfo_mesh_traction_tensor.argtypes = [C.c_int, _ip, _dp, _dp, _dp, _dp,
                                            _dp, C.c_int, C.c_int, _dp,
                                            C.c_void_p]
    L.rsfo_read_patches.restype = C.c_int
    L.rsfo_read_patches.argtypes = [C.c_char_p, C.c_int, _dp, _fp, _fp, _dp,
                                    _dp, _ip]
    L.rsfo_vel.restype = C.c_double
    L.rsfo_vel.argtypes = [C.c_double] * 5
    L.rsfo_state_rate.restype = C.c_double
    L.rsfo_state_rate.argtypes = [C.POINTER(Problem), C.c_int, C.c_double,
                                  C.c_double, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]
    L.rsfo_rhs.restype = None
    L.rsfo_rhs.argtypes = [C.POINTER(Problem), _dp, _dp, C.c_void_p,
                           C.c_void_p]
    L.rsfo_domain_ok.restype = C.c_int
    L.rsfo_domain_ok.argtypes = [C.POINTER(Problem), _dp]
    L.rsfo_patch_rhs_c.restype = None
    L.rsfo_patch_rhs_c.argtypes = [C.POINTER(Problem), C.c_int, _dp, _dp, _dp, _dp]
    L.rsfo_patch_jacobian.restype = None
    L.rsfo_patch_jacobian.argtypes = [C.POINTER(Problem), C.c_int, _dp, _dp] + [_dp] * 6 + \
        [C.POINTER(C.c_double)]
    L.rsfo_replay.restype = None
    L.rsfo_replay.argtypes = [C.POINTER(Problem), C.c_int, _dp, _dp, C.c_long, _ip, C.c_int,
                              C.c_void_p, C.c_void_p]
    L.rsfo_adjoint.restype = C.c_double
    L.rsfo_adjoint.argtypes = [C.POINTER(Problem), C.c_int, _dp, _dp, C.c_long, _ip, C.c_int,
                               _dp, _dp, _dp, _dp, _dp]
    L.rsfo_solve.restype = C.c_int
    L.rsfo_solve.argtypes = [C.POINTER(Problem), C.POINTER(Options), _dp,
                             C.POINTER(Result)]
    L.rsfo_imex_parts.restype = None
    L.rsfo_imex_parts.argtypes = [C.POINTER(Problem), _dp, _dp, _dp]
    L.rsfo_imex_block.restype = None
    L.rsfo_imex_block.argtypes = [C.POINTER(Problem), C.c_int, _dp, _dp]
    L.rsfo_imex_newton.restype = C.c_int
    L.rsfo_imex_newton.argtypes = [C.POINTER(Problem), C.c_int, _dp, C.c_double,
                                   C.c_double, _dp, _dp, _dp]
    L.rsfo_rk_attempt.restype = C.c_int
    L.rsfo_rk_attempt.argtypes = [C.POINTER(Problem), C.c_int, C.c_double,
                                  _dp, C.c_double, _dp, _dp, _dp, _dp,
                                  C.POINTER(C.c_double)]


# ------------------------------------------------------------------ geometry
def read_patches(fname, nmax=10_000_000):
    """Parse a reference patch file; returns dict of arrays."""
    n_lines = sum(1 for _ in open(fname))
    nmax = min(nmax, n_lines)
    xyz = np.zeros(3 * nmax)
    st = np.zeros(nmax, np.float32)
    dp = np.zeros(nmax, np.float32)
    ll = np.zeros(nmax)
    ww = np.zeros(nmax)
    gr = np.zeros(nmax, np.int32)
    n = lib().rsfo_read_patches(fname.encode(), nmax, xyz, st, dp, ll, ww, gr)
    if n < 0:
        raise OSError(fname)
    return dict(xyz=xyz[:3 * n].reshape(n, 3).copy(), strike=st[:n].copy(),
                dip=dp[:n].copy(), l=ll[:n].copy(), w=ww[:n].copy(),
                group=gr[:n].copy())


def dc3d(alpha, x, y, z, depth, dip, al1, al2, aw1, aw2, d1, d2, full=0):
    u = np.zeros(12)
    iret = lib().rsfo_dc3d(alpha, x, y, z, depth, dip, al1, al2, aw1, aw2,
                           d1, d2, int(full), u)
    return u, iret


def patch_frame(strike, dip):
    fr = np.zeros(13)
    lib().rsfo_patch_frame(float(strike), float(dip), fr)
    return fr


def traction_tensor(g, full_space=0, nthreads=8, want_iret=False):
    """T[i, j, slip(strike,dip), trac(strike,dip,normal)], internal units."""
    n = len(g["l"])
    T = np.zeros(n * n * 6)
    ir = np.zeros(2 * n * n, np.int32) if want_iret else None
    lib().rsfo_traction_tensor(
        n, np.ascontiguousarray(g["xyz"], np.float64).ravel(),
        np.ascontiguousarray(g["strike"], np.float32),
        np.ascontiguousarray(g["dip"], np.float32),
        np.ascontiguousarray(g["l"], np.float64),
        np.ascontiguousarray(g["w"], np.float64), int(full_space),
        int(nthreads), T,
        ir.ctypes.data_as(C.c_void_p) if want_iret else None)
    T = T.reshape(n, n, 2, 3)
    if want_iret:
        return T, ir.reshape(n, n, 2)
    return T


def tri_frame(vert):
    """(normal, strike, dip, area) of a triangle, vert = 9 coordinates."""
    v = np.ascontiguousarray(vert, np.float64).ravel()
    vn, vs, vd, ar = np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(1)
    lib().rsfo_tri_frame(v[0:3].copy(), v[3:6].copy(), v[6:9].copy(), vn, vs, vd, ar)
    return vn, vs, vd, float(ar[0])


def tri_stress(P, vert, ss, ds, ts, nu=0.25, mu=1.0e4):
    """Half-space stress tensor at P from slip on a triangle; (S, rc)."""
    v = np.ascontiguousarray(vert, np.float64).ravel()
    S = np.zeros(9)
    lam = 2.0 * mu * nu / (1.0 - 2.0 * nu)
    rc = lib().rsfo_tri_stress(np.ascontiguousarray(P, np.float64), v[0:3].copy(),
                               v[3:6].copy(), v[6:9].copy(), ss, ds, ts, nu, 2.0 * mu, lam, S)
    return S.reshape(3, 3), rc


def mesh_receivers(g):
    """(is_tri, receiver points, 13-frames) of a mesh dict: rectangles use
    their centroid and float-angle frame, triangles (l, w < 0, vertices in
    g["vert"]) their vertex mean and TDCS vectors, as read_geometry does
    (src/read_geometry.c:148-215, get_tri_prop_based_on_gh,
    src/green/eval_triangle_nw.c:99-140)."""
    n = len(g["l"])
    l, w = np.asarray(g["l"]), np.asarray(g["w"])
    is_tri = ((l < 0) & (w < 0)).astype(np.int32)
    xyz = np.array(g["xyz"], np.float64).reshape(n, 3)
    fr = np.zeros((n, 13))
    vert = np.asarray(g["vert"], np.float64).reshape(n, 9) if "vert" in g else np.zeros((n, 9))
    for i in range(n):
        if is_tri[i]:
            v = vert[i]
            xyz[i] = (v[0:3] + v[3:6] + v[6:9]) / 3.0
            vn, vs, vd, _ = tri_frame(v)
            fr[i, 2:5], fr[i, 5:8], fr[i, 8:11] = vn, vs, vd
        else:
            fr[i] = patch_frame(g["strike"][i], g["dip"][i])
    return is_tri, xyz, fr, vert


def set_tri_vertical_tol(tol=0.0):
    """vertical-edge tolerance [rad] of the triangles' free-surface
    correction for later evaluations; 0 restores the reference's 5e-15"""
    lib().rsfo_set_tri_vertical_tol(float(tol))


def mesh_traction_pairs(g, recv, src, full_space=0, nthreads=8, rx=None):
    """T[p, slip, trac] for the (recv[p], src[p]) pairs of a mesh of
    rectangles and triangles; rx: mesh_receivers(g), when already formed"""
    n = len(g["l"])
    is_tri, xyz, fr, vert = mesh_receivers(g) if rx is None else rx
    recv = np.ascontiguousarray(recv, np.int32)
    src = np.ascontiguousarray(src, np.int32)
    assert recv.shape == src.shape and recv.min() >= 0 and src.min() >= 0
    assert recv.max() < n and src.max() < n
    T = np.zeros(len(recv) * 6)
    lib().rsfo_mesh_traction_pairs(
        n, is_tri, xyz.ravel(), fr.ravel(), np.ascontiguousarray(g["l"], np.float64),
        np.ascontiguousarray(g["w"], np.float64), vert.ravel(), int(full_space),
        int(nthreads), len(recv), recv, src, T)
    return T.reshape(-1, 2, 3)


def mesh_traction_tensor(g, full_space=0, nthreads=8, want_iret=False):
    """T[i, j, slip, trac] for a mesh of rectangles and triangles."""
    n = len(g["l"])
    is_tri, xyz, fr, vert = mesh_receivers(g)
    T = np.zeros(n * n * 6)
    ir = np.zeros(2 * n * n, np.int32) if want_iret else None
    lib().rsfo_mesh_traction_tensor(
        n, is_tri, xyz.ravel(), fr.ravel(), np.ascontiguousarray(g["l"], np.float64),
        np.ascontiguousarray(g["w"], np.float64), vert.ravel(), int(full_space),
        int(nthreads), T, ir.ctypes.data_as(C.c_void_p) if want_iret else None)
    T = T.reshape(n, n, 2, 3)
    if want_iret:
        return T, ir.reshape(n, n, 2)
    return T


# ------------------------------------------------------------------- physics
class RSFProblem:
    """Keeps numpy arrays alive behind the ctypes Problem struct."""

    def __init__(self, K, a, b, dc, sinc, *, ncomp=1, Kn=None, sinc_n=None,
                 v0=1e-6, f0=0.6, eta=0.0, law=1, vmin_state=1e-16,
                 sato_beta=1e-2, kt_vc=None, calc_sigma_dot=False,
                 limit_sigma=False, min_sigma=1e6, max_sigma=300e6,
                 nthreads=8):
        self._keep = []

        def arr(x):
            if x is None:
                return None
            x = np.ascontiguousarray(x, np.float64)
            self._keep.append(x)
            return x.ctypes.data

        n = len(a)
        self.n, self.ncomp = n, ncomp
        self.p = Problem(n, ncomp, arr(K), arr(Kn), arr(a), arr(b), arr(dc),
                         arr(sinc), arr(sinc_n), v0, f0, eta, vmin_state,
                         sato_beta, 1e-2 * v0 if kt_vc is None else kt_vc,
                         law, int(calc_sigma_dot), int(limit_sigma),
                         min_sigma, max_sigma, nthreads)

    @property
    def dim(self):
        return 6 if self.ncomp == 2 else 4

    def rhs(self, X):
        X = np.ascontiguousarray(X, np.float64)
        F = np.zeros_like(X)
        lib().rsfo_rhs(C.byref(self.p), X, F, None, None)
        return F

    def domain_ok(self, X):
        return bool(lib().rsfo_domain_ok(C.byref(self.p),
                                         np.ascontiguousarray(X, np.float64)))

    def state_rate(self, i, psi, v):
        d1, d2 = C.c_double(), C.c_double()
        S = lib().rsfo_state_rate(C.byref(self.p), i, psi, v, C.byref(d1),
                                  C.byref(d2))
        return S, d1.value, d2.value

    def imex_parts(self, X):
        """(G, Fimpl) of the IMEX split (scalar layout); G + Fimpl = rhs."""
        X = np.ascontiguousarray(X, np.float64)
        G, Fi = np.zeros_like(X), np.zeros_like(X)
        lib().rsfo_imex_parts(C.byref(self.p), X, G, Fi)
        return G, Fi

    def imex_block(self, i, x):
        """dFimpl/dX of cell i at its state x (4x4)."""
        J = np.zeros(16)
        lib().rsfo_imex_block(C.byref(self.p), i, np.ascontiguousarray(x, np.float64), J)
        return J.reshape(4, 4)

    def imex_newton(self, i, Z, hg, guess, rtol=1e-8, atol=(1e-12, 1e-3, 1e-3, 1e-3)):
        """Stage equation Y = Z + hg Fimpl(Y) of cell i; returns (Y, Fi, iterations)."""
        y = np.zeros(4)
        y[:2] = guess[:2]
        fi = np.zeros(4)
        it = lib().rsfo_imex_newton(C.byref(self.p), i, np.ascontiguousarray(Z, np.float64),
                                    hg, rtol, np.ascontiguousarray(atol, np.float64), y, fi)
        return y, fi, it

    def rk_attempt(self, X, F0, h, method=0, rtol=1e-8, atol=None):
        atol = np.asarray(atol if atol is not None else
                          [1e-12, 1e-3, 1e-3, 1e-3, 1e-3, 1e-3][:self.dim],
                          np.float64)
        Xn = np.zeros_like(X)
        Fn = np.zeros_like(X)
        en = C.c_double()
        ok = lib().rsfo_rk_attempt(C.byref(self.p), method, rtol,
                                   np.ascontiguousarray(atol), h,
                                   np.ascontiguousarray(X, np.float64),
                                   np.ascontiguousarray(F0, np.float64), Xn,
                                   Fn, C.byref(en))
        return Xn, Fn, en.value, bool(ok)

    def solve(self, X0, t_end, *, method=0, rtol=1e-8, atol=None, t0=0.0,
              dt_init=1.0, dt_max=1e10, dt_monitor=5 * SEC_PER_YEAR,
              rdx_monitor=1e-4, adx_monitor=0.0, vel_event=1e-3,
              vel_event_hyst=0.5, rupture_vth=None, max_steps=10_000_000,
              max_domain_reject=5000, max_events=4096, max_monitor=200000,
              log_steps=0, cat_area=None, cat_mu=0.0):
        X = np.array(X0, np.float64, copy=True)
        at = (C.c_double * 6)(*(list(atol) if atol is not None else
                                [1e-12, 1e-3, 1e-3, 1e-3, 1e-3, 1e-3][:self.dim]
                                + [0.0] * (6 - self.dim)))
        o = Options(method, rtol, at, t0, t_end, dt_init, dt_max, dt_monitor,
                    rdx_monitor, adx_monitor, vel_event, vel_event_hyst,
                    vel_event if rupture_vth is None else rupture_vth,
                    max_steps, max_domain_reject, max_events, max_monitor)
        ev = np.zeros(max_events * 8)
        mon = np.zeros(max_monitor * 10)
        r = Result()
        r.ev = ev.ctypes.data
        r.mon = mon.ctypes.data
        cat = np.zeros(max_events * 11)
        r.cat = cat.ctypes.data
        area = None
        if cat_area is not None:
            area = np.ascontiguousarray(cat_area, np.float64)
            r.cat_area = area.ctypes.data
        r.cat_mu = cat_mu
        hlog = np.zeros(max(int(log_steps), 1))
        if log_steps:
            r.hlog = hlog.ctypes.data
            r.hlog_cap = int(log_steps)
        lib().rsfo_solve(C.byref(self.p), C.byref(o), X, C.byref(r))
        out = dict(X=X, steps=r.steps, rejects=r.rejects,
                   domain_rejects=r.domain_rejects, rhs_evals=r.rhs_evals,
                   t_final=r.t_final, dt_final=r.dt_final, status=r.status,
                   events=ev[:8 * r.nevent].reshape(-1, 8).copy(),
                   monitor=mon[:10 * r.nmon].reshape(-1, 10).copy(),
                   catalog=cat[:11 * r.ncat].reshape(-1, 11).copy())
        if log_steps:
            out["hlog"] = hlog[:min(r.steps, int(log_steps))].copy()
        return out

    # ---- discrete adjoint of the replayed integration (rsfo_adj.c)
    def patch_rhs_c(self, i, x, td):
        """f, v of patch i by the complex restatement on the real axis."""
        f, v = np.zeros(self.dim), np.zeros(self.ncomp)
        lib().rsfo_patch_rhs_c(C.byref(self.p), i, np.ascontiguousarray(x, np.float64),
                               np.ascontiguousarray(td, np.float64), f, v)
        return f, v

    def patch_jacobian(self, i, x, td):
        """Complex-step Jacobians of patch i: dict Jx, Jtd, Jp (a, b, dc),
        Vx, Vp, Lx, La."""
        d, nc = self.dim, self.ncomp
        Jx, Jtd, Jp = np.zeros(d * d), np.zeros(d * nc), np.zeros(d * 3)
        Vx, Vp, Lx, La = np.zeros(nc * d), np.zeros(nc), np.zeros(d), C.c_double()
        lib().rsfo_patch_jacobian(C.byref(self.p), i, np.ascontiguousarray(x, np.float64),
                                  np.ascontiguousarray(td, np.float64), Jx, Jtd, Jp, Vx, Vp,
                                  Lx, C.byref(La))
        return dict(Jx=Jx.reshape(d, d), Jtd=Jtd.reshape(d, nc), Jp=Jp.reshape(d, 3),
                    Vx=Vx.reshape(nc, d), Vp=Vp, Lx=Lx, La=La.value)

    def replay(self, X0, steps, recv=(), method=0):
        """Replay the steps from X0: (trace [(N+1) x nrecv] of ln|v|, X_end)."""
        steps = np.ascontiguousarray(steps, np.float64)
        recv = np.ascontiguousarray(recv, np.int32)
        trace = np.zeros((len(steps) + 1) * max(len(recv), 1))
        Xe = np.zeros(len(X0))
        lib().rsfo_replay(C.byref(self.p), method, np.ascontiguousarray(X0, np.float64), steps,
                          len(steps), recv, len(recv), trace.ctypes.data, Xe.ctypes.data)
        return trace.reshape(len(steps) + 1, -1)[:, :len(recv)].copy(), Xe

    def adjoint(self, X0, steps, recv, target, weight, method=0):
        """J and dJ/da, dJ/db, dJ/ddc of the replay misfit."""
        steps = np.ascontiguousarray(steps, np.float64)
        recv = np.ascontiguousarray(recv, np.int32)
        ga, gb, gd = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        J = lib().rsfo_adjoint(C.byref(self.p), method, np.ascontiguousarray(X0, np.float64),
                               steps, len(steps), recv, len(recv),
                               np.ascontiguousarray(target, np.float64).ravel(),
                               np.ascontiguousarray(weight, np.float64), ga, gb, gd)
        return J, ga, gb, gd


def vel(tau, sigma, psi, a, v0):
    return lib().rsfo_vel(tau, sigma, psi, a, v0)
