"""Host-side mirror of the rsf_solve driver for the dense GPU path.

``FaultSystem`` owns one librsfg handle and follows the call sequence of
rsf_solve_run (src/rsf_solve.c:85-668): geometry -> operator assembly
(calc_petsc_Isn_matrices) -> per-patch friction -> backslip loading ->
initial state -> adaptive integration with monitor and event output.
All compute runs in librsfg on the GPU; this module only marshals arrays.
"""
import ctypes as C
import os

import numpy as np

from . import lib as L

SEC_PER_YEAR = 31557600.0
G_INTERNAL = 1.0e4   # interact's internal shear modulus (src/includes/properties.h:23)


class FaultSystem:
    def __init__(self, patches, device=0):
        """patches: ctypes array of lib.Patch or a dict of arrays
        (xyz, strike, dip, l, w[, rake, group, vert]); l, w < 0 with the
        9 vertex coordinates in vert marks a triangle."""
        if isinstance(patches, dict):
            patches = L.patches_from_arrays(
                patches["xyz"], patches["strike"], patches["dip"],
                patches["l"], patches["w"], patches.get("rake"),
                patches.get("group"), patches.get("vert"))
        self._lib = L.load()
        self.n = len(patches)
        self._patches = patches
        hh = C.c_void_p()
        L.check(self._lib.rsfg_create(device, patches, self.n, C.byref(hh)),
                "rsfg_create")
        self.h = hh
        self.ncomp = 1
        self.G = 1
        self.precisions = 0
        self.assembly_ms = None
        self._members = []

    def member(self):
        """A member system sharing these patches and the assembled operators
        (rsfg_create_member): own friction, loading and state."""
        m = FaultSystem.__new__(FaultSystem)
        m._lib, m.n, m._patches = self._lib, self.n, self._patches
        hh = C.c_void_p()
        L.check(self._lib.rsfg_create_member(self.h, C.byref(hh)), "rsfg_create_member")
        m.h = hh
        m.ncomp, m.G, m.precisions, m.assembly_ms = self.ncomp, self.G, self.precisions, None
        m.dim = getattr(self, "dim", 2 * self.ncomp + 2)
        m._members = []
        m._parent = self
        self._members.append(m)
        return m

    def close(self):
        for m in getattr(self, "_members", []):
            m.close()
        if self.h and getattr(self, "_owned", True):
            self._lib.rsfg_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------ operator
    def assemble(self, slip_mode=L.SLIP_STRIKE, *, shear_modulus=3.204e10,
                 full_space=False, with_normal=False, store=L.PREC_FP64,
                 scale=None, nf_tol=0.0):
        a = L.Assembly(slip_mode, int(with_normal), int(full_space), store,
                       shear_modulus / G_INTERNAL if scale is None else scale, nf_tol)
        ms = C.c_double()
        L.check(self._lib.rsfg_assemble(self.h, C.byref(a), C.byref(ms)),
                "rsfg_assemble")
        self.ncomp = 2 if slip_mode == L.SLIP_VECTOR else 1
        self.G = self.ncomp + int(bool(with_normal))
        self.precisions = store | L.PREC_FP64
        self.assembly_ms = ms.value
        self.dim = 2 * self.ncomp + 2
        return ms.value

    def operator_bytes(self, prec):
        """bytes of the stored operator that one product with precision prec
        reads (the bf16 near-field format counts its lo blocks)"""
        b = C.c_double()
        L.check(self._lib.rsfg_operator_bytes(self.h, prec, C.byref(b)), "rsfg_operator_bytes")
        return b.value

    def dims(self):
        r, c, ld = C.c_int(), C.c_int(), C.c_int()
        L.check(self._lib.rsfg_operator_dims(self.h, C.byref(r), C.byref(c),
                                             C.byref(ld)), "dims")
        return r.value, c.value, ld.value

    def entries(self, rows, cols):
        rows = np.ascontiguousarray(rows, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        out = np.zeros(len(rows) * len(cols))
        L.check(self._lib.rsfg_entries(self.h, len(rows), rows, len(cols),
                                       cols, out), "rsfg_entries")
        return out.reshape(len(cols), len(rows)).T.copy()

    def dense(self):
        r, c, _ = self.dims()
        return self.entries(np.arange(r), np.arange(c))

    def dense_normal(self):
        """the assembled normal-traction ("In") operator, n x cols (needs
        with_normal at assembly)"""
        r, c, _ = self.dims()
        return self.entries(r + np.arange(self.n), np.arange(c))

    def set_triangle_tolerance(self, vertical_tol=0.0):
        """vertical-edge tolerance [rad] of the triangles' free-surface
        correction for later assemblies and traction_pairs; 0 restores the
        reference's 5e-15 (3e-3 keeps near-vertical edges' far entries
        accurate; see include/rsfg.h)"""
        L.check(self._lib.rsfg_set_triangle_tolerance(self.h, float(vertical_tol)),
                "rsfg_set_triangle_tolerance")

    def traction_pairs(self, recv, src, full_space=False):
        recv = np.ascontiguousarray(recv, np.int32)
        src = np.ascontiguousarray(src, np.int32)
        out = np.zeros(6 * len(recv))
        L.check(self._lib.rsfg_traction_pairs(self.h, len(recv), recv, src,
                                              int(full_space), out),
                "rsfg_traction_pairs")
        return out.reshape(-1, 2, 3)

    def matvec(self, x, prec=L.PREC_FP64):
        _, c, _ = self.dims()
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.ncomp * self.n)
        L.check(self._lib.rsfg_matvec(self.h, prec, x, y), "rsfg_matvec")
        return y

    def matvec_t(self, x, prec=L.PREC_FP64):
        _, c, _ = self.dims()
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(c)
        L.check(self._lib.rsfg_matvec_t(self.h, prec, x, y), "rsfg_matvec_t")
        return y

    def matvec_bench(self, prec=L.PREC_FP64, nrep=50):
        ms = C.c_double()
        L.check(self._lib.rsfg_matvec_bench(self.h, prec, nrep, C.byref(ms)),
                "rsfg_matvec_bench")
        return ms.value

    def read_bench(self, nbytes=4 << 30, nrep=20):
        """ms per launch of a plain read stream over nbytes on this device
        (the products' load pattern without arithmetic)"""
        ms = C.c_double()
        L.check(self._lib.rsfg_read_bench(self.h, float(nbytes), nrep, C.byref(ms)),
                "rsfg_read_bench")
        return ms.value

    # ------------------------------------------------------------ physics
    def set_friction(self, a, b, dc, *, law=1, v0=1e-6, f0=0.6, eta=None,
                     shear_modulus=3.204e10, s_wave_speed=3464.0, rd_fac=1.0,
                     vmin_state=1e-16, sato_beta=1e-2, kt_vc=-1.0,
                     calc_sigma_dot=False, limit_sigma=False, min_sigma=1e6,
                     max_sigma=300e6):
        if eta is None:   # src/rsf_init.c:427
            eta = rd_fac * shear_modulus / (2.0 * s_wave_speed)
        self.eta = eta
        f = L.Friction(law, v0, f0, eta, vmin_state, sato_beta, kt_vc,
                       int(calc_sigma_dot), int(limit_sigma), min_sigma,
                       max_sigma)
        n = self.n
        bc = lambda v: np.ascontiguousarray(np.broadcast_to(v, (n,)), np.float64)
        L.check(self._lib.rsfg_set_friction(self.h, C.byref(f), bc(a), bc(b),
                                            bc(dc)), "rsfg_set_friction")

    def set_loading(self, vpl=1e-9, vpl_patch=None, taudot_add=None,
                    sigdot_add=None):
        keep = []

        def ptr(v):
            if v is None:
                return None
            v = np.ascontiguousarray(np.broadcast_to(v, (self.n,)), np.float64)
            keep.append(v)
            return v.ctypes.data_as(C.c_void_p)

        L.check(self._lib.rsfg_set_loading(self.h, vpl, ptr(vpl_patch),
                                           ptr(taudot_add), ptr(sigdot_add)),
                "rsfg_set_loading")

    def rhs(self, X, prec=L.PREC_FP64):
        X = np.ascontiguousarray(X, np.float64)
        F = np.zeros_like(X)
        ok = C.c_int()
        L.check(self._lib.rsfg_rhs(self.h, prec, X, F, C.byref(ok)),
                "rsfg_rhs")
        return F, bool(ok.value)

    def set_state(self, X, t=0.0):
        L.check(self._lib.rsfg_set_state(self.h, t,
                                         np.ascontiguousarray(X, np.float64)),
                "rsfg_set_state")

    def write_checkpoint(self, fname, *, slip_mode=0, law=1, step=0, dt=0.0):
        """rsf_solve checkpoint of the current state (PETSc binary)."""
        L.check(self._lib.rsfg_write_checkpoint(self.h, os.fsencode(fname), slip_mode, law,
                                                step, dt), "rsfg_write_checkpoint")

    def read_checkpoint(self, fname):
        """Loads the state of an rsf_solve checkpoint; returns its metadata."""
        sm, law, step = C.c_int(), C.c_int(), C.c_long()
        t, dt = C.c_double(), C.c_double()
        L.check(self._lib.rsfg_read_checkpoint(self.h, os.fsencode(fname), C.byref(sm),
                                               C.byref(law), C.byref(step), C.byref(t),
                                               C.byref(dt)), "rsfg_read_checkpoint")
        return dict(slip_mode=sm.value, law=law.value, step=step.value, t=t.value, dt=dt.value)

    # ---- sensitivities (discrete adjoint of the replayed integration)
    def set_step_log(self, capacity):
        """Keep the accepted step sizes of the following solves (0: off)."""
        L.check(self._lib.rsfg_set_step_log(self.h, int(capacity)), "rsfg_set_step_log")

    def step_log(self):
        """Accepted step sizes of the last solve [s]."""
        k = C.c_longlong()
        L.check(self._lib.rsfg_get_step_log(self.h, None, 0, C.byref(k)), "rsfg_get_step_log")
        out = np.zeros(k.value)
        if k.value:
            L.check(self._lib.rsfg_get_step_log(self.h, out.ctypes.data, k.value, C.byref(k)),
                    "rsfg_get_step_log")
        return out

    def _adjoint_problem(self, nsteps, recv, target, weight, method, prec, checkpoint_every):
        keep = []
        recv = np.ascontiguousarray(recv, np.int32)
        keep.append(recv)
        pb = L.AdjointProblem(method, prec, checkpoint_every, len(recv),
                              recv.ctypes.data if len(recv) else None, None, None)
        if target is not None:
            t = np.ascontiguousarray(target, np.float64).reshape(nsteps + 1, len(recv))
            w = np.ascontiguousarray(weight, np.float64).reshape(nsteps + 1)
            keep += [t, w]
            pb.target, pb.weight = t.ctypes.data, w.ctypes.data
        return pb, keep

    def replay(self, steps, recv=(), target=None, weight=None, *, method=L.RK5DP,
               prec=L.PREC_FP64):
        """Replays the steps from the current state (left unchanged):
        dict(J (None without a target), trace [(N+1) x nrecv] ln|v| at the
        receivers, X_end)."""
        steps = np.ascontiguousarray(steps, np.float64)
        N = len(steps)
        pb, keep = self._adjoint_problem(N, recv, target, weight, method, prec, 0)
        trace = np.zeros((N + 1) * max(pb.nrecv, 1))
        Xe = np.zeros(self.n * self.dim)
        J = C.c_double()
        L.check(self._lib.rsfg_replay(self.h, steps, N, C.byref(pb),
                                      C.byref(J) if target is not None else None,
                                      trace.ctypes.data, Xe.ctypes.data), "rsfg_replay")
        return dict(J=J.value if target is not None else None,
                    trace=trace.reshape(N + 1, -1)[:, :pb.nrecv].copy(), X_end=Xe)

    def adjoint_gradient(self, steps, recv, target, weight, *, method=L.RK5DP,
                         prec=L.PREC_FP64, checkpoint_every=256):
        """J and dJ/da, dJ/db, dJ/dD_c per patch of the replay misfit
        (rsfg_adjoint_gradient)."""
        steps = np.ascontiguousarray(steps, np.float64)
        N = len(steps)
        pb, keep = self._adjoint_problem(N, recv, target, weight, method, prec,
                                         checkpoint_every)
        ga, gb, gd = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        J = C.c_double()
        L.check(self._lib.rsfg_adjoint_gradient(self.h, steps, N, C.byref(pb), C.byref(J),
                                                ga, gb, gd), "rsfg_adjoint_gradient")
        return dict(J=J.value, grad_a=ga, grad_b=gb, grad_dc=gd)

    def adjoint_bench(self, prec=L.PREC_FP64, nrep=20):
        """mean duration [ms] of the backward sweep's K^T w product"""
        ms = C.c_double()
        L.check(self._lib.rsfg_adjoint_bench(self.h, prec, nrep, C.byref(ms)),
                "rsfg_adjoint_bench")
        return ms.value

    def get_state(self):
        X = np.zeros(self.n * self.dim)
        t = C.c_double()
        L.check(self._lib.rsfg_get_state(self.h, C.byref(t), X),
                "rsfg_get_state")
        return X, t.value

    @staticmethod
    def solver_options(**kw):
        o = L.SolverOpts()
        L.load().rsfg_solver_defaults(C.byref(o))
        for k, v in kw.items():
            if k == "atol":
                for i, a in enumerate(v):
                    o.atol[i] = a
            else:
                setattr(o, k, v)
        return o

    def solve(self, t_end, **kw):
        o = self.solver_options(**kw)
        st = L.Stats()
        L.check(self._lib.rsfg_solve(self.h, C.byref(o), t_end, C.byref(st)),
                "rsfg_solve")
        return {k: getattr(st, k) for k, _ in L.Stats._fields_}

    def solve_continue(self, t_end, **kw):
        """Continues the last solve to t_end (rsfg_solve_continue): the same
        accepted steps as one solve to the final t_end; method and prec must
        be those of the continued solve."""
        o = self.solver_options(**kw)
        st = L.Stats()
        L.check(self._lib.rsfg_solve_continue(self.h, C.byref(o), t_end, C.byref(st)),
                "rsfg_solve_continue")
        return {k: getattr(st, k) for k, _ in L.Stats._fields_}

    def stage_bench(self, prec=L.PREC_FP64, nrep=50):
        ms = C.c_double()
        L.check(self._lib.rsfg_stage_bench(self.h, prec, nrep, C.byref(ms)),
                "rsfg_stage_bench")
        return ms.value

    def events(self, cap=4096):
        rows = np.zeros(8 * cap)
        n = C.c_int(cap)
        L.check(self._lib.rsfg_events(self.h, rows, C.byref(n)), "events")
        return rows[:8 * n.value].reshape(-1, 8).copy()

    def catalog(self, cap=4096):
        """rows of rsfg_catalog (11 columns: onset t [s], arrest t [s],
        ruptured cells, area, mean and max |slip change|, mean and max
        stress drop [Pa], peak |v|, M0, Mw)"""
        rows = np.zeros(11 * cap)
        n = C.c_int(cap)
        L.check(self._lib.rsfg_catalog(self.h, rows, C.byref(n)), "catalog")
        return rows[:11 * n.value].reshape(-1, 11).copy()

    def monitor(self, cap=1 << 18):
        rows = np.zeros(10 * cap)
        n = C.c_int(cap)
        L.check(self._lib.rsfg_monitor(self.h, rows, C.byref(n)), "monitor")
        return rows[:10 * n.value].reshape(-1, 10).copy()


class Ensemble:
    """Batched runs over shared operators (rsfg_ensemble_*): members_per_geom
    members for each geometry (FaultSystems with an FP32 operator).  Members
    are configured like single systems (set_friction, set_loading,
    set_state) and read out the same way (events, get_state)."""

    def __init__(self, geoms, members_per_geom=32):
        self._lib = L.load()
        self.geoms = list(geoms)
        arr = (C.c_void_p * len(self.geoms))(*[g.h for g in self.geoms])
        e = C.c_void_p()
        L.check(self._lib.rsfg_ensemble_create(arr, len(self.geoms), members_per_geom,
                                               C.byref(e)), "rsfg_ensemble_create")
        self.e = e
        self.nm = members_per_geom
        self.prec = L.PREC_FP32
        self.members = []
        for gi, g in enumerate(self.geoms):
            for k in range(members_per_geom):
                hh = C.c_void_p()
                L.check(self._lib.rsfg_ensemble_member(self.e, gi * members_per_geom + k,
                                                       C.byref(hh)), "rsfg_ensemble_member")
                m = FaultSystem.__new__(FaultSystem)
                m._lib, m.n, m._patches = self._lib, g.n, g._patches
                m.h, m._owned, m._members = hh, False, []
                m.ncomp, m.G, m.precisions, m.assembly_ms = g.ncomp, g.G, g.precisions, None
                m.dim = getattr(g, "dim", 2 * g.ncomp + 2)
                self.members.append(m)

    def set_operator(self, prec, nf_tol=0.0):
        """Operator of the batched product: PREC_FP32 or PREC_BF16NF (bf16 hi
        tiles plus bf16 lo in the near-field segments of each row, chosen with
        energy tolerance nf_tol, <= 0 for 2^-6).  Returns the operator bytes
        one product streams."""
        b = C.c_double()
        L.check(self._lib.rsfg_ensemble_set_operator(self.e, prec, nf_tol, C.byref(b)),
                "rsfg_ensemble_set_operator")
        self.prec = prec
        return b.value

    def solve(self, t_end, **kw):
        kw.setdefault("prec", self.prec)
        o = FaultSystem.solver_options(**kw)
        st = (L.Stats * len(self.members))()
        L.check(self._lib.rsfg_ensemble_solve(self.e, C.byref(o), t_end, st),
                "rsfg_ensemble_solve")
        return [{k: getattr(s, k) for k, _ in L.Stats._fields_} for s in st]

    def tune(self, **kw):
        """Times every tile shape on this ensemble's stage and keeps the
        fastest: returns (choice, per-shape ms, shapes)."""
        o = FaultSystem.solver_options(**kw)
        ch = C.c_int()
        ncfg = C.c_int()
        self._lib.rsfg_ensemble_config(self.e, None, C.byref(ncfg), None)
        ms = (C.c_double * ncfg.value)()
        L.check(self._lib.rsfg_ensemble_tune(self.e, C.byref(o), C.byref(ch), ms),
                "rsfg_ensemble_tune")
        return ch.value, list(ms), self.shapes()

    def bench(self, nrep=20, **kw):
        """(product ms, epilogue ms) of the current shape by HIP events"""
        o = FaultSystem.solver_options(**kw)
        mp, me = C.c_double(), C.c_double()
        L.check(self._lib.rsfg_ensemble_bench(self.e, C.byref(o), nrep, C.byref(mp), C.byref(me)),
                "rsfg_ensemble_bench")
        return mp.value, me.value

    def shapes(self):
        ncfg, cur = C.c_int(), C.c_int()
        self._lib.rsfg_ensemble_config(self.e, C.byref(cur), C.byref(ncfg), None)
        sh = (C.c_int * (7 * ncfg.value))()
        self._lib.rsfg_ensemble_config(self.e, None, None, sh)
        return [tuple(sh[7 * i:7 * i + 7]) for i in range(ncfg.value)]

    def select(self, cfg):
        """use candidate shape cfg (an index into shapes()) without tuning"""
        L.check(self._lib.rsfg_ensemble_select(self.e, cfg), "rsfg_ensemble_select")

    def config(self):
        cur = C.c_int()
        self._lib.rsfg_ensemble_config(self.e, C.byref(cur), None, None)
        return cur.value

    def close(self):
        for m in self.members:
            m.h = None
        if self.e:
            self._lib.rsfg_ensemble_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_patch_file(fname):
    """Patches of a reference geometry file (rsfg_read_patch_file): a dict
    of arrays xyz, strike, dip (float32, as read), l, w, group, vert (nine
    vertex coordinates of triangles, zeros for rectangles)."""
    arr, k = L.read_patch_file(fname)
    return dict(xyz=np.array([[p.x, p.y, p.z] for p in arr[:k]]).reshape(k, 3),
                strike=np.array([p.strike for p in arr[:k]], np.float32),
                dip=np.array([p.dip for p in arr[:k]], np.float32),
                l=np.array([p.l for p in arr[:k]]), w=np.array([p.w for p in arr[:k]]),
                group=np.array([p.group for p in arr[:k]], np.int32),
                vert=np.array([list(p.vert) for p in arr[:k]]).reshape(k, 9))


def read_table(fname, n, ncols):
    """Per-patch input table of rsf_solve (rsf "a b", ic "tau v", dc,
    sigma0, vpl, "tau_dot sigma_dot"): an (n, ncols) array."""
    out = np.zeros(n * ncols)
    L.check(L.load().rsfg_read_table(os.fsencode(fname), n, ncols, out), "rsfg_read_table")
    return out.reshape(n, ncols)


def initial_state(tau, sigma, vel, a, v0, ncomp=1):
    """State from per-cell (tau, sigma, v), src/rsf_solve.c:378-423:
    psi = a ln(2 v0 / v sinh(tau / (sigma a))), slip = 0."""
    tau, sigma, vel, a = (np.asarray(v, np.float64) for v in (tau, sigma, vel, a))
    n = len(a)
    psi = a * np.log(2.0 * v0 / vel * np.sinh(tau / sigma / a))
    if ncomp == 1:
        X = np.zeros((n, 4))
        X[:, 0], X[:, 1], X[:, 2] = psi, tau, sigma
    else:
        X = np.zeros((n, 6))
        X[:, 0], X[:, 1], X[:, 3] = psi, tau, sigma
    return X.ravel()
