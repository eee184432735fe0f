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
