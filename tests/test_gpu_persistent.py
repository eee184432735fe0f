"""GPU tests of the persistent stage and product kernels (x held in LDS, one
workgroup per CU) against the per-patch kernels, in every slip layout the
stage kernel is built for: the environment switch RSFG_STAGE_PERSIST=1
selects the persistent kernels wherever they can run (here on faults small
enough that the size rule alone would pick the per-patch kernel), =0 the
per-patch kernels.  Both form every row sum in the same order and share the
epilogue, so the accepted steps and the states agree bit for bit; the event
rows agree in their times and slip-rate columns.  The c3 cycle tests cover
the persistent kernels where the size rule selects them.
"""
import os

import numpy as np
import pytest

from helpers import SEC_PER_YEAR, initial_state

pytestmark = pytest.mark.gpu

G_BP5 = 3.204e10


def _with_persist(flag, fn):
    old = os.environ.get("RSFG_STAGE_PERSIST")
    os.environ["RSFG_STAGE_PERSIST"] = flag
    try:
        return fn()
    finally:
        if old is None:
            del os.environ["RSFG_STAGE_PERSIST"]
        else:
            os.environ["RSFG_STAGE_PERSIST"] = old


def _dipping_grid(ds_km=2.0, nx=40, nz=12, dip=30.0, top_km=1.0):
    d = np.radians(dip)
    i, j = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    i, j = i.ravel(), j.ravel()
    ds = ds_km * 1e3
    s = (j + 0.5) * ds
    xyz = np.stack([s * np.cos(d), (i + 0.5 - nx / 2) * ds, -top_km * 1e3 - s * np.sin(d)], 1)
    n = nx * nz
    return dict(xyz=xyz, strike=np.zeros(n, np.float32), dip=np.full(n, dip, np.float32),
                l=np.full(n, ds / 2), w=np.full(n, ds / 2),
                along=(i + 0.5 - nx / 2) * ds_km, depth=top_km + s / 1e3 * np.sin(d))


def _vector_cycle(store):
    """planar 100 x 40 km fault at 2 km (1000 patches, two slip components)"""
    from interact_amd import faults, lib
    from interact_amd.model import FaultSystem
    g = faults.planar_grid(2.0)
    f = faults.cycle_fields(g)
    fs = FaultSystem(g)
    fs.assemble(lib.SLIP_VECTOR, shear_modulus=f["shear_modulus"], store=store)
    fs.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"])
    fs.set_loading(f["vpl"])
    X0 = initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=2)
    return fs, X0


def _dipping(layout, store):
    """30-degree dipping fault (480 patches) with normal rows and the
    normal-stress evolution on; scalar strike slip or two components"""
    from interact_amd import faults, lib
    from interact_amd.model import FaultSystem
    g = _dipping_grid()
    f = faults.bp5_fields(g)
    fs = FaultSystem(g)
    fs.assemble(layout, shear_modulus=G_BP5, with_normal=True, store=store)
    fs.set_friction(f["a"], f["b"], f["dc"], v0=1e-6, f0=0.6, eta=f["eta"],
                    calc_sigma_dot=True)
    fs.set_loading(1e-9)
    nc = 2 if layout == lib.SLIP_VECTOR else 1
    X0 = initial_state(f["tau"], 25e6, f["vel"], f["a"], 1e-6, ncomp=nc)
    return fs, X0


def _bp5_scalar(store):
    """planar fault at 2 km, scalar strike slip, no normal rows"""
    from interact_amd import faults, lib
    from interact_amd.model import FaultSystem
    g = faults.planar_grid(2.0)
    f = faults.bp5_fields(g, prestress="make_bp5")
    fs = FaultSystem(g)
    fs.assemble(lib.SLIP_STRIKE, shear_modulus=G_BP5, store=store)
    fs.set_friction(f["a"], f["b"], f["dc"], v0=1e-6, f0=0.6, eta=f["eta"])
    fs.set_loading(1e-9)
    X0 = initial_state(f["tau"], 25e6, f["vel"], f["a"], 1e-6)
    return fs, X0


def _compare(fs, X0, t_end, prec, method=None):
    from interact_amd import lib
    kw = dict(rtol=1e-8, prec=prec, steps_per_launch=32)
    if method is not None:
        kw["method"] = method
    out = {}
    for flag in ("1", "0"):
        def run():
            fs.set_state(X0)
            st = fs.solve(t_end, **kw)
            X, t = fs.get_state()
            return st, X.copy(), t, fs.events().copy()
        out[flag] = _with_persist(flag, run)
    (s1, X1, t1, e1), (s0, X0_, t0, e0) = out["1"], out["0"]
    assert s1["status"] == 0 and s0["status"] == 0
    assert (s1["steps"], s1["rejects"]) == (s0["steps"], s0["rejects"])
    assert t1 == t0
    assert np.array_equal(X1, X0_)
    assert e1.shape == e0.shape
    if len(e0):
        assert np.array_equal(e1[:, :3], e0[:, :3]) and np.array_equal(e1[:, 5], e0[:, 5])
    # the products through the same switch
    D = fs.dims()[1]
    x = np.sin(np.arange(D) + 1.0) + 0.3 * np.cos(3.0 * np.arange(D))
    y1 = _with_persist("1", lambda: fs.matvec(x, prec))
    y0 = _with_persist("0", lambda: fs.matvec(x, prec))
    assert np.array_equal(y1, y0)
    return s1


@pytest.mark.parametrize("prec_name", ["fp32", "bf16x2", "bf16", "bf16nf"])
def test_vector_layout_cycle_through_an_event(prec_name):
    """two slip components, two rows per patch: 40 yr through the first
    event near 31.8 yr, every FP32-input format"""
    from interact_amd import accuracy, lib
    store = lib.PREC_FP32 | lib.PREC_BF16X2 | lib.PREC_BF16 | lib.PREC_BF16NF
    fs, X0 = _vector_cycle(store)
    _compare(fs, X0, 40.0 * SEC_PER_YEAR, lib.PREC_BY_NAME[prec_name])
    assert len(accuracy.onsets(fs.events())) == 1
    fs.close()


def test_vector_layout_rk3bs():
    from interact_amd import lib
    fs, X0 = _vector_cycle(lib.PREC_FP32)
    _compare(fs, X0, 5.0 * SEC_PER_YEAR, lib.PREC_FP32, method=lib.RK3BS)
    fs.close()


@pytest.mark.parametrize("layout_name", ["SLIP_STRIKE", "SLIP_VECTOR"])
def test_normal_rows_and_sigma_evolution(layout_name):
    """normal rows (scalar: two rows per patch; two components: three) and
    the normal-stress evolution on a dipping fault, through the seeded
    nucleation"""
    from interact_amd import lib
    fs, X0 = _dipping(getattr(lib, layout_name), lib.PREC_FP32 | lib.PREC_BF16NF)
    for prec in (lib.PREC_FP32, lib.PREC_BF16NF):
        _compare(fs, X0, 0.02 * SEC_PER_YEAR, prec)
    fs.close()


def test_scalar_layout_one_row_per_patch():
    from interact_amd import lib
    fs, X0 = _bp5_scalar(lib.PREC_FP32 | lib.PREC_BF16)
    for prec in (lib.PREC_FP32, lib.PREC_BF16):
        _compare(fs, X0, 2.0 * SEC_PER_YEAR, prec)
    fs.close()


def test_column_blocks_beyond_the_lds_copy():
    """x longer than the LDS copy (36864 floats): a 100 x 48 km fault at
    0.5 km, two slip components, 38400 columns in two column blocks.  The
    persistent kernels add each row's two block sums; the per-patch kernels
    sum the whole row in one go, so the products agree to FP64 rounding
    (not bit for bit), and 0.5 yr of the cycle takes the same steps with
    states equal to 1e-10."""
    from interact_amd import faults, lib
    from interact_amd.model import FaultSystem
    g = faults.planar_grid(0.5, 100.0, 48.0)
    n = len(g["l"])
    assert 2 * n > 36864
    f = faults.cycle_fields(g)
    store = lib.PREC_FP32 | lib.PREC_BF16X2 | lib.PREC_BF16 | lib.PREC_BF16NF
    fs = FaultSystem(g)
    fs.assemble(lib.SLIP_VECTOR, shear_modulus=f["shear_modulus"], store=store)
    i = np.arange(2 * n)
    x = np.sin(i + 1.0) + 0.3 * np.cos(3.0 * i)
    y64 = fs.matvec(x, lib.PREC_FP64)
    for name in ("fp32", "bf16x2", "bf16", "bf16nf"):
        p = lib.PREC_BY_NAME[name]
        y1 = _with_persist("1", lambda: fs.matvec(x, p))
        y0 = _with_persist("0", lambda: fs.matvec(x, p))
        assert np.abs(y1 - y0).max() <= 1e-13 * np.abs(y0).max(), name
        err = np.linalg.norm(y1 - y64) / np.linalg.norm(y64)
        assert err < {"fp32": 1e-7, "bf16x2": 5e-6, "bf16": 5e-3, "bf16nf": 5e-5}[name], name
    fs.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"])
    fs.set_loading(f["vpl"])
    X0 = initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=2)
    out = {}
    for flag in ("1", "0"):
        fs.set_state(X0)
        st = _with_persist(flag, lambda: fs.solve(0.5 * SEC_PER_YEAR, rtol=1e-8,
                                                  prec=lib.PREC_BF16NF))
        out[flag] = (st["steps"], st["rejects"], fs.get_state()[0])
    assert out["1"][:2] == out["0"][:2]
    X1, X0s = out["1"][2].reshape(n, -1), out["0"][2].reshape(n, -1)
    for k in range(X1.shape[1]):    # each field against its own scale
        assert np.abs(X1[:, k] - X0s[:, k]).max() <= 1e-10 * np.abs(X0s[:, k]).max() + 1e-300
    fs.close()


@pytest.mark.parametrize("layout", ["strike_40000", "vector_normal_19200"])
def test_column_blocks_other_layouts(layout):
    """Column blocks with one row per patch (strike slip, 40000 patches:
    40000 columns) and with three (two components and the normal row,
    19200 patches: 38400 columns): persistent products against the
    per-patch kernels to FP64 rounding in every FP32-input format."""
    from interact_amd import faults, lib
    from interact_amd.model import FaultSystem
    store = lib.PREC_FP32 | lib.PREC_BF16X2 | lib.PREC_BF16 | lib.PREC_BF16NF
    if layout == "strike_40000":
        g = faults.planar_grid(0.5, 100.0, 100.0)
        mode, normal = lib.SLIP_STRIKE, False
    else:
        g = faults.planar_grid(0.5, 100.0, 48.0)
        mode, normal = lib.SLIP_VECTOR, True
    fs = FaultSystem(g)
    fs.assemble(mode, shear_modulus=G_BP5, with_normal=normal, store=store)
    rows, cols, _ = fs.dims()
    assert cols > 36864
    i = np.arange(cols)
    x = np.sin(i + 1.0) + 0.3 * np.cos(3.0 * i)
    for name in ("fp32", "bf16x2", "bf16", "bf16nf"):
        p = lib.PREC_BY_NAME[name]
        y1 = _with_persist("1", lambda: fs.matvec(x, p))
        y0 = _with_persist("0", lambda: fs.matvec(x, p))
        assert y1.shape == y0.shape
        assert np.abs(y1 - y0).max() <= 1e-13 * np.abs(y0).max(), (layout, name)
    fs.close()
