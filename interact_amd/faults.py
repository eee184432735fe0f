"""Seeded fault meshes and friction fields of the benchmark configurations.

Planar meshes follow the layout of the reference's BP5 generator
(seas_bp_tests/bp5/make_bp5.py): a vertical strike-slip fault in the x = 0
plane, strike 0 (strike direction +y), dip 90, square cells of side ds,
patch k = i * nz + j with i along strike and j down dip, half lengths
l = w = ds / 2 [m].
"""
import numpy as np

SEC_PER_YEAR = 31557600.0


def planar_grid(ds_km, length_km=100.0, depth_km=40.0):
    nx, nz = int(round(length_km / ds_km)), int(round(depth_km / ds_km))
    i, j = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    i, j = i.ravel(), j.ravel()
    y = (i + 0.5 - nx / 2.0) * ds_km * 1e3
    z = -(j + 0.5) * ds_km * 1e3
    n = nx * nz
    half = ds_km * 1e3 / 2.0
    xyz = np.stack([np.zeros(n), y, z], axis=1)
    return dict(xyz=xyz, strike=np.zeros(n, np.float32),
                dip=np.full(n, 90.0, np.float32), l=np.full(n, half),
                w=np.full(n, half), nx=nx, nz=nz, ds_km=ds_km,
                along=(i + 0.5 - nx / 2.0) * ds_km, depth=(j + 0.5) * ds_km)


def rough_offset(nx, nz, ds_km, hurst=0.8, alpha=1e-3, seed=20261202):
    """Self-affine random offset field on the (nx + 1) x (nz + 1) node grid
    of a planar_grid outline: 2-D Fourier synthesis with amplitude spectrum
    |k|^-(H + 1) (Hurst exponent H) and seeded Gaussian coefficients (PCG64),
    zero mean, scaled to an rms of alpha times the along-strike length
    (amplitude-to-wavelength ratio alpha at the longest wavelength) [m]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ds = ds_km * 1e3
    kx = np.fft.fftfreq(nx + 1, d=ds)[:, None]
    kz = np.fft.fftfreq(nz + 1, d=ds)[None, :]
    k = np.sqrt(kx**2 + kz**2)
    amp = np.zeros_like(k)
    amp[k > 0] = k[k > 0] ** -(hurst + 1.0)
    coef = rng.standard_normal(k.shape) + 1j * rng.standard_normal(k.shape)
    h = np.real(np.fft.ifft2(coef * amp))
    h -= h.mean()
    return h * (alpha * nx * ds / np.sqrt(np.mean(h**2)))


def planar_tri_grid(ds_km, length_km=100.0, depth_km=40.0, offset=None):
    """The planar_grid outline meshed with two triangles per square cell
    (2 nx nz triangles, triangle 2 (i nz + j) + t for cell (i, j)).  Nodes
    at y = (i - nx/2) ds, z = -j ds, x = offset[i, j] (0 for the plane; the
    rough variant passes rough_offset).  Cell (i, j) with nodes A = (i, j+1),
    B = (i+1, j+1), C = (i, j), D = (i+1, j) splits into (A, B, C) and
    (B, D, C): on the plane their TDCS normal is +x, strike +y and dip +z,
    the rectangles' frame.  Patches carry l = w = -1 (the reference's
    triangle marker), the vertex mean as xyz, and vert[9]."""
    nx, nz = int(round(length_km / ds_km)), int(round(depth_km / ds_km))
    ds = ds_km * 1e3
    off = np.zeros((nx + 1, nz + 1)) if offset is None else np.asarray(offset, np.float64)

    def node(i, j):
        return np.stack([off[i, j], (i - nx / 2.0) * ds, -j * ds * np.ones_like(i * 1.0)], axis=-1)

    i, j = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    i, j = i.ravel(), j.ravel()
    A, B, C, D = node(i, j + 1), node(i + 1, j + 1), node(i, j), node(i + 1, j)
    t1 = np.concatenate([A, B, C], axis=1)
    t2 = np.concatenate([B, D, C], axis=1)
    vert = np.stack([t1, t2], axis=1).reshape(-1, 9)
    n = len(vert)
    xyz = (vert[:, 0:3] + vert[:, 3:6] + vert[:, 6:9]) / 3.0
    return dict(xyz=xyz, strike=np.zeros(n, np.float32), dip=np.full(n, 90.0, np.float32),
                l=np.full(n, -1.0), w=np.full(n, -1.0), vert=vert, nx=nx, nz=nz,
                ds_km=ds_km, along=xyz[:, 1] / 1e3, depth=-xyz[:, 2] / 1e3)


def write_patch_file(fname, g):
    """Reference patch-file format "x y z strike dip L W group"; triangles
    (l, w < 0) append their nine vertex coordinates (src/read_geometry.c:
    148-175), written with round-trip precision."""
    grp = g.get("group", np.zeros(len(g["l"]), int))
    with open(fname, "w") as f:
        for k in range(len(g["l"])):
            x, y, z = g["xyz"][k]
            line = (f"{x:.6e} {y:.6e} {z:.6e} {float(g['strike'][k]):.6f} "
                    f"{float(g['dip'][k]):.6f} {g['l'][k]:.6e} {g['w'][k]:.6e} "
                    f"{int(grp[k])}")
            if g["l"][k] < 0 and g["w"][k] < 0:
                line += " " + " ".join(f"{v:.17e}" for v in g["vert"][k])
            f.write(line + "\n")


def bp5_fields(g, a0=0.004, a_max=0.04, b0=0.03, dc0=0.14, dc_nuc=0.13,
               f0=0.6, v0=1e-6, vpl=1e-9, sigma=25e6, shear_modulus=3.204e10,
               cs=3464.0, vnuc=3e-2, prestress="make_bp5"):
    """SEAS BP5-QD fields.

    prestress="make_bp5": tau = sigma a asinh(.) + eta v, the committed
    ic_bp5_*.in files (seas_bp_tests/bp5/make_bp5.py:95);
    prestress="hbi": tau = sigma (a asinh(.) + eta_MPa v), the convention of
    HBI's generator (hbi_tests/bp5r_param_gen.py:51-53), which the
    reference's published BP5 recurrences (234.2937 yr 5dp and 234.2722 yr
    3bs at 1 km, 236.81 yr at 2 km) were computed with.
    """
    x, dep = g["along"], g["depth"]
    r = np.maximum(np.abs(dep - 10) - 6, np.abs(x) - 30) / 2.0
    a = np.clip(a0 + r * (a_max - a0), a0, a_max)
    b = np.full_like(a, b0)
    nuc = (np.abs(x + 24) < 6) & (np.abs(dep - 10) < 6)
    v = np.where(nuc, vnuc, vpl)
    dc = np.where(nuc, dc_nuc, dc0)
    eta = shear_modulus / (2.0 * cs)
    tau = sigma * a * np.arcsinh(0.5 * v / v0 * np.exp((f0 + b0 * np.log(v0 / vpl)) / a)) + eta * v
    if prestress == "hbi":
        # eta in MPa s/m multiplied by the normal stress in MPa
        tau = tau + (sigma / 1e6 - 1.0) * eta * v
    return dict(a=a, b=b, dc=dc, tau=tau, vel=v, sigma=np.full_like(a, sigma),
                f0=f0, v0=v0, vpl=vpl, eta=eta, shear_modulus=shear_modulus, cs=cs)


def cycle_fields(g, seed=20261202, a_vw=0.004, b0=0.03, amb_excess=0.04,
                 taper_km=2.0, dc0=0.14, dc_nuc=0.13, f0=0.6, v0=1e-6,
                 vpl=1e-9, sigma=25e6, shear_modulus=32e9, cs=3464.0,
                 perturb=1e-3, box=(30.0, 10.0, 6.0), nuc=(-24.0, 10.0, 6.0)):
    """Friction fields of the planar benchmark cycles.

    a = a_vw and b = b0 inside the central seismogenic box (|along| <= 30 km,
    |depth - 10| <= 6 km, i.e. 60 x 12 km); outside, a - b ramps linearly over
    taper_km to +amb_excess with b held at b0 (so a -> b0 + amb_excess);
    D_c = dc0 with dc_nuc in a 12 x 12 km nucleation square (the BP5 one);
    initial slip rate vpl (1 + perturb N(0,1)) from a seeded PCG64 stream,
    state at steady state for that rate, and the quasi-dynamic prestress
    tau = sigma a asinh(v/(2 v0) e^(psi_ss/a)) + eta v.
    """
    x, dep = g["along"], g["depth"]
    half_len, cdep, half_dep = box
    r = np.maximum(np.abs(dep - cdep) - half_dep, np.abs(x) - half_len) / taper_km
    amb_in = a_vw - b0
    amb = np.clip(amb_in + r * (amb_excess - amb_in), amb_in, amb_excess)
    b = np.full_like(x, b0)
    a = amb + b
    nx0, nd0, nh = nuc
    nucm = (np.abs(x - nx0) < nh) & (np.abs(dep - nd0) < nh)
    dc = np.where(nucm, dc_nuc, dc0)
    rng = np.random.Generator(np.random.PCG64(seed))
    v = vpl * (1.0 + perturb * rng.standard_normal(len(x)))
    eta = shear_modulus / (2.0 * cs)
    psi_ss = f0 + b * np.log(v0 / v)
    tau = sigma * a * np.arcsinh(0.5 * v / v0 * np.exp(psi_ss / a)) + eta * v
    return dict(a=a, b=b, dc=dc, tau=tau, vel=v, sigma=np.full_like(x, sigma),
                f0=f0, v0=v0, vpl=vpl, eta=eta, shear_modulus=shear_modulus, cs=cs)


def ensemble_fields(g, length_km, depth_km, member, geom_index=0, seed=20261202):
    """Friction fields of one ensemble member on a planar outline of
    length_km x depth_km: the benchmark fields with the seismogenic box,
    taper and nucleation square scaled to the outline (box |along| <=
    0.3 L, depth 0.25 W +- 0.15 W; nucleation square of half-size 0.06 L at
    along = -0.24 L) and three seeded draws per member: a - b in the box
    uniform in [-0.03, -0.02] (a = 0.004), D_c log-uniform in [0.05, 0.3] m
    (0.93 D_c in the nucleation square), sigma_n uniform in [20, 40] MPa.
    The draws come from a PCG64 stream keyed by (seed, geometry, member)."""
    rng = np.random.Generator(np.random.PCG64([seed, geom_index, member]))
    amb = rng.uniform(-0.03, -0.02)
    dc = float(np.exp(rng.uniform(np.log(0.05), np.log(0.3))))
    sigma = rng.uniform(20e6, 40e6)
    L, W = float(length_km), float(depth_km)
    f = cycle_fields(g, seed=seed + 1000 * geom_index + member, a_vw=0.004, b0=0.004 - amb,
                     dc0=dc, dc_nuc=0.93 * dc, sigma=sigma,
                     box=(0.3 * L, 0.25 * W, 0.15 * W), nuc=(-0.24 * L, 0.25 * W, 0.06 * L))
    f.update(amb=amb, dc_draw=dc, sigma_draw=sigma)
    return f
