"""Write the golden vectors of tests/golden from the reference's own code.

Runs only in a container that has the reference sources (default
/root/reference) and the tools of oracle/Makefile.ref built from them
(oracle/_ref/ref_kgen, oracle/_ref/makefault).  Nothing here is needed at
test time: the outputs are committed as data.

    make -f oracle/Makefile.ref REF=/root/reference
    python tests/golden/make_golden.py [--ref /root/reference]

Outputs
  geom_fvh.in.gz        makefault -n 40 -m 25 twice (second at x=1, group 1):
                        the 2000-patch mesh of scaling_tests/full_vs_half_test.sh
  ktensor_grid.npz      reference tractions T[i,j,slip,trac] (internal units)
                        on a 12 x 5 planar grid with 1 km cells, half space
  ktensor_random.npz    same on 40 random rectangles (strike, dip, size),
                        half and full space
  ktensor_tri.npz       reference tractions on triangle meshes (Nikkhoo-Walter
                        sources, TDCS receiver frames): planar split grid,
                        rough (Hurst 0.8) grid, random triangles incl. a
                        surface-touching, a horizontal and a vertical-edge
                        one, and a mixed rectangle + triangle mesh (half and
                        full space)
  tri_vertical_edge.npz reference tractions among six triangles of the
                        32000-triangle rough mesh (0.5 km cells, Hurst 0.8
                        offset): three receivers near the surface or the
                        ends, each with a far source one of whose edges lies
                        within 1e-5 rad of vertical
  dc3d_cases.npz        direct DC3D calls: random points, points on edges,
                        on negative edge extensions, at the surface, and
                        dips 0 / 90 / 89.9999995 (cos snap) / 45
  bp5_2km/              the reference's BP5 2 km input files (data)
  ref_results/          the reference's published result files used as
                        fixtures (events, slider physics, |K 1| answers, the
                        head of a dense monitor file)
"""
import argparse
import gzip
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFBIN = os.path.join(ROOT, "oracle", "_ref")


def run(args, stdin=None):
    return subprocess.run(args, input=stdin, capture_output=True, text=True,
                          check=True).stdout


def write_geom(fname, xyz, strike, dip, l, w, grp=None):
    with open(fname, "w") as f:
        for k in range(len(l)):
            g = 0 if grp is None else grp[k]
            f.write("%.12e %.12e %.12e %.6f %.6f %.12e %.12e %d\n" %
                    (xyz[k][0], xyz[k][1], xyz[k][2], strike[k], dip[k], l[k],
                     w[k], g))


def kmat(geom_file, full):
    n = sum(1 for line in open(geom_file) if line.strip())
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as tf:
        out = tf.name
    try:
        run([os.path.join(REFBIN, "ref_kgen"), "kmat", geom_file, str(full), out])
        raw = np.fromfile(out, dtype=np.uint8)
    finally:
        os.unlink(out)
    T = raw[:n * n * 48].view(np.float64).reshape(n, n, 2, 3)
    ir = raw[n * n * 48:].view(np.int32).reshape(n, n, 2)
    return T, ir


def write_mesh(fname, g):
    """rectangles as write_geom; triangles (l, w < 0) with their nine
    vertex coordinates appended (src/read_geometry.c:148-175)"""
    with open(fname, "w") as f:
        for k in range(len(g["l"])):
            x, y, z = g["xyz"][k]
            line = "%.12e %.12e %.12e %.6f %.6f %.12e %.12e 0" % (
                x, y, z, g["strike"][k], g["dip"][k], g["l"][k], g["w"][k])
            if g["l"][k] < 0 and g["w"][k] < 0:
                line += " " + " ".join("%.17e" % v for v in g["vert"][k])
            f.write(line + "\n")


def triangles(tmp):
    sys.path.insert(0, ROOT)
    from interact_amd import faults
    rng = np.random.default_rng(4242)
    sets = {}
    sets["planar"] = faults.planar_tri_grid(2.0, 12.0, 6.0)
    sets["rough"] = faults.planar_tri_grid(
        1.0, 8.0, 5.0, offset=faults.rough_offset(8, 5, 1.0, alpha=2e-2, seed=7))
    m = 30
    c = np.stack([rng.uniform(-4e3, 4e3, m), rng.uniform(-4e3, 4e3, m),
                  rng.uniform(-6e3, -1.2e3, m)], 1)
    vert = np.repeat(c, 3, axis=0).reshape(m, 9) + rng.uniform(-900, 900, (m, 9))
    vert[:, 2::3] = np.minimum(vert[:, 2::3], -50.0)
    vert[0] = [0, 0, 0, 800, 0, 0, 300, 0, -900]           # touches the surface
    vert[1] = [1500, 1500, -2000, 2500, 1500, -2000, 1500, 2600, -2000]   # horizontal
    vert[2] = [-2500, 0, -1000, -2500, 0, -2200, -1700, 600, -1500]       # vertical edge
    sets["random"] = dict(xyz=(vert[:, 0:3] + vert[:, 3:6] + vert[:, 6:9]) / 3.0,
                          strike=np.zeros(m), dip=np.full(m, 90.0), l=np.full(m, -1.0),
                          w=np.full(m, -1.0), vert=vert)
    mr, mt = 10, 10
    xyz = np.stack([rng.uniform(-4e3, 4e3, mr), rng.uniform(-4e3, 4e3, mr),
                    rng.uniform(-6e3, -1.5e3, mr)], 1)
    mixed = dict(xyz=np.concatenate([xyz, sets["random"]["xyz"][3:3 + mt]]),
                 strike=np.concatenate([np.round(rng.uniform(0, 360, mr), 3), np.zeros(mt)]),
                 dip=np.concatenate([np.round(rng.uniform(5, 90, mr), 3), np.full(mt, 90.0)]),
                 l=np.concatenate([rng.uniform(150, 700, mr), np.full(mt, -1.0)]),
                 w=np.concatenate([rng.uniform(150, 500, mr), np.full(mt, -1.0)]),
                 vert=np.concatenate([np.zeros((mr, 9)), sets["random"]["vert"][3:3 + mt]]))
    sets["mixed"] = mixed
    out = {}
    for name, g in sets.items():
        gfile = os.path.join(tmp, name + ".in")
        write_mesh(gfile, g)
        T, ir = kmat(gfile, 0)
        out[name + "_geom"] = open(gfile).read()
        out[name + "_T"], out[name + "_iret"] = T, ir
        if name == "mixed":
            Tf, irf = kmat(gfile, 1)
            out[name + "_Tfull"], out[name + "_iretfull"] = Tf, irf
    np.savez_compressed(os.path.join(HERE, "ktensor_tri.npz"), **out)


def vertical_edges(tmp):
    """the reference's tractions among patches 0, 154, 31952, 31709, 3994
    and 3761 of the rough 0.5 km triangle mesh; sources 154, 31709 and 3761
    have an edge 6e-6, 1e-5 and 1e-5 rad from vertical"""
    sys.path.insert(0, ROOT)
    from interact_amd import faults
    g = faults.planar_tri_grid(0.5, offset=faults.rough_offset(200, 80, 0.5))
    ids = np.array([0, 154, 31952, 31709, 3994, 3761])
    vert = np.asarray(g["vert"]).reshape(-1, 9)[ids]
    m = len(ids)
    sub = dict(xyz=(vert[:, 0:3] + vert[:, 3:6] + vert[:, 6:9]) / 3.0, strike=np.zeros(m),
               dip=np.full(m, 90.0), l=np.full(m, -1.0), w=np.full(m, -1.0), vert=vert)
    gfile = os.path.join(tmp, "vertical.in")
    write_mesh(gfile, sub)
    T, ir = kmat(gfile, 0)
    np.savez_compressed(os.path.join(HERE, "tri_vertical_edge.npz"), ids=ids, vert=vert, T=T,
                        iret=ir, geom=open(gfile).read())


def monitor_excerpt(ref, rr, nrows=40):
    """the header and first nrows rows of a dense rsf_monitor.dat written by
    the reference (test_hmatrix/out_2f_2.0km_sep4.0km_260yr_np4): fixture
    of the monitor-file writer's column layout"""
    src = os.path.join(ref, "test_hmatrix", "out_2f_2.0km_sep4.0km_260yr_np4", "dense",
                       "rsf_monitor.dat")
    with open(src) as f:
        lines = f.readlines()
    head = [l for l in lines if l.startswith("#")][:1]
    data = [l for l in lines if not l.startswith("#")][:nrows]
    with open(os.path.join(rr, "monitor_2f_2km_dense_head.dat"), "w") as f:
        f.writelines(head + data)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--only", choices=["tri", "vertical"], default=None,
                    help="write only one group of fixtures")
    args = ap.parse_args()
    ref = args.ref
    if not os.path.exists(os.path.join(REFBIN, "ref_kgen")):
        sys.exit("build oracle/_ref first: make -f oracle/Makefile.ref")
    tmp = tempfile.mkdtemp()
    if args.only == "tri":
        triangles(tmp)
        shutil.rmtree(tmp)
        print("triangle fixtures written to", HERE)
        return
    if args.only == "vertical":
        vertical_edges(tmp)
        shutil.rmtree(tmp)
        print("vertical-edge fixture written to", HERE)
        return

    # 1. the full_vs_half mesh (scaling_tests/full_vs_half_test.sh:85-86)
    mf = os.path.join(REFBIN, "makefault")
    txt = run([mf, "-n", "40", "-m", "25"]) + run([mf, "-n", "40", "-m", "25",
                                                  "-x", "1", "-grp", "1"])
    with gzip.open(os.path.join(HERE, "geom_fvh.in.gz"), "wt") as f:
        f.write(txt)

    # 2. planar grid, 12 x 5 cells of 1 km (BP5 layout)
    nx, nz, ds = 12, 5, 1000.0
    ii, jj = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    n = nx * nz
    xyz = np.stack([np.zeros(n), (ii + 0.5 - nx / 2) * ds, -(jj + 0.5) * ds], 1)
    gfile = os.path.join(tmp, "grid.in")
    write_geom(gfile, xyz, np.zeros(n), np.full(n, 90.0), np.full(n, ds / 2),
               np.full(n, ds / 2))
    T, ir = kmat(gfile, 0)
    np.savez_compressed(os.path.join(HERE, "ktensor_grid.npz"),
                        geom=open(gfile).read(), T=T, iret=ir)

    # 3. random rectangles (strike, dip, size), half and full space
    rng = np.random.default_rng(20261202)
    m = 40
    xyz = np.stack([rng.uniform(-4e3, 4e3, m), rng.uniform(-4e3, 4e3, m),
                    rng.uniform(-7e3, -1.5e3, m)], 1)
    strike = np.round(rng.uniform(0, 360, m), 3)
    dip = np.round(rng.uniform(5, 90, m), 3)
    l = rng.uniform(150, 700, m)
    w = rng.uniform(150, 500, m)
    gfile = os.path.join(tmp, "rand.in")
    write_geom(gfile, xyz, strike, dip, l, w)
    Th, irh = kmat(gfile, 0)
    Tf, irf = kmat(gfile, 1)
    np.savez_compressed(os.path.join(HERE, "ktensor_random.npz"),
                        geom=open(gfile).read(), T_half=Th, iret_half=irh,
                        T_full=Tf, iret_full=irf)

    # 4. direct DC3D calls incl. special cases
    cases = []
    alpha = 2.0 / 3.0
    for k in range(200):   # generic points
        cases.append([alpha, rng.uniform(-3, 3), rng.uniform(-3, 3),
                      -rng.uniform(0, 4), rng.uniform(0.5, 4),
                      rng.choice([0.0, 30.0, 45.0, 70.0, 90.0, 89.9999995]),
                      -1.0, 1.0, -0.7, 0.7, 1.0, 0.0, 0.0, int(k % 3 == 0)])
    for dip in (0.0, 45.0, 90.0, 89.9999995):
        for d1, d2 in ((1.0, 0.0), (0.0, 1.0), (0.7, -0.3)):
            cd, sd = np.cos(np.radians(dip)), np.sin(np.radians(dip))
            depth = 2.0
            for (x, ep) in ((1.0, 0.0), (-1.0, 0.3), (1.0, 0.7), (0.2, -0.7),
                            (-1.5, 0.7), (-1.5, 0.0), (0.3, 1.2), (3.0, -0.7)):
                # point in the fault plane at along-strike x, along-dip ep
                y = ep * cd
                d = depth - ep * sd
                cases.append([alpha, x, y, -d, depth, dip, -1.0, 1.0, -0.7,
                              0.7, d1, d2, 0.0, 0])
            cases.append([alpha, 0.5, -0.4, 0.0, depth, dip, -1.0, 1.0, -0.7,
                          0.7, d1, d2, 0.0, 0])   # free surface
            cases.append([alpha, 0.3, 0.2, -1.0, 0.6, dip, -1.0, 1.0, -0.6,
                          0.6, d1, d2, 0.0, 0])   # surface-breaking source
    cases = np.array(cases, dtype=np.float64)
    lines = "\n".join(" ".join("%.17g" % v for v in c) for c in cases) + "\n"
    out = run([os.path.join(REFBIN, "ref_kgen"), "dc3d"], stdin=lines)
    res = np.array([[float(v) for v in line.split()] for line in out.strip().split("\n")])
    np.savez_compressed(os.path.join(HERE, "dc3d_cases.npz"), args=cases,
                        u=res[:, :12], iret=res[:, 12].astype(np.int32))

    # 5. reference data files (inputs and published results)
    bp = os.path.join(HERE, "bp5_2km")
    os.makedirs(bp, exist_ok=True)
    src = os.path.join(ref, "seas_bp_tests", "bp5")
    for f in ("geom_bp5_2km.in", "rsf_bp5_2km.dat", "ic_bp5_2km.in",
              "dc_bp5_2km.in"):
        shutil.copyfile(os.path.join(src, f), os.path.join(bp, f))
    rr = os.path.join(HERE, "ref_results")
    os.makedirs(rr, exist_ok=True)
    shutil.copyfile(os.path.join(src, "ntest.0", "rsf_events.1km.0.inf.dat"),
                    os.path.join(rr, "bp5_1km_dense_3bs_rtol1e-4_events.dat"))
    shutil.copyfile(os.path.join(ref, "slider", "slider_physics.dat"),
                    os.path.join(rr, "slider_physics.dat"))
    shutil.copyfile(os.path.join(ref, "slider", "geom_slider10.in"),
                    os.path.join(rr, "geom_slider10.in"))
    shutil.copyfile(os.path.join(ref, "scaling_tests", "full_vs_half.csv"),
                    os.path.join(rr, "full_vs_half.csv"))
    monitor_excerpt(ref, rr)
    # 6. triangle meshes
    triangles(tmp)
    vertical_edges(tmp)
    shutil.rmtree(tmp)
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
