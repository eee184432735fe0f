"""Benchmark of the dense rsf_solve hot path on one MI355X.

Workload (default, the large dense configuration): vertical strike-slip
fault 100 x 40 km at 0.5 km patches = 16000 patches x 2 slip components =
32000 DOF, dense K of 32000 x 32000 (8.2 GB in FP64); aging-law
rate-and-state friction with radiation damping; seeded initial slip rate
vpl (1 + 1e-3 N(0,1)); rk5dp with relative tolerance 1e-8 over 200 yr.
--config c2 runs the same fault at 1 km patches (4000 patches, 8000 DOF)
over 400 yr.  The headline cycle is timed as --steps continued segments of
equal length (rsfg_solve_continue, whose accepted steps are exactly those
of one solve over the whole window): one bench "step" is one segment, and
the timed steps together are always the whole cycle; --warmup segments of
the same length run first from the initial state, untimed.

Reported: simulated years per wall-clock hour of the mixed-precision path
as `value` (K as bf16 hi for every entry plus the bf16 lo of each patch's
near-field column chunks, FP32 products summed per 16-byte slice and
accumulated in FP64, FP64 state and error control); the FP64 path, FP32 and
bf16 hi+lo storage (and bf16 hi alone on c2) beside it with the error of
each against the FP64 path at matched events (first-recurrence time, peak
slip rate) and in K v; the roofline of the headline fused stage kernel
(K v + RHS epilogue) measured with HIP events on its stream; operator
assembly throughput; and the CPU oracle timed on the host cores on a
bounded sample.

--config c5 runs the sensitivity workload: on the 1 km fault (8000 DOF), the
gradient of the log slip-rate misfit at 64 seeded receivers over a 50 yr
interseismic window (after the first event) with respect to a - b and D_c of
every patch (discrete adjoint of the accepted rk5dp steps, FP64 operator,
checkpoints every 256 steps), checked against central differences on 32
random directions and the oracle's FP64 reference adjoint on the window's
first steps.  --config c4 runs the batched ensemble instead: four planar outlines (40 x 20,
60 x 24, 80 x 32 and 100 x 40 km at 1 km patches, 1600 to 8000 DOF) with 32
seeded friction draws each, 100 yr per member, all 128 members in one
ensemble (K [v_1 .. v_32] per geometry on the matrix cores, FP32 operators in MFMA tiles),
against the same 128 members run one after another through the single-fault
solver; then the same ensemble on the two-byte operator (bf16 hi tiles plus
near-field lo tiles on bf16 MFMAs) against single runs on the bf16nf operator
(--no-two-byte skips that leg).

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--config c3|c2|c4|c5]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SEC_PER_YEAR = 31557600.0
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the stage kernel on the 1 km (8000 DOF: per-patch
# kernel) and 0.5 km (32000 DOF: persistent kernel, FP64 per-patch)
# workloads: FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors that a plain
# 256 MiB device copy reads on the same counters
# (profiles/pmc_traffic_planar1km_8000dof.md,
# profiles/kernels_planar0.5km_32000dof.md)
STAGE_TRAFFIC = {"c2": {"fp64": 517.6e6, "fp32": 262.1e6, "bf16x2": 262.2e6, "bf16": 132.6e6,
                        "bf16nf": 145.9e6},
                 "c3": {"fp64": 8.2031e9, "fp32": 4.1035e9, "bf16x2": 4.1035e9,
                        "bf16": 2.0694e9, "bf16nf": 2.1443e9}}
# HBM bytes per launch of the FP32 product k_matvec_p<float,2> (same counters)
MATVEC_TRAFFIC = {"c2": {"fp32": 260.1e6}, "c3": {"fp32": 4.0965e9}}
# HBM bytes per launch of the c4 batched product by unit shape (the shapes
# the tuner picks on these boxes), from the same two counter runs of
# `--config c4 --kernels-only` with the copy-kernel factors above
# (profiles/k_prod_b_ensemble_c4.md)
PRODUCT_TRAFFIC_C4 = {(32, 4, 1, 0, 8, 1, 1): 449.9e6, (32, 4, 1, 0, 4, 1, 2): 459.4e6,
                      (64, 4, 1, 1024, 2, 1, 0): 462.8e6, (32, 8, 1, 1024, 2, 1, 0): 462.5e6,
                      (32, 8, 1, 0, 4, 1, 1): 464.8e6, (64, 4, 1, 1024, 4, 1, 0): 465.1e6,
                      (32, 8, 1, 0, 8, 1, 1): 468.0e6, (32, 8, 2, 0, 4, 1, 1): 473.7e6}
# (the two deeper-ring shapes: profiles/k_prod_sk_h_ensemble_c4.md); the
# same for the two-byte product (FP32 partials) by stream-K shape
# (profiles/k_prod_sk_h_ensemble_c4.md)
PRODUCT_TRAFFIC_C4_H = {(64, 4, 1, 0, 4, 1, 1): 277.0e6, (64, 4, 1, 0, 4, 1, 2): 277.0e6,
                        (64, 4, 1, 0, 2, 1, 2): 278.5e6, (32, 8, 1, 0, 4, 1, 1): 271.8e6,
                        (32, 4, 1, 0, 8, 1, 1): 265.2e6, (32, 4, 1, 0, 4, 1, 2): 268.4e6,
                        (32, 8, 1, 0, 8, 1, 1): 275.1e6, (32, 8, 2, 0, 4, 1, 1): 281.0e6}
# HBM bytes per launch of the backward sweep's K^T w at c5 (FP64 operator),
# from the same two counter runs of `--config c5 --kernels-only`
# (profiles/adjoint_planar1km_8000dof.md)
ADJ_KT_TRAFFIC_C5 = 521.2e6
# the mixed path reported as value: bf16 hi for every entry plus the bf16 lo
# of each patch's near-field column chunks (compensated where it matters)
HEADLINE = "bf16nf"
STAGE_TYPES = {"fp64": "double", "fp32": "float", "bf16x2": "bf16x2_t", "bf16": "bf16_t",
               "bf16nf": "bf16nf_t"}
DTYPES = {"fp64": "f64 K.v + f64 state",
          "fp32": "f32 K (f32 slice products, f64 accumulation) + f64 state",
          "bf16x2": "bf16 hi+lo K (f32 slice products, f64 accumulation) + f64 state",
          "bf16": "bf16 K (f32 slice products, f64 accumulation) + f64 state",
          "bf16nf": "bf16 hi K + bf16 lo in near-field chunks (f32 slice products, f64 "
                    "accumulation) + f64 state"}
# algorithmic FP64 operations per (receiver, source) pair of the Okada path
# (two DC3D evaluations + stress + tractions, half space, vertical coplanar
# patches), counted on the oracle by oracle/opcount.cpp
F_PAIR = 6205
# the same for a Nikkhoo-Walter triangle pair (two half-space triangle
# stresses for strike and dip slip, each the main and image dislocations and
# the free-surface correction of three edges, plus tractions), counted by
# oracle/opcount_tri.cpp on receivers off the source triangle of a rough
# fault; it includes the per-triangle frame and edge angles that the oracle,
# like the reference, evaluates within every pair
F_PAIR_TRI = 30559
FP64_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3
METRIC = "sim-years/hour + K·v HBM GB/s, mixed-prec vs FP64 GPU at matched event times"


# the persistent kernels (x copied into LDS, one workgroup per CU; x longer
# than the 144 KiB LDS copy in column blocks) run for the FP32-input formats;
# the stage kernel only from 3 patches per wave up (more than 4096 patches
# on the 256 CUs of an MI355X), below which the per-patch kernel is faster
PERSIST_MIN_PATCHES = 2 * 256 * 8 + 1


def persistent(name, D):
    # every FP32-input format (x longer than the LDS copy runs in column blocks)
    return name != "fp64"


def stage_kernel(name, D, n):
    if persistent(name, D) and n >= PERSIST_MIN_PATCHES:
        return f"k_stage_p<{STAGE_TYPES[name]},2,2> (persistent: x in LDS, fused K v + RHS epilogue)"
    return f"k_stage<{STAGE_TYPES[name]},2,2> (fused K v + RHS epilogue)"


def matvec_kernel(name, D):
    if persistent(name, D):
        return f"k_matvec_p<{STAGE_TYPES[name]},2> (persistent: x in LDS)"
    return f"k_matvec<{STAGE_TYPES[name]},2>"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def first_onset(ev, t_min_yr=0.0):
    on = ev[ev[:, 1] > 0, 0] / SEC_PER_YEAR
    on = on[on > t_min_yr]
    return float(on[0]) if len(on) else float("nan")


def peak_after(ev, t_yr):
    """peak |v| of the event whose onset is at t_yr (from its arrest row)."""
    rows = ev[(ev[:, 0] / SEC_PER_YEAR >= t_yr - 1e-9) & (ev[:, 1] < 0)]
    return float(rows[0, 5]) if len(rows) else float("nan")


def host_cpu():
    """CPU model, sockets, physical cores and the cores this process may use
    (lscpu, else /proc/cpuinfo)."""
    import subprocess
    info = dict(model="unknown", sockets=None, physical_cores=None,
                threads_per_core=None, affinity_cpus=len(os.sched_getaffinity(0)),
                nproc=os.cpu_count())
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name", info["model"])
        sockets = int(kv.get("Socket(s)", "0") or 0)
        cps = int(kv.get("Core(s) per socket", "0") or 0)
        tpc = int(kv.get("Thread(s) per core", "0") or 0)
        if sockets and cps:
            info.update(sockets=sockets, physical_cores=sockets * cps,
                        threads_per_core=tpc or None)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    if info["model"] == "unknown":
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    info["model"] = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    return info


def cpu_baseline(g, f, K, rhs_evals_ref, years, budget_s, log_fn):
    """Time the FP64 CPU oracle on the same workload (same operator values,
    same state): dense K v + RHS evaluations for about budget_s seconds at
    one thread and at all physical cores this process may use, each scaled
    by the RHS count of the FP64 GPU run to simulated years/hour; assembly
    throughput timed on a 400-patch sample of the mesh at both counts."""
    import oracle
    from interact_amd.model import initial_state
    cpu = host_cpu()
    allowed = cpu["affinity_cpus"]
    # all physical cores, as far as the affinity mask lets this process use them
    full = max(1, min(cpu["physical_cores"] or allowed, allowed))
    n = len(g["l"])
    D = K.shape[0]
    rng = np.random.default_rng(0)
    sub = 400
    idx = np.sort(rng.choice(n, sub, replace=False))
    gs = {k: g[k][idx] for k in ("xyz", "strike", "dip", "l", "w")}
    xpl = np.zeros(D)
    xpl[0::2] = -f["vpl"]
    X = initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=2)
    res = {}
    for threads in (1, full):
        t0 = time.perf_counter()
        oracle.traction_tensor(gs, 0, threads)
        t_pairs = (time.perf_counter() - t0) / (sub * sub)
        P = oracle.RSFProblem(K, f["a"], f["b"], f["dc"], K @ xpl, ncomp=2, eta=f["eta"],
                              v0=f["v0"], f0=f["f0"], nthreads=threads)
        P.rhs(X)
        nev, t0 = 0, time.perf_counter()
        while nev < 2 or time.perf_counter() - t0 < budget_s / 2:
            P.rhs(X)
            nev += 1
        t_rhs = (time.perf_counter() - t0) / nev
        res[threads] = (t_rhs, t_pairs, nev)
        log_fn(f"cpu oracle, {threads} threads: {nev} RHS -> {t_rhs*1e3:.2f} ms/RHS, "
               f"{t_pairs*1e9:.0f} ns/pair")
        del P
    t_rhs, t_pairs, nev = res[full]
    t1, p1, n1 = res[1]
    rate = lambda t: years / (rhs_evals_ref * t / 3600.0)
    return dict(value=rate(t_rhs), unit="sim-years/hour", cores=full, kind="port",
                value_1thread=rate(t1), cores_1thread=1,
                sample=(f"FP64 right-hand sides (dense {D}x{D} K v + RHS, OpenMP over rows) "
                        f"on this workload's state: {nev} at {full} threads ({t_rhs*1e3:.2f} ms "
                        f"each) and {n1} at 1 thread ({t1*1e3:.1f} ms each), scaled by the "
                        f"{rhs_evals_ref} RHS evaluations the FP64 GPU run needed for "
                        f"{years:g} yr; assembly on a {sub}x{sub} sample: {t_pairs*1e9:.0f} "
                        f"ns/pair at {full} threads, {p1*1e9:.0f} ns/pair at 1"),
                rhs_ms=t_rhs * 1e3, rhs_ms_1thread=t1 * 1e3,
                assembly_pairs_per_s=1.0 / t_pairs, assembly_pairs_per_s_1thread=1.0 / p1,
                host=cpu)


def kernels_only(fs, store, rough_ds=None):
    """Launch each dominant kernel a few times and one plain device copy of
    256 MiB (read 256 MiB, write 256 MiB) that calibrates the FETCH_SIZE /
    WRITE_SIZE counters.  The order is fixed: assembly (already done by the
    caller), then per precision 5 stage and 5 matvec launches, then the copy,
    then (rough_ds given) two assemblies of the rough triangle mesh."""
    import ctypes as C
    from interact_amd import lib
    for name in ("fp64", "fp32", "bf16x2", "bf16", "bf16nf"):
        p = lib.PREC_BY_NAME[name]
        if p != lib.PREC_FP64 and not (store & p):
            continue
        fs.stage_bench(p, 5)
        fs.matvec_bench(p, 5)
    # device-to-device copy through the HIP runtime librsfg already loaded
    # (the runtime's copy kernel: reads n bytes, writes n bytes)
    hip = C.CDLL("libamdhip64.so")
    n = 256 << 20
    src, dst = C.c_void_p(), C.c_void_p()
    ok = hip.hipMalloc(C.byref(src), C.c_size_t(n)) == 0
    ok = ok and hip.hipMalloc(C.byref(dst), C.c_size_t(n)) == 0
    ok = ok and hip.hipMemcpy(dst, src, C.c_size_t(n), 3) == 0   # hipMemcpyDeviceToDevice
    ok = ok and hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    if not ok:
        raise SystemExit("kernels-only: device copy failed")
    log(f"kernels-only: copy of {n} bytes done")
    if rough_ds is not None:
        fs.close()
        rough_triangle_assembly(rough_ds, products=False)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="c2/c3: timed segments of the headline cycle (together always the "
                         "whole cycle; default 20); c4/c5: timed repetitions (default 1)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="c2/c3: untimed segments of the same length from the initial state "
                         "before the timed cycle (default 5); c4/c5: untimed repetitions "
                         "(default 0)")
    ap.add_argument("--config", choices=["c2", "c3", "c4", "c5"], default="c3",
                    help="c3: 0.5 km patches, 32000 DOF, 200 yr (default); c2: 1 km patches, "
                         "8000 DOF, 400 yr; c4: ensemble of 4 geometries x 32 friction "
                         "draws, 100 yr each; c5: adjoint gradient on the 1 km fault")
    ap.add_argument("--no-oracle-c1", action="store_true",
                    help="skip the c1 cycle (2000 DOF, 300 yr) on the device against the "
                         "FP64 oracle's cycle")
    ap.add_argument("--years", type=float, default=None)
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--fd-directions", type=int, default=32,
                    help="c5: random directions of the central-difference check")
    ap.add_argument("--no-two-byte", action="store_true",
                    help="c4: skip the ensemble on the two-byte operator (bf16 hi + near-field "
                         "lo tiles on bf16 MFMAs) and its single runs on the bf16nf operator")
    ap.add_argument("--singles", type=int, default=None,
                    help="c4: members also run one after another through the single-fault "
                         "solver (default all 128)")
    ap.add_argument("--rough-cycle-years", type=float, default=0.0,
                    help="c3: also run cycles of this length on the rough triangle mesh "
                         "(64000 DOF) in FP64, FP32 and bf16nf (default 0: products only)")
    ap.add_argument("--rough-only", action="store_true",
                    help="c3: only the rough triangle mesh (assembly, products, and the cycles "
                         "of --rough-cycle-years); prints its JSON object")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precisions", default=None,
                    help="cycle runs, comma separated (fp64 first is the reference); "
                         "default fp64,bf16nf,fp32,bf16x2 (c2 adds bf16)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="one assembly, a few launches of the stage and matvec kernels "
                         "per precision and a 256 MiB device copy of known size, for "
                         "hardware-counter runs; no cycle, no JSON bench line (c4: the "
                         "tuner's launches of every tile shape; c5: one gradient over 64 "
                         "fixed steps and 20 K^T w launches)")
    args = ap.parse_args()
    cyc = args.config in ("c2", "c3")
    if args.steps is None:
        args.steps = 20 if cyc else 1
    if args.warmup is None:
        args.warmup = 5 if cyc else 0
    if args.gpus != 1:
        print(json.dumps({"error": "runs on one GPU only: the row-block sharding over "
                          "several GPUs is not built", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.config == "c4":
        if args.kernels_only:
            ensemble_kernels_only(args)
            return
        print(json.dumps(ensemble_bench(args)))
        return
    if args.config == "c5":
        if args.kernels_only:
            adjoint_kernels_only(args)
            return
        print(json.dumps(adjoint_bench(args)))
        return
    args.ds = {"c2": 1.0, "c3": 0.5}[args.config]
    if args.rough_only:
        if args.config != "c3":
            raise SystemExit("--rough-only belongs to c3")
        print(json.dumps(rough_triangle_assembly(args.ds, args.rough_cycle_years, args.rtol)))
        return
    if args.years is None:
        args.years = {"c2": 400.0, "c3": 200.0}[args.config]
    if args.precisions is None:
        args.precisions = ("fp64,bf16nf,fp32,bf16x2,bf16" if args.config == "c2"
                           else "fp64,bf16nf,fp32,bf16x2")
    if args.steps < 1 or args.warmup < 0:
        raise SystemExit("--steps must be >= 1 and --warmup >= 0")

    from interact_amd import faults, lib
    from interact_amd.model import FaultSystem, initial_state

    g = faults.planar_grid(args.ds)
    f = faults.cycle_fields(g)
    n = len(g["l"])
    D = 2 * n
    store = lib.PREC_FP32 | lib.PREC_BF16X2 | lib.PREC_BF16 | lib.PREC_BF16NF
    fs = FaultSystem(g)
    asm_ms = fs.assemble(lib.SLIP_VECTOR, shear_modulus=f["shear_modulus"], store=store)
    asm_ms = fs.assemble(lib.SLIP_VECTOR, shear_modulus=f["shear_modulus"], store=store)
    pairs = n * n
    log(f"assembly: {n} patches, {pairs} pairs in {asm_ms:.3f} ms = {pairs/asm_ms*1e3:.3e} pairs/s")
    fs.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"])
    fs.set_loading(f["vpl"])
    X0 = initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=2)
    fs.set_state(X0)
    if args.kernels_only:
        kernels_only(fs, store, args.ds if args.config == "c3" else None)
        return

    # K v accuracy against the FP64 operator on x_i = sin(i+1) + 0.3 cos(3i)
    i = np.arange(D)
    xa = np.sin(i + 1.0) + 0.3 * np.cos(3.0 * i)
    y64 = fs.matvec(xa, lib.PREC_FP64)
    kv = {}
    for name in ("fp64", "fp32", "bf16x2", "bf16", "bf16nf"):
        p = lib.PREC_BY_NAME[name]
        y = fs.matvec(xa, p)
        # median of three measurements of 50 launches each (HIP events)
        mv_ms = float(np.median([fs.matvec_bench(p, 50) for _ in range(3)]))
        st_ms = float(np.median([fs.stage_bench(p, 50) for _ in range(3)]))
        bytes_alg = fs.operator_bytes(p) + 16.0 * D
        kv[name] = dict(relerr=float(np.linalg.norm(y - y64) / np.linalg.norm(y64)),
                        matvec_ms=mv_ms, stage_ms=st_ms,
                        matvec_gbs=bytes_alg / mv_ms / 1e6, stage_gbs=bytes_alg / st_ms / 1e6,
                        bytes=bytes_alg)
        log(f"K.v {name}: relerr {kv[name]['relerr']:.2e}  matvec {mv_ms*1e3:.1f} us "
            f"{kv[name]['matvec_gbs']:.0f} GB/s  stage {st_ms*1e3:.1f} us {kv[name]['stage_gbs']:.0f} GB/s")
    # what a plain read of 4 GiB reaches on this device with the products'
    # load pattern and no arithmetic (rsfg_read_bench), in the same process
    rd_bytes = 4 << 30
    rd_ms = float(np.median([fs.read_bench(rd_bytes, 20) for _ in range(3)]))
    read_ceil = dict(gbs=rd_bytes / rd_ms / 1e6, bytes=rd_bytes, launch_us=rd_ms * 1e3,
                     frac=rd_bytes / rd_ms / 1e6 / HBM_PEAK_GBS,
                     kernel="k_read_stream (16-byte nontemporal loads, 8 in flight per lane, "
                            "2 workgroups of 8 waves per CU)")
    log(f"read stream: {rd_bytes} bytes in {rd_ms*1e3:.1f} us = {read_ceil['gbs']:.0f} GB/s "
        f"({read_ceil['frac']:.3f} of {HBM_PEAK_GBS:.0f})")

    t_end = args.years * SEC_PER_YEAR
    seg = t_end / args.steps
    opts = dict(rtol=args.rtol, steps_per_launch=64)
    runs = {}
    for name in args.precisions.split(","):
        p = lib.PREC_BY_NAME[name]
        if name == HEADLINE:
            # warmup: W segments from the initial state (graph capture, caches)
            fs.set_state(X0)
            for k in range(args.warmup):
                (fs.solve if k == 0 else fs.solve_continue)(min(k + 1, args.steps) * seg,
                                                              prec=p, **opts)
            # timed: the whole cycle as K continued segments (one step each);
            # their accepted steps are exactly those of one solve to t_end
            fs.set_state(X0)
            walls = []
            for k in range(args.steps):
                t0 = time.perf_counter()
                st = (fs.solve if k == 0 else fs.solve_continue)((k + 1) * seg, prec=p, **opts)
                walls.append(time.perf_counter() - t0)
            wall = float(np.sum(walls))
            ms_step = wall / args.steps * 1e3
        else:
            fs.set_state(X0)
            t0 = time.perf_counter()
            st = fs.solve(t_end, prec=p, **opts)
            wall = time.perf_counter() - t0
            ms_step = None
        ev = fs.events()
        t1 = first_onset(ev, 1.0)
        runs[name] = dict(wall_s=wall, steps=st["steps"], rejects=st["rejects"],
                          rhs_evals=st["rhs_evals"], events=int(len(ev)),
                          first_recurrence_yr=t1, peak_slip_rate=peak_after(ev, t1),
                          sim_years_per_hour=args.years / (wall / 3600.0))
        if ms_step is not None:
            runs[name].update(segments=args.steps, ms_per_segment=ms_step)
        log(f"cycle {name}: {args.years:g} yr in {wall:.2f} s "
            f"({runs[name]['sim_years_per_hour']:.3e} sim-yr/h), steps {st['steps']} "
            f"rejects {st['rejects']}, first event {t1:.4f} yr, peak v {runs[name]['peak_slip_rate']:.4f}")
    ref = runs.get("fp64")
    errors = {}
    for name, r in runs.items():
        e = dict(kv_relerr=kv[name]["relerr"])
        if ref is not None:
            e["first_recurrence_dt_yr"] = r["first_recurrence_yr"] - ref["first_recurrence_yr"]
            e["peak_slip_rate_rel"] = (r["peak_slip_rate"] - ref["peak_slip_rate"]) / ref["peak_slip_rate"]
        errors[name] = e

    if HEADLINE not in runs:
        raise SystemExit(f"--precisions must include the headline {HEADLINE}")
    hname = HEADLINE
    head = runs[hname]
    stg = kv[hname]
    traffic = STAGE_TRAFFIC[args.config].get(hname)
    roof = dict(kernel=stage_kernel(hname, D, n), bound="hbm",
                achieved=stg["stage_gbs"], peak=HBM_PEAK_GBS, unit="GB/s",
                frac=stg["stage_gbs"] / HBM_PEAK_GBS, traffic=traffic,
                bytes_per_launch=stg["bytes"], launch_us=stg["stage_ms"] * 1e3,
                frac_of_read_stream=stg["stage_gbs"] / read_ceil["gbs"])
    # the FP32 product K v on its own (the storage the north-star K v figure
    # names), HIP events on the handle's stream
    k32 = kv["fp32"]
    roof_kv32 = dict(kernel=matvec_kernel("fp32", D) + ", K v of the FP32 operator", bound="hbm",
                     achieved=k32["matvec_gbs"], peak=HBM_PEAK_GBS, unit="GB/s",
                     frac=k32["matvec_gbs"] / HBM_PEAK_GBS,
                     traffic=MATVEC_TRAFFIC[args.config].get("fp32"),
                     bytes_per_launch=k32["bytes"], launch_us=k32["matvec_ms"] * 1e3,
                     frac_of_read_stream=k32["matvec_gbs"] / read_ceil["gbs"])
    out = {
        "metric": METRIC,
        "value": head["sim_years_per_hour"],
        "unit": "sim-years/hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_segment"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[hname],
        "data": "synthetic (seeded fault and friction fields)",
        "config": {"workload": f"{args.config}: planar strike-slip fault 100x40 km, {args.ds:g} km "
                               f"patches, {n} patches x 2 slip components = {D} DOF, dense K "
                               f"{D}x{D}, {args.years:g} yr, rk5dp rtol {args.rtol:g}; one step "
                               f"= one of {args.steps} continued segments of "
                               f"{args.years / args.steps:g} yr (together the whole cycle)",
                   "headline_precision": hname, "parallelism": "1 GPU"},
        "speedup_mixed_vs_fp64": (head["sim_years_per_hour"] / ref["sim_years_per_hour"]) if ref else None,
        "speedup_vs_fp64": ({k: r["sim_years_per_hour"] / ref["sim_years_per_hour"]
                             for k, r in runs.items()} if ref else None),
        "runs": runs,
        "kv": kv,
        "errors_vs_fp64": errors,
        "assembly": dict(pairs=pairs, ms=asm_ms, pairs_per_s=pairs / asm_ms * 1e3,
                         f_pair=F_PAIR, tflops=pairs / asm_ms * 1e3 * F_PAIR / 1e12,
                         frac_fp64_peak=pairs / asm_ms * 1e3 * F_PAIR / 1e12 / FP64_PEAK_TFLOPS),
        "roofline": roof,
        "roofline_kv_fp32": roof_kv32,
        "hbm_read_stream": read_ceil,
    }
    if not args.no_cpu and ref is not None:
        K64 = fs.dense()
        out["cpu_baseline"] = cpu_baseline(g, f, K64, ref["rhs_evals"], args.years,
                                           args.cpu_budget, log)
        del K64
    fs.close()
    if args.config == "c3":
        out["assembly_rough_triangles"] = rough_triangle_assembly(
            args.ds, args.rough_cycle_years, args.rtol)
    if not args.no_oracle_c1:
        out["errors_vs_oracle_c1"] = oracle_c1(args.rtol)
    print(json.dumps(out))


def oracle_c1(rtol):
    """c1 (2 km patches, 2000 DOF, 300 yr) on the device in every storage
    format against the FP64 oracle's cycle on the host cores: onsets,
    recurrence, peak slip rates and the log10 max|v| trace, scored as the
    reference's rsf_accuracy.py does."""
    import oracle
    from interact_amd import accuracy, faults, lib
    from interact_amd.model import FaultSystem, initial_state
    g = faults.planar_grid(2.0)
    f = faults.cycle_fields(g)
    n = len(g["l"])
    store = lib.PREC_FP32 | lib.PREC_BF16X2 | lib.PREC_BF16NF
    fs = FaultSystem(g)
    fs.assemble(lib.SLIP_VECTOR, shear_modulus=f["shear_modulus"], store=store)
    fs.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"])
    fs.set_loading(f["vpl"])
    X0 = initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=2)
    K = fs.dense()
    xpl = np.zeros(2 * n)
    xpl[0::2] = -f["vpl"]
    threads = len(os.sched_getaffinity(0))
    P = oracle.RSFProblem(K, f["a"], f["b"], f["dc"], K @ xpl, ncomp=2, eta=f["eta"],
                          v0=f["v0"], f0=f["f0"], nthreads=threads)
    t_end = 300 * SEC_PER_YEAR
    t0 = time.perf_counter()
    ro = P.solve(X0, t_end, rtol=rtol)
    t_or = time.perf_counter() - t0
    out = dict(workload="c1: 100x40 km at 2 km patches, 1000 patches x 2 = 2000 DOF, 300 yr, "
                        f"rk5dp rtol {rtol:g}; the device's FP64 operator in the oracle",
               oracle=dict(wall_s=t_or, threads=threads, steps=ro["steps"],
                           onsets_yr=[float(x) for x in accuracy.onsets(ro["events"])]))
    for name in ("fp64", "fp32", "bf16x2", "bf16nf"):
        fs.set_state(X0)
        t0 = time.perf_counter()
        st = fs.solve(t_end, rtol=rtol, prec=lib.PREC_BY_NAME[name], steps_per_launch=64)
        wall = time.perf_counter() - t0
        sc = accuracy.score(ro["monitor"], ro["events"], fs.monitor(), fs.events())
        sc.update(steps=st["steps"], wall_s=wall)
        out[name] = sc
        log(f"c1 vs oracle, {name}: onsets {sc['donset']} yr, peaks {sc['dpeak']}, "
            f"trace max {sc['maxdev']:.3g} rms {sc['rmsdev']:.3g} on the monitor rows, "
            f"max {sc['maxdev_uniform']:.3g} rms {sc['rmsdev_uniform']:.3g} on a uniform "
            f"grid (log10 v), {wall:.2f} s")
    log(f"c1 oracle cycle: {ro['steps']} steps in {t_or:.1f} s on {threads} threads")
    fs.close()
    return out


ENS_OUTLINES = [(40, 20), (60, 24), (80, 32), (100, 40)]   # km
ENS_MEMBERS = 32


def ensemble_bench(args):
    """c4: one ensemble of 4 geometries x 32 members against the same
    members run one at a time through the single-fault FP32 solver."""
    from interact_amd import faults, lib
    from interact_amd.model import Ensemble, FaultSystem, initial_state
    import oracle
    years = 100.0 if args.years is None else args.years
    t_end = years * SEC_PER_YEAR
    nm = ENS_MEMBERS
    two = not args.no_two_byte
    geoms, gdat = [], []
    for (L, W) in ENS_OUTLINES:
        g = faults.planar_grid(1.0, L, W)
        fs = FaultSystem(g)
        fs.assemble(lib.SLIP_VECTOR, shear_modulus=32e9,
                    store=lib.PREC_FP32 | (lib.PREC_BF16NF if two else 0))
        geoms.append(fs)
        gdat.append((g, L, W))
    npat = [len(g["l"]) for g, _, _ in gdat]
    log(f"ensemble: geometries {npat} patches, {nm} members each")
    ens = Ensemble(geoms, nm)
    members = []
    for gi, (g, L, W) in enumerate(gdat):
        for k in range(nm):
            f = faults.ensemble_fields(g, L, W, k, gi)
            X0 = initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=2)
            members.append((gi, f, X0))

    def configure(m, f, X0):
        m.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"])
        m.set_loading(f["vpl"])
        m.set_state(X0)

    for m, (gi, f, X0) in zip(ens.members, members):
        configure(m, f, X0)
    t0 = time.perf_counter()
    choice, ms, shapes = ens.tune(rtol=args.rtol)
    tune_s = time.perf_counter() - t0
    log("tuner: " + ", ".join(f"{shapes[i]} {ms[i]*1e3:.1f} us" for i in range(len(ms))
                              if ms[i] > 0) + f" -> {shapes[choice]} ({tune_s:.2f} s)")
    prod_ms, epi_ms = ens.bench(20, rtol=args.rtol)
    log(f"tuned shape: product {prod_ms * 1e3:.1f} us, epilogue {epi_ms * 1e3:.1f} us")
    for _ in range(args.warmup):
        for m, (gi, f, X0) in zip(ens.members, members):
            m.set_state(X0)
        ens.solve(t_end, rtol=args.rtol, steps_per_launch=64)
    walls = []
    for _ in range(args.steps):
        for m, (gi, f, X0) in zip(ens.members, members):
            m.set_state(X0)
        t0 = time.perf_counter()
        st = ens.solve(t_end, rtol=args.rtol, steps_per_launch=64)
        walls.append(time.perf_counter() - t0)
    wall = float(np.mean(walls))
    M = len(members)
    ens_rate = M * years / (wall / 3600.0)
    steps = [s["steps"] + s["rejects"] for s in st]
    log(f"ensemble: {M} members x {years:g} yr in {wall:.2f} s ({ens_rate:.3e} sim-yr/h); "
        f"attempts per member min {min(steps)} max {max(steps)}")
    ev32 = [m.events() for m in ens.members]

    def onset_diff(evs_a, evs_b):
        """worst onset difference over all events, members with another
        event count, and the first-onset differences of every member
        (median, 90th percentile, worst)"""
        worst, nother, first = 0.0, 0, []
        for ea, eb in zip(evs_a, evs_b):
            oa = ea[ea[:, 1] > 0, 0] / SEC_PER_YEAR if len(ea) else np.zeros(0)
            ob = eb[eb[:, 1] > 0, 0] / SEC_PER_YEAR if len(eb) else np.zeros(0)
            k = min(len(oa), len(ob))
            if k:
                worst = max(worst, float(np.abs(oa[:k] - ob[:k]).max()))
                first.append(abs(float(oa[0] - ob[0])))
            nother += int(len(oa) != len(ob))
        f = np.array(first) if first else np.zeros(1)
        return worst, nother, dict(median=float(np.median(f)), p90=float(np.percentile(f, 90)),
                                   worst=float(f.max()))

    if two:
        # the same members on the two-byte operator
        bh = ens.set_operator(lib.PREC_BF16NF)
        t0 = time.perf_counter()
        choice_h, ms_h, _ = ens.tune(rtol=args.rtol)
        tune_h = time.perf_counter() - t0
        log("two-byte tuner: " + ", ".join(f"{shapes[i]} {ms_h[i]*1e3:.1f} us"
                                           for i in range(len(ms_h)) if ms_h[i] > 0) +
            f" -> {shapes[choice_h]} ({tune_h:.2f} s)")
        prod_h, epi_h = ens.bench(20, rtol=args.rtol)
        log(f"two-byte tuned shape: product {prod_h * 1e3:.1f} us, epilogue {epi_h * 1e3:.1f} us")
        for _ in range(args.warmup):
            for m, (gi, f, X0) in zip(ens.members, members):
                m.set_state(X0)
            ens.solve(t_end, rtol=args.rtol, steps_per_launch=64)
        walls_h = []
        for _ in range(args.steps):
            for m, (gi, f, X0) in zip(ens.members, members):
                m.set_state(X0)
            t0 = time.perf_counter()
            st_h = ens.solve(t_end, rtol=args.rtol, steps_per_launch=64)
            walls_h.append(time.perf_counter() - t0)
        wall_h = float(np.mean(walls_h))
        evh = [m.events() for m in ens.members]
        on_h32, n_h32, first_h32 = onset_diff(evh, ev32)
        steps_h = [q["steps"] + q["rejects"] for q in st_h]
        log(f"two-byte ensemble: {M} members x {years:g} yr in {wall_h:.2f} s "
            f"({M * years / (wall_h / 3600.0):.3e} sim-yr/h); attempts per member min "
            f"{min(steps_h)} max {max(steps_h)}; onsets vs the FP32 ensemble: worst "
            f"{on_h32:.2e} yr, first onsets {first_h32}, {n_h32} members with another count")
    # the same members one after another through the single-fault solver
    nsing = M if args.singles is None else min(args.singles, M)
    single_s, worst_on, n_on, ev_s32 = 0.0, 0.0, 0, []
    for i in range(nsing):
        gi, f, X0 = members[i]
        s = geoms[gi].member()
        configure(s, f, X0)
        t0 = time.perf_counter()
        s.solve(t_end, rtol=args.rtol, prec=lib.PREC_FP32, steps_per_launch=64)
        single_s += time.perf_counter() - t0
        e1, e2 = s.events(), ev32[i]
        ev_s32.append(e1)
        o1 = e1[e1[:, 1] > 0, 0] / SEC_PER_YEAR if len(e1) else np.zeros(0)
        o2 = e2[e2[:, 1] > 0, 0] / SEC_PER_YEAR if len(e2) else np.zeros(0)
        k = min(len(o1), len(o2))
        if k:
            worst_on = max(worst_on, float(np.abs(o1[:k] - o2[:k]).max()))
        n_on += int(len(o1) != len(o2))
        s.close()
        if (i + 1) % nm == 0:
            log(f"singles: {i + 1} of {nsing} done, {single_s:.1f} s")
    seq_rate = nsing * years / (single_s / 3600.0)
    log(f"singles: {nsing} runs one after another in {single_s:.2f} s ({seq_rate:.3e} sim-yr/h); "
        f"onsets vs ensemble: worst {worst_on:.2e} yr, {n_on} members with another count")
    if two:
        # the same members one after another on the single-fault bf16nf operator
        single_h, evs = 0.0, []
        for i in range(nsing):
            gi, f, X0 = members[i]
            s = geoms[gi].member()
            configure(s, f, X0)
            t0 = time.perf_counter()
            s.solve(t_end, rtol=args.rtol, prec=lib.PREC_BF16NF, steps_per_launch=64)
            single_h += time.perf_counter() - t0
            evs.append(s.events())
            s.close()
            if (i + 1) % nm == 0:
                log(f"bf16nf singles: {i + 1} of {nsing} done, {single_h:.1f} s")
        on_hs, n_hs, first_hs = onset_diff(evh[:nsing], evs)
        on_ss, n_ss, first_ss = onset_diff(evs, ev_s32)
        seq_h = nsing * years / (single_h / 3600.0)
        log(f"bf16nf singles: {nsing} runs in {single_h:.2f} s ({seq_h:.3e} sim-yr/h); onsets vs "
            f"the two-byte ensemble: worst {on_hs:.2e} yr, first onsets {first_hs}, {n_hs} "
            f"members with another count; vs the FP32 single runs: worst {on_ss:.2e} yr, first "
            f"onsets {first_ss}, {n_ss} members with another count")
    # roofline of the batched product kernel at the tuned shape (HIP events
    # around repeated launches of a middle stage, on the ensemble's stream)
    launch_s = prod_ms * 1e-3
    flops = sum(2.0 * (2 * n) * (2 * n) * nm for n in npat)
    k_bytes = sum((2 * n) * (2 * n) * 4.0 for n in npat)
    # 2 flops per operator byte-quarter and 32 members: 16 FLOP/B, below the
    # FP32 MFMA ridge (157.3 TFLOP/s / 8 TB/s = 19.7 FLOP/B), so the roofline
    # that bounds the product is the operator stream
    kname = "k_prod_sk" if shapes[choice][6] > 0 else "k_prod_b"
    roof = dict(kernel=f"{kname} (K [v_1..v_{nm}] on v_mfma_f32_16x16x4_f32 over MFMA-tiled "
                       f"operators), shape {shapes[choice]}",
                bound="hbm", achieved=k_bytes / launch_s / 1e9, peak=HBM_PEAK_GBS,
                unit="GB/s", frac=k_bytes / launch_s / 1e9 / HBM_PEAK_GBS,
                traffic=PRODUCT_TRAFFIC_C4.get(tuple(shapes[choice])) if nm == 32 else None,
                bytes_per_launch=k_bytes, launch_us=launch_s * 1e6,
                flops_per_launch=flops, mfma_tflops=flops / launch_s / 1e12,
                mfma_frac=flops / launch_s / 1e12 / FP32_MFMA_PEAK_TFLOPS,
                stage_us=ms[choice] * 1e3, epilogue_us=epi_ms * 1e3)
    out = {
        "metric": "ensemble sim-years/hour vs 128 sequential single runs",
        "value": ens_rate,
        "unit": "sim-years/hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 K (f32 MFMA, f64 accumulation every 256 columns) + f64 state",
        "data": "synthetic (seeded friction draws)",
        "config": {"workload": f"c4: {len(npat)} planar strike-slip outlines "
                               f"{[f'{L}x{W} km' for L, W in ENS_OUTLINES]} at 1 km patches "
                               f"({[2 * n for n in npat]} DOF) x {nm} friction draws, "
                               f"{years:g} yr each, rk5dp rtol {args.rtol:g}",
                   "parallelism": "1 GPU"},
        "sequential_single_runs": dict(runs=nsing, wall_s=single_s, sim_years_per_hour=seq_rate),
        "speedup_vs_sequential": ens_rate / seq_rate,
        "parity_vs_single": dict(worst_onset_diff_yr=worst_on, members_with_other_count=n_on),
        "tuner": dict(choice=list(shapes[choice]), ms={str(tuple(shapes[i])): ms[i]
                                                        for i in range(len(ms))}, wall_s=tune_s),
        "attempts_per_member": dict(min=min(steps), max=max(steps)),
        "roofline": roof,
    }
    if two:
        rate_h = M * years / (wall_h / 3600.0)
        pad = sum(2.0 * (-(-2 * n // 256) * 256) ** 2 for n in npat)   # hi tiles as streamed
        lo_bytes = bh - pad
        hb = sum((2 * n) * (2 * n) * 2.0 for n in npat) + lo_bytes
        ls = prod_h * 1e-3
        out["two_byte"] = dict(
            operator="bf16 hi of every entry (16 x 32 tiles, v_mfma_f32_16x16x32_bf16) + bf16 lo "
                     "tiles of each 16-row block in the 256-column segments its rows need "
                     "(energy tolerance 2^-6); slip rates as three bf16 terms summing to FP32; "
                     "FP32 accumulation, FP64 every 256 columns",
            value=rate_h, unit="sim-years/hour", wall_s=wall_h,
            speedup_vs_fp32_ensemble=wall / wall_h,
            sequential_single_runs_bf16nf=dict(runs=nsing, wall_s=single_h,
                                               sim_years_per_hour=seq_h),
            speedup_vs_sequential_bf16nf=rate_h / seq_h,
            speedup_vs_sequential_fp32=rate_h / seq_rate,
            parity_vs_fp32_ensemble=dict(worst_onset_diff_yr=on_h32,
                                         first_onset_diff_yr=first_h32,
                                         members_with_other_count=n_h32),
            parity_vs_single_bf16nf=dict(worst_onset_diff_yr=on_hs,
                                         first_onset_diff_yr=first_hs,
                                         members_with_other_count=n_hs),
            single_bf16nf_vs_single_fp32=dict(worst_onset_diff_yr=on_ss,
                                              first_onset_diff_yr=first_ss,
                                              members_with_other_count=n_ss),
            attempts_per_member=dict(min=min(steps_h), max=max(steps_h)),
            tuner=dict(choice=list(shapes[choice_h]), wall_s=tune_h,
                       ms={str(tuple(shapes[i])): ms_h[i] for i in range(len(ms_h)) if ms_h[i] > 0}),
            roofline=dict(kernel=f"k_prod_sk_h (two-byte operator), shape {shapes[choice_h]}",
                          bound="hbm", achieved=hb / ls / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                          frac=hb / ls / 1e9 / HBM_PEAK_GBS,
                          traffic=PRODUCT_TRAFFIC_C4_H.get(tuple(shapes[choice_h])),
                          bytes_per_launch=hb, lo_bytes=lo_bytes, streamed_bytes=bh,
                          launch_us=prod_h * 1e3, epilogue_us=epi_h * 1e3,
                          stage_us=ms_h[choice_h] * 1e3))
    if not args.no_cpu:
        # FP64 oracle right-hand sides on each geometry, scaled by every
        # member's RHS count from the ensemble run
        threads = len(os.sched_getaffinity(0))
        t_rhs = []
        for gi, (g, L, W) in enumerate(gdat):
            K = geoms[gi].dense()
            f = faults.ensemble_fields(g, L, W, 0, gi)
            xpl = np.zeros(K.shape[0])
            xpl[0::2] = -f["vpl"]
            P = oracle.RSFProblem(K, f["a"], f["b"], f["dc"], K @ xpl, ncomp=2, eta=f["eta"],
                                  v0=f["v0"], f0=f["f0"], nthreads=threads)
            X = members[gi * nm][2]
            P.rhs(X)
            nev, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < args.cpu_budget / len(gdat):
                P.rhs(X)
                nev += 1
            t_rhs.append((time.perf_counter() - t0) / nev)
            del K, P
        cpu_s = sum(st[i]["rhs_evals"] * t_rhs[members[i][0]] for i in range(M))
        log("cpu oracle: " + ", ".join(f"{t*1e3:.2f} ms/RHS" for t in t_rhs) +
            f" -> {cpu_s:.0f} s for the ensemble's RHS evaluations")
        out["cpu_baseline"] = dict(
            value=M * years / (cpu_s / 3600.0), unit="sim-years/hour", cores=threads,
            kind="port",
            sample=(f"FP64 right-hand sides (dense K v + RHS, OpenMP over rows, {threads} "
                    f"threads) timed for {args.cpu_budget / len(gdat):.1f} s on each geometry "
                    f"({', '.join(f'{t*1e3:.2f}' for t in t_rhs)} ms each), scaled by every "
                    f"member's RHS count in the ensemble run"))
    ens.close()
    for g in geoms:
        g.close()
    return out


def ensemble_kernels_only(args):
    """c4 kernels for counter and trace runs: the ensemble's setup (no
    solve), the tuner, which launches a middle stage (product + epilogue)
    21 times per unit shape, and 20 more launches of each for the tuned
    shape timed by HIP events."""
    from interact_amd import faults, lib
    from interact_amd.model import Ensemble, FaultSystem, initial_state
    geoms, gdat = [], []
    for (L, W) in ENS_OUTLINES:
        g = faults.planar_grid(1.0, L, W)
        fs = FaultSystem(g)
        fs.assemble(lib.SLIP_VECTOR, shear_modulus=32e9, store=lib.PREC_FP32)
        geoms.append(fs)
        gdat.append((g, L, W))
    ens = Ensemble(geoms, ENS_MEMBERS)
    for gi, (g, L, W) in enumerate(gdat):
        for k in range(ENS_MEMBERS):
            f = faults.ensemble_fields(g, L, W, k, gi)
            m = ens.members[gi * ENS_MEMBERS + k]
            m.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"])
            m.set_loading(f["vpl"])
            m.set_state(initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=2))
    choice, ms, shapes = ens.tune(rtol=args.rtol)
    log("kernels-only c4: " + ", ".join(f"{shapes[i]} {ms[i]*1e3:.1f} us" for i in range(len(ms))
                                         if ms[i] > 0))
    pm, pe = ens.bench(20, rtol=args.rtol)
    log(f"kernels-only c4: tuned shape {shapes[choice]}: product {pm*1e3:.1f} us, "
        f"epilogue {pe*1e3:.1f} us by HIP events (20 launches each)")
    if not args.no_two_byte:
        bh = ens.set_operator(lib.PREC_BF16NF)
        choice, ms, shapes = ens.tune(rtol=args.rtol)
        log("kernels-only c4 two-byte: " + ", ".join(
            f"{shapes[i]} {ms[i]*1e3:.1f} us" for i in range(len(ms)) if ms[i] > 0))
        pm, pe = ens.bench(20, rtol=args.rtol)
        log(f"kernels-only c4 two-byte: {bh / 1e6:.1f} MB streamed per product; tuned shape "
            f"{shapes[choice]}: product {pm*1e3:.1f} us, epilogue {pe*1e3:.1f} us by HIP "
            f"events (20 launches each)")
    ens.close()
    for g in geoms:
        g.close()


def adjoint_kernels_only(args):
    """c5 kernels for counter and trace runs: one gradient over 64 steps of
    0.05 yr from the initial (interseismic) state of the 1 km fault and 20
    launches of the backward sweep's K^T w."""
    from interact_amd import faults, lib
    from interact_amd.model import FaultSystem, initial_state
    g = faults.planar_grid(1.0)
    f = faults.cycle_fields(g)
    n = len(g["l"])
    fs = FaultSystem(g)
    fs.assemble(lib.SLIP_VECTOR, shear_modulus=f["shear_modulus"], store=lib.PREC_FP32)
    fs.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"])
    fs.set_loading(f["vpl"])
    fs.set_state(initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=2))
    steps = np.full(64, 0.05 * SEC_PER_YEAR)
    recv = np.sort(np.random.default_rng(20261202).choice(n, 64, replace=False)).astype(np.int32)
    tr = fs.replay(steps, recv)["trace"]
    weight = np.full(65, 1.0 / 65)
    r = fs.adjoint_gradient(steps, recv, tr + 0.01, weight, prec=lib.PREC_FP64)
    ms = fs.adjoint_bench(lib.PREC_FP64, 20)
    log(f"kernels-only c5: J {r['J']:.6e}, K^T w {ms * 1e3:.1f} us")
    fs.close()


def adjoint_bench(args):
    """c5: gradient of the receiver misfit on the 1 km fault."""
    from interact_amd import faults, lib
    from interact_amd.model import FaultSystem, initial_state
    import oracle
    window = 50.0
    rtol = args.rtol
    g = faults.planar_grid(1.0)
    f = faults.cycle_fields(g)
    n = len(g["l"])
    D = 2 * n
    fs = FaultSystem(g)
    fs.assemble(lib.SLIP_VECTOR, shear_modulus=f["shear_modulus"], store=lib.PREC_FP32)
    fric = dict(v0=f["v0"], f0=f["f0"], eta=f["eta"])
    fs.set_friction(f["a"], f["b"], f["dc"], **fric)
    fs.set_loading(f["vpl"])
    X0 = initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=2)
    fs.set_state(X0)
    # the window starts after the first event (its arrest)
    fs.solve(20.0 * SEC_PER_YEAR, rtol=rtol, prec=lib.PREC_FP64, steps_per_launch=64)
    t_first = float("nan")
    for _ in range(20):
        fs.solve(fs.get_state()[1] + 2.0 * SEC_PER_YEAR, rtol=rtol, prec=lib.PREC_FP64,
                 steps_per_launch=64)
        ev = fs.events()
        arr = ev[ev[:, 1] < 0, 0] if len(ev) else []
        if len(arr):
            t_first = first_onset(ev)
            break
    fs.solve(fs.get_state()[1] + 1.0 * SEC_PER_YEAR, rtol=rtol, prec=lib.PREC_FP64,
             steps_per_launch=64)
    Xs, ts = fs.get_state()
    log(f"adjoint: first event {t_first:.3f} yr, window starts {ts / SEC_PER_YEAR:.3f} yr")
    fs.set_step_log(1_000_000)
    fs.solve(ts + window * SEC_PER_YEAR, rtol=rtol, prec=lib.PREC_FP64, steps_per_launch=64)
    steps = fs.step_log()
    ev = fs.events()
    n_on = int(np.sum(ev[:, 1] > 0)) if len(ev) else 0
    N = len(steps)
    rng = np.random.default_rng(20261202)
    recv = np.sort(rng.choice(n, 64, replace=False)).astype(np.int32)
    # target: the same window with perturbed b and D_c, on its own steps
    bt = f["b"] * (1 + 0.02 * rng.standard_normal(n))
    dt = f["dc"] * (1 + 0.05 * rng.standard_normal(n))
    fs.set_friction(f["a"], bt, dt, **fric)
    fs.set_state(Xs, ts)
    fs.solve(ts + 1.01 * window * SEC_PER_YEAR, rtol=rtol, prec=lib.PREC_FP64, steps_per_launch=64)
    st_t = fs.step_log()
    fs.set_state(Xs, ts)
    tr_t = fs.replay(st_t, recv)["trace"]
    tt = np.concatenate([[0.0], np.cumsum(st_t)])
    tm = np.concatenate([[0.0], np.cumsum(steps)])
    target = np.stack([np.interp(tm, tt, tr_t[:, r]) for r in range(len(recv))], 1)
    weight = np.zeros(N + 1)
    weight[1:] = steps / steps.sum()
    fs.set_friction(f["a"], f["b"], f["dc"], **fric)
    fs.set_state(Xs, ts)
    log(f"adjoint: {N} steps over {window:g} yr, {n_on} onsets in the window, 64 receivers")

    def grad():
        return fs.adjoint_gradient(steps, recv, target, weight, prec=lib.PREC_FP64,
                                   checkpoint_every=256)
    for _ in range(args.warmup):
        grad()
    walls = []
    for _ in range(max(args.steps, 1)):
        t0 = time.perf_counter()
        r = grad()
        walls.append(time.perf_counter() - t0)
    wall = float(np.mean(walls))
    t0 = time.perf_counter()
    J_rep = fs.replay(steps, recv, target, weight)["J"]
    replay_s = time.perf_counter() - t0
    g_amb, g_dc = -r["grad_b"], r["grad_dc"]
    log(f"adjoint: J {r['J']:.6e} (replay {J_rep:.6e}); gradient in {wall:.2f} s, "
        f"replay alone {replay_s:.2f} s; |dJ/d(a-b)| {np.linalg.norm(g_amb):.3e} "
        f"|dJ/dDc| {np.linalg.norm(g_dc):.3e}")
    # the same gradient with the FP32 operator (FP64 state and adjoint)
    fs.set_state(Xs, ts)
    fs.adjoint_gradient(steps, recv, target, weight, prec=lib.PREC_FP32, checkpoint_every=256)
    t0 = time.perf_counter()
    r32 = fs.adjoint_gradient(steps, recv, target, weight, prec=lib.PREC_FP32,
                              checkpoint_every=256)
    wall32 = time.perf_counter() - t0
    fp32 = dict(gradient_s=wall32, J_rel=abs(r32["J"] - r["J"]) / abs(r["J"]),
                **{k + "_rel": float(np.linalg.norm(r32[k] - r[k]) / np.linalg.norm(r[k]))
                   for k in ("grad_a", "grad_b", "grad_dc")})
    log(f"adjoint: FP32 operator gradient in {wall32:.2f} s; relative differences to FP64 "
        f"J {fp32['J_rel']:.1e}, grad_a {fp32['grad_a_rel']:.1e}, "
        f"grad_b {fp32['grad_b_rel']:.1e}, grad_dc {fp32['grad_dc_rel']:.1e}")
    # central differences on 32 random directions of (a - b at fixed a, D_c)
    amb0 = f["a"] - f["b"]
    eps = 1e-5
    rels = []
    nd = max(0, args.fd_directions)
    for k in range(nd):
        d1 = rng.standard_normal(n) * np.abs(amb0)
        d2 = rng.standard_normal(n) * f["dc"]
        Js = []
        for sgn in (1.0, -1.0):
            fs.set_friction(f["a"], f["a"] - (amb0 + sgn * eps * d1), f["dc"] + sgn * eps * d2,
                            **fric)
            fs.set_state(Xs, ts)
            Js.append(fs.replay(steps, recv, target, weight)["J"])
        fdv = (Js[0] - Js[1]) / (2 * eps)
        ad = g_amb @ d1 + g_dc @ d2
        rels.append(abs(fdv - ad) / abs(ad))
        if k % 8 == 7 or k == nd - 1:
            log(f"adjoint: {k + 1} of {nd} directions, worst relative difference {max(rels):.2e}")
    fs.set_friction(f["a"], f["b"], f["dc"], **fric)
    fs.set_state(Xs, ts)
    # the oracle's FP64 reference adjoint on the window's first steps
    Np = min(16, N)
    K64 = fs.dense()
    xpl = np.zeros(D)
    xpl[0::2] = -f["vpl"]
    threads = len(os.sched_getaffinity(0))
    P = oracle.RSFProblem(K64, f["a"], f["b"], f["dc"], K64 @ xpl, ncomp=2, eta=f["eta"],
                          v0=f["v0"], f0=f["f0"], nthreads=threads)
    tp, wp = target[:Np + 1], weight[:Np + 1]
    t0 = time.perf_counter()
    Jo, gao, gbo, gdo = P.adjoint(Xs, steps[:Np], recv, tp, wp, 0)
    cpu_s = time.perf_counter() - t0
    rp = fs.adjoint_gradient(steps[:Np], recv, tp, wp, prec=lib.PREC_FP64)
    ref_err = dict(J=abs(rp["J"] - Jo) / abs(Jo),
                   grad_a=float(np.linalg.norm(rp["grad_a"] - gao) / np.linalg.norm(gao)),
                   grad_b=float(np.linalg.norm(rp["grad_b"] - gbo) / np.linalg.norm(gbo)),
                   grad_dc=float(np.linalg.norm(rp["grad_dc"] - gdo) / np.linalg.norm(gdo)))
    log(f"adjoint: oracle FP64 reference adjoint on {Np} steps in {cpu_s:.1f} s on {threads} "
        f"threads; relative differences {ref_err}")
    del K64, P
    # roofline of the backward sweep's K^T w (HIP events on the handle's stream)
    kt_ms = fs.adjoint_bench(lib.PREC_FP64, 20)
    kt_bytes = float(D) * D * 8 + 8.0 * D * 2
    fs.close()
    rate = window / (wall / 3600.0)
    cpu_rate = window / (cpu_s / Np * N / 3600.0)
    return {
        "metric": "adjoint gradient: simulated years/hour of replay + backward sweep",
        "value": rate,
        "unit": "sim-years/hour",
        "n_gpus": 1,
        "steps": max(args.steps, 1),
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 K v, K^T w and state",
        "data": "synthetic (seeded fault, friction, receivers and target)",
        "config": {"workload": f"c5: planar fault 100x40 km, 1 km patches, {D} DOF; "
                               f"ln|v| misfit at 64 receivers over a {window:g} yr window "
                               f"after the first event ({N} rk5dp steps at rtol {rtol:g}); "
                               f"gradient in a - b and D_c of {n} patches",
                   "parallelism": "1 GPU"},
        "adjoint": dict(J=r["J"], steps=N, onsets_in_window=n_on, gradient_s=wall,
                        replay_s=replay_s, fd_directions=nd, fd_eps=eps,
                        fd_worst_rel=float(max(rels)) if rels else None,
                        fd_median_rel=float(np.median(rels)) if rels else None,
                        oracle_steps=Np, oracle_rel=ref_err, first_event_yr=t_first,
                        window_start_yr=ts / SEC_PER_YEAR, fp32_operator=fp32),
        "roofline": dict(kernel="k_adj_kt<double> (K^T w of the backward sweep, row-split "
                                "partials)", bound="hbm", achieved=kt_bytes / kt_ms / 1e6,
                         peak=HBM_PEAK_GBS, unit="GB/s", frac=kt_bytes / kt_ms / 1e6 / HBM_PEAK_GBS,
                         traffic=ADJ_KT_TRAFFIC_C5, bytes_per_launch=kt_bytes,
                         launch_us=kt_ms * 1e3),
        "cpu_baseline": dict(value=cpu_rate, unit="sim-years/hour", cores=threads, kind="port",
                             sample=(f"the oracle's FP64 reference adjoint (replay + backward, "
                                     f"OpenMP products on {threads} threads) on the window's "
                                     f"first {Np} steps in {cpu_s:.1f} s, scaled to the "
                                     f"{N} steps of the window")),
    }


ROUGH_TRI_TOL = 3e-3   # vertical-edge tolerance [rad] of the rough mesh's products and cycles


def rough_triangle_assembly(ds_km, cycle_years=0.0, rtol=1e-8, precisions="fp64,fp32,bf16nf",
                            products=True):
    """FP64 assembly throughput on the rough variant of the same outline:
    two Nikkhoo-Walter triangles per cell (32000 at 0.5 km) on a seeded
    Hurst 0.8 self-affine surface with amplitude-to-wavelength ratio 1e-3,
    two slip components (64000 DOF, K 32.8 GB in FP64).  Timed with the
    reference's vertical-edge rule; then assembled again with the 3e-3 rad
    tolerance (the reference's rule leaves rounding noise near the mesh's
    near-vertical cell edges) in FP64, FP32 and bf16nf for the K v products,
    and (cycle_years > 0) cycles of the benchmark's friction fields."""
    from interact_amd import faults, lib
    from interact_amd.model import FaultSystem, initial_state
    nx, nz = int(round(100.0 / ds_km)), int(round(40.0 / ds_km))
    g = faults.planar_tri_grid(ds_km, offset=faults.rough_offset(nx, nz, ds_km))
    n = len(g["l"])
    D = 2 * n
    fs = FaultSystem(g)
    fs.assemble(lib.SLIP_VECTOR, shear_modulus=32e9)
    ms = fs.assemble(lib.SLIP_VECTOR, shear_modulus=32e9)
    pairs = n * n
    log(f"assembly, rough triangles: {n} triangles, {pairs} pairs in {ms:.1f} ms = "
        f"{pairs / ms * 1e3:.3e} pairs/s")
    rate = pairs / ms * 1e3
    out = dict(triangles=n, pairs=pairs, ms=ms, pairs_per_s=rate,
               entries_per_s=4 * rate, dof=D, f_pair=F_PAIR_TRI,
               tflops=rate * F_PAIR_TRI / 1e12,
               frac_fp64_peak=rate * F_PAIR_TRI / 1e12 / FP64_PEAK_TFLOPS)
    if not products:
        fs.close()
        return out
    f = faults.cycle_fields(g)
    fs.set_triangle_tolerance(ROUGH_TRI_TOL)
    fs.assemble(lib.SLIP_VECTOR, shear_modulus=f["shear_modulus"],
                store=lib.PREC_FP32 | lib.PREC_BF16NF)
    fs.set_friction(f["a"], f["b"], f["dc"], v0=f["v0"], f0=f["f0"], eta=f["eta"])
    fs.set_loading(f["vpl"])
    X0 = initial_state(f["tau"], f["sigma"], f["vel"], f["a"], f["v0"], ncomp=2)
    fs.set_state(X0)
    i = np.arange(D)
    xa = np.sin(i + 1.0) + 0.3 * np.cos(3.0 * i)
    y64 = fs.matvec(xa, lib.PREC_FP64)
    kv = {}
    for name in ("fp64", "fp32", "bf16nf"):
        p = lib.PREC_BY_NAME[name]
        y = fs.matvec(xa, p)
        mv_ms = float(np.median([fs.matvec_bench(p, 20) for _ in range(3)]))
        st_ms = float(np.median([fs.stage_bench(p, 20) for _ in range(3)]))
        by = fs.operator_bytes(p) + 16.0 * D
        kv[name] = dict(relerr=float(np.linalg.norm(y - y64) / np.linalg.norm(y64)),
                        matvec_ms=mv_ms, stage_ms=st_ms, matvec_gbs=by / mv_ms / 1e6,
                        stage_gbs=by / st_ms / 1e6, bytes=by,
                        matvec_frac=by / mv_ms / 1e6 / HBM_PEAK_GBS)
        log(f"rough K.v {name}: relerr {kv[name]['relerr']:.2e}  matvec {mv_ms*1e3:.1f} us "
            f"{kv[name]['matvec_gbs']:.0f} GB/s  stage {st_ms*1e3:.1f} us "
            f"{kv[name]['stage_gbs']:.0f} GB/s")
    out["vertical_edge_tol"] = ROUGH_TRI_TOL
    out["kv"] = kv
    if cycle_years > 0:
        cyc = {}
        for name in precisions.split(","):
            fs.set_state(X0)
            t0 = time.perf_counter()
            st = fs.solve(cycle_years * SEC_PER_YEAR, rtol=rtol, prec=lib.PREC_BY_NAME[name],
                          steps_per_launch=64)
            wall = time.perf_counter() - t0
            ev = fs.events()
            on = [float(t) / SEC_PER_YEAR for t in (ev[ev[:, 1] > 0, 0] if len(ev) else [])]
            cyc[name] = dict(wall_s=wall, steps=st["steps"], rejects=st["rejects"],
                             onsets_yr=on, sim_years_per_hour=cycle_years / wall * 3600.0)
            log(f"rough cycle {name}: {cycle_years:g} yr in {wall:.2f} s "
                f"({cycle_years / wall * 3600:.3e} sim-yr/h), steps {st['steps']} rejects "
                f"{st['rejects']}, onsets {[round(t, 4) for t in on]} yr")
        out["cycles"] = dict(years=cycle_years, rtol=rtol, runs=cyc)
    fs.close()
    return out


if __name__ == "__main__":
    main()
