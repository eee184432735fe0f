"""Build librsfg.so for gfx950 in-tree with hipcc (no JIT, no cache dirs).

    python -m interact_amd.build          # compile interact_amd/librsfg.so
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["assemble.hip", "solver.hip", "ensemble.hip", "adjoint.hip", "api.hip"]
LIB = os.path.join(HERE, "librsfg.so")
ARCH = os.environ.get("RSFG_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    # keep the compiler from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build(verbose=False, jobs=None):
    """Compiles each source to an object in parallel (no kernel is launched
    outside the translation unit that defines it), then links the library."""
    import tempfile
    from concurrent.futures import ThreadPoolExecutor
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    with tempfile.TemporaryDirectory(prefix="rsfg_build_") as d:
        objs = [os.path.join(d, os.path.basename(s) + ".o") for s in srcs]
        cflags = [f for f in HIPCC_FLAGS if f != "-shared"]
        cmds = [[hipcc()] + cflags + ["-c", s, "-o", o] for s, o in zip(srcs, objs)]
        if verbose:
            for c in cmds:
                print(" ".join(c), flush=True)
        jobs = jobs or min(len(cmds), os.cpu_count() or 1)
        with ThreadPoolExecutor(jobs) as ex:
            res = list(ex.map(lambda c: subprocess.run(c, capture_output=True, text=True), cmds))
        for c, r in zip(cmds, res):
            if r.stderr and (verbose or r.returncode):
                sys.stderr.write("".join(l + "\n" for l in r.stderr.splitlines()
                                         if "packed-fp32-ops" not in l))
            if r.returncode:
                raise subprocess.CalledProcessError(r.returncode, c)
        link = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}"] + objs + ["-o", tmp]
        if verbose:
            print(" ".join(link), flush=True)
        subprocess.run(link, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
    sys.exit(0)
