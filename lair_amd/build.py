"""Builds the gfx950 shared library lair_amd/lib/liblair_dev.so with hipcc.

Every HIP source is compiled with packed-FP32 instruction selection disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops): float code stays on scalar
v_fma_f32 / v_mul_f32.  hipcc's host pass warns that it does not know the
feature; only the device pass uses it.
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["kernels.hip", "rpanel.hip", "solve_kernels.hip", "gemm_kernels.hip", "complex.hip", "batched.hip",
           "driver.hip"]
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "liblair_dev.so")
# Input generator for the benchmark and large tests (links rocSOLVER/rocBLAS;
# never loaded by the solver library itself).
GEN_SOURCES = ["gen_inputs.hip"]
GEN_LIB = os.path.join(LIB_DIR, "liblair_gen.so")
ARCH = os.environ.get("LAIR_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def command(out=LIB, extra=()):
    """One hipcc line building the whole library (kept for reference and for
    callers that want a single command); build() compiles the sources in
    parallel instead.  Every translation unit holds its own kernels and their
    launchers, so no device-code linking across units is needed."""
    srcs = [os.path.join(HERE, "csrc", s) for s in SOURCES]
    return [hipcc(), *FLAGS, "-shared", *extra, "-o", out, *srcs]


def gen_command(out=GEN_LIB):
    srcs = [os.path.join(HERE, "csrc", s) for s in GEN_SOURCES]
    return [hipcc(), "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-o", out, *srcs, "-lrocsolver", "-lrocblas"]


def _run(cmd, tmp, final, verbose):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed building {os.path.basename(final)}")
    if verbose:
        sys.stderr.write(res.stderr)
    os.replace(tmp, final)


def _compile(src, obj, verbose):
    res = subprocess.run([hipcc(), *FLAGS, "-c", "-o", obj, src], capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed compiling {os.path.basename(src)}")
    if verbose:
        sys.stderr.write(res.stderr)


def build(verbose=False):
    """Compiles every source of the library in parallel, links it, then builds
    the input-generator library."""
    os.makedirs(LIB_DIR, exist_ok=True)
    jobs = max(1, min(len(SOURCES) + 1, int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with tempfile.TemporaryDirectory(prefix="lair_build_") as tmp:
        srcs = [os.path.join(HERE, "csrc", s) for s in SOURCES]
        objs = [os.path.join(tmp, os.path.splitext(s)[0] + ".o") for s in SOURCES]
        with ThreadPoolExecutor(jobs) as pool:
            gen = pool.submit(_run, gen_command(GEN_LIB + ".tmp"), GEN_LIB + ".tmp", GEN_LIB, verbose)
            for f in [pool.submit(_compile, s, o, verbose) for s, o in zip(srcs, objs)]:
                f.result()
            gen.result()
        _run([hipcc(), *FLAGS, "-shared", "-o", LIB + ".tmp", *objs], LIB + ".tmp", LIB, verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
