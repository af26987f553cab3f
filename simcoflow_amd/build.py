"""Build the gfx950 shared library libsimcogpu.so in-tree with hipcc.

    python -m simcoflow_amd.build

Every HIP source is compiled for gfx950 only, with packed-FP32 instruction
selection disabled; the projection-coefficient, interface-transport and API
files are compiled with FP contraction off so their arithmetic rounds
exactly as written.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libsimcogpu.so")
ARCH = "gfx950"

SOURCES = [
    "sg_api.hip",
    "sg_poisson.hip",
    "sg_mg.hip",
    "sg_pcg.hip",
    "sg_cycle.hip",
    "sg_clsvof.hip",
    "sg_predictor.hip",
    "sg_momentum.hip",
    "sg_nn.hip",
    "sg_learned.hip",
    "sg_step.hip",
    "sg_particles.hip",
    "sg_state.hip",
    "sg_geometry.hip",
]
EXTRA = {
    "sg_poisson.hip": ["-ffp-contract=off"],
    "sg_clsvof.hip": ["-ffp-contract=off"],
    "sg_predictor.hip": ["-ffp-contract=off"],
    "sg_api.hip": ["-ffp-contract=off"],
    "sg_step.hip": ["-ffp-contract=off"],
    "sg_particles.hip": ["-ffp-contract=off"],
    "sg_geometry.hip": ["-ffp-contract=off"],
    # FP32 atomic adds of the weight gradients as hardware atomics (device
    # memory, coarse-grained)
    "sg_nn.hip": ["-munsafe-fp-atomics"],
}
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def _compile(src):
    obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + COMMON + EXTRA.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_fortran(verbose)
    return LIB


FORTRAN_DIR = os.path.join(os.path.dirname(HERE), "fortran")
FORTRAN_OUT = os.path.join(FORTRAN_DIR, "_build")
FORTRAN_TEST = os.path.join(FORTRAN_OUT, "test_projection")
FORTRAN_TESTS = ("test_projection", "test_transport")


def amdflang():
    for cand in (os.environ.get("AMDFLANG"), "/opt/rocm/bin/amdflang", "/opt/rocm/lib/llvm/bin/amdflang"):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran(verbose=False):
    """The Fortran host binding (iso_c_binding module, drop-in bodies) and
    its test driver, linked against libsimcogpu.so; skipped when no Fortran
    compiler is installed."""
    fc = amdflang()
    if fc is None:
        if verbose:
            print("amdflang not found: Fortran host binding not built")
        return None
    os.makedirs(FORTRAN_OUT, exist_ok=True)
    lib_srcs = [os.path.join(FORTRAN_DIR, f) for f in ("simcogpu_mod.f90", "sg_projection_dropin.f90")]
    for t in FORTRAN_TESTS:
        exe = os.path.join(FORTRAN_OUT, t)
        cmd = [fc, "-O2", "-module-dir", FORTRAN_OUT] + lib_srcs + [os.path.join(FORTRAN_DIR, t + ".f90")] + [
            "-L" + OUT_DIR, "-lsimcogpu", "-Wl,-rpath," + OUT_DIR, "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=FORTRAN_OUT)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed:\n{r.stderr}")
        if verbose:
            print("built", exe)
    return FORTRAN_TEST


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
