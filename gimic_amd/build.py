"""Build libgimic_hip.so (HIP kernels for gfx950 + C-ABI) in-tree with hipcc."""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libgimic_hip.so")
ARCH = os.environ.get("GIMIC_OFFLOAD_ARCH", "gfx950")

SOURCES = ["basis_eval.hip", "contract_f64.hip", "contract_mx2.hip", "finalize.hip", "selftest.hip",
           "surrogate.hip", "tiles.hip", "planes.hip", "refine.hip",
           "api.cpp", "surrogate_api.cpp", "host_basis.cpp", "quadrature.cpp"]

# Packed FP32 VALU ops stay disabled in device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          f"--offload-arch={ARCH}", "-I", CSRC, "-I", os.path.join(HERE, "..", "include")]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(OBJ, src + ".o")
    lang = [] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [hipcc()] + lang + COMMON + DEVICE_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in (res.stderr or "").splitlines()
                    if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    return obj, err


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        results = list(ex.map(_compile, SOURCES))
    for _, err in results:
        if err.strip() and verbose:
            print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + \
        [o for o, _ in results]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n" + res.stderr)
    os.replace(tmp, LIB)
    return LIB


FLANG = "/opt/rocm/lib/llvm/bin/flang"
FDIR = os.path.join(HERE, "fortran")


def build_fortran():
    """Fortran iso_c_binding module and demo caller (skipped without flang)."""
    if not os.path.exists(FLANG):
        return None
    out = os.path.join(FDIR, "fgimic_hip_demo")
    cmd = [FLANG, "-O2", "-module-dir", OBJ, os.path.join(FDIR, "gimic_hip.f90"),
           os.path.join(FDIR, "fgimic_hip_demo.f90"), "-L", HERE, "-lgimic_hip",
           "-Wl,-rpath,$ORIGIN/..", "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("flang failed:\n" + res.stderr)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran())
