"""Build libppmg.so (HIP, gfx950) in-tree: ppm_legacy_amd/libppmg.so.

    python -m ppm_legacy_amd.build            # compile every csrc/*.hip, link
    python -m ppm_legacy_amd.build --fortran  # also the Fortran modules (flang)

Every HIP translation unit is compiled with packed-FP32 instruction
selection disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libppmg.so")
ARCH = os.environ.get("PPMG_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


HIP_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unused-command-line-argument",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [hipcc(), *HIP_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    warn = [l for l in r.stderr.splitlines() if "warning" in l and "packed-fp32-ops" not in l]
    return obj, warn


def build(verbose=True, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, warn in ex.map(_compile, srcs):
            objs.append(obj)
            if verbose:
                for w in warn:
                    print(w, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_fortran(verbose=True):
    """Compile the Fortran drop-in modules and their test driver with flang."""
    fdir = os.path.join(HERE, "fortran")
    flang = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
    if not os.path.exists(flang):
        raise RuntimeError("flang not found")
    out = os.path.join(BUILD, "fortran")
    os.makedirs(out, exist_ok=True)
    order = ["ppm_module_ppmg.f90", "ppm_module_core.f90", "ppm_module_interp_p2m.f90",
             "ppm_module_interp_m2p.f90", "ppm_module_rmsh.f90", "ppm_module_util_rank.f90",
             "ppm_module_map_field_ghost.f90", "ppm_module_util_sort.f90", "ppm_module_neighlist.f90"]
    srcs = [os.path.join(fdir, f) for f in order]
    objs = []
    for s in srcs:
        o = os.path.join(out, os.path.basename(s) + ".o")
        cmd = [flang, "-O2", "-fPIC", "-J", out, "-c", s, "-o", o]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed for {s}:\n{r.stdout}\n{r.stderr}")
        objs.append(o)
    lib = os.path.join(HERE, "libppm_fortran.a")
    if os.path.exists(lib):
        os.remove(lib)
    subprocess.run(["ar", "rcs", lib, *objs], check=True)
    # the test client of the drop-in (tests/fortran), linked against it
    drv = os.path.join(ROOT, "tests", "fortran", "ppmg_fortran_check.f90")
    if os.path.exists(drv):
        exe = os.path.join(ROOT, "tests", "fortran", "ppmg_fortran_check")
        cmd = [flang, "-O2", "-I", out, drv, lib, "-L", HERE, "-lppmg",
               "-Wl,-rpath,$ORIGIN/../../ppm_legacy_amd", "-Wl,-rpath,/opt/rocm/lib", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang link failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {lib}")
    return lib


def build_c_client(verbose=True):
    """Compile the C client of include/ppmg.h (tests/c_client) with gcc
    against the built library."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "c_client", "ppmg_c_check.c")
    if not os.path.exists(src):
        return None
    exe = os.path.join(root, "tests", "c_client", "ppmg_c_check")
    cmd = ["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", "-I", os.path.join(root, "include"), src,
           "-L", HERE, "-lppmg", "-lm", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/../../ppm_legacy_amd",
           "-Wl,-rpath,/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 or r.stderr.strip():
        raise RuntimeError(f"C client build failed or warned:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {exe}")
    return exe


if __name__ == "__main__":
    build()
    if "--fortran" in sys.argv:
        build_fortran()
