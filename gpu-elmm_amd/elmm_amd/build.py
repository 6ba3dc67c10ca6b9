"""Compile the gfx950 HIP library (libelmm_hip.so) in-tree.

Sources: gpu-elmm_amd/csrc/*.hip (device kernels + engine + C ABI) and
gpu-elmm_amd/csrc/*.cpp (host-only case set-up).  Output:
gpu-elmm_amd/lib/libelmm_hip.so.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")
LIBNAME = "libelmm_hip.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off keeps every multiply and add separately rounded (the
# reference's arithmetic; *_f32.hip files, FP32 kernels of the precision form,
# use -ffp-contract=fast); packed FP32 instructions are disabled.
HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]
CXX_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall"]


def lib_path():
    return os.path.join(LIBDIR, LIBNAME)


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False, jobs=None):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    inc = ["-I", INCLUDE, "-I", CSRC]
    jobs_list = []
    for f in sorted(os.listdir(CSRC)):
        src = os.path.join(CSRC, f)
        obj = os.path.join(OBJDIR, f + ".o")
        if f.endswith(".hip"):
            flags = HIP_FLAGS
            if f.endswith("_f32.hip"):
                # FP32 instantiations of the precision form: multiply-add
                # pairs may fuse (the FP64 form keeps every operation rounded)
                flags = [x if x != "-ffp-contract=off" else "-ffp-contract=fast" for x in HIP_FLAGS]
            jobs_list.append([HIPCC] + flags + inc + ["-c", src, "-o", obj])
        elif f.endswith(".cpp"):
            jobs_list.append(["g++"] + CXX_FLAGS + inc + ["-c", src, "-o", obj])
    n = jobs or min(8, len(jobs_list))
    with ThreadPoolExecutor(max_workers=n) as ex:
        outs = list(ex.map(_run, jobs_list))
    if verbose:
        for o in outs:
            if o.strip():
                sys.stdout.write(o)
    objs = [c[-1] for c in jobs_list]
    tmp = lib_path() + ".tmp"
    # --no-undefined: a missing definition fails the build instead of the load
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-Wl,--no-undefined", "-o", tmp] + objs)
    os.replace(tmp, lib_path())
    build_fortran_host()
    shutil.rmtree(OBJDIR, ignore_errors=True)
    return lib_path()


FORTRAN = os.path.join(PKG, "fortran")
FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")


def host_path():
    return os.path.join(LIBDIR, "elmm_host")


def build_fortran_host():
    """The iso_c_binding module and the namelist host (gpu-elmm_amd/fortran),
    linked against the library; skipped when amdflang is absent."""
    if not os.path.exists(FC):
        return None
    mod = os.path.join(OBJDIR, "fmod")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for f in ("elmm_gpu.f90", "elmm_host.f90"):
        obj = os.path.join(OBJDIR, f + ".o")
        _run([FC, "-O2", "-module-dir", mod, "-I", mod, "-c", os.path.join(FORTRAN, f), "-o", obj])
        objs.append(obj)
    tmp = host_path() + ".tmp"
    _run([FC, "-o", tmp] + objs + ["-L", LIBDIR, "-lelmm_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, host_path())
    return host_path()


if __name__ == "__main__":
    print(build(verbose=True))
