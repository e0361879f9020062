"""Build recipes for the native parts (run from __graft_entry__.build()).

  * libbiefhip.so   -- HIP kernels + C ABI, gfx950 only (hipcc)
  * libbiefmesh.so  -- host-only synthetic mesh helper (g++)
  * libbiefshim.so  -- the Fortran BIEF-signature host layer over the C ABI
                       (flang, ISO_C_BINDING), when flang is present
Everything is built in-tree so the libraries travel with the sources.
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 instructions are kept out of the device code
HIP_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed ({' '.join(cmd)}):\n{r.stdout[-6000:]}")
    return r.stdout


def build_hip(verbose=False):
    out = os.path.join(HERE, "libbiefhip.so")
    srcs = [os.path.join(CSRC, f) for f in ("bief_hip.hip", "layout.cpp")]
    tmp = out + ".tmp"
    log = _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), *srcs, "-o", tmp])
    os.replace(tmp, out)
    if verbose:
        print(log)
    return out


def build_mesh():
    out = os.path.join(HERE, "libbiefmesh.so")
    _run(["g++", "-O2", "-fPIC", "-shared", os.path.join(CSRC, "meshgen.cpp"), "-o", out])
    return out


def have_flang():
    return shutil.which("flang") is not None or os.path.exists("/opt/rocm/bin/amdflang")


def _fc():
    return shutil.which("flang") or "/opt/rocm/bin/amdflang"


def build_shim():
    """Fortran host layer with BIEF's SOLVE/MATVEC/MATRBL signatures over the
    C ABI (ISO_C_BINDING), stand-alone types: libbiefshim.so."""
    if not have_flang():
        return None
    out = os.path.join(HERE, "libbiefshim.so")
    mod = os.path.join(HERE, "_build", "fmod")
    os.makedirs(mod, exist_ok=True)
    srcs = [os.path.join(FSRC, f) for f in ("bief_hip_types.f90", "bief_hip_capi.f90",
                                            "bief_hip_shim.F90")]
    _run([_fc(), "-O2", "-cpp", "-DBHIP_OWN_TYPES", "-fPIC", "-shared", "-module-dir", mod,
          *srcs, "-L", HERE, "-lbiefhip", "-Wl,-rpath,$ORIGIN", "-o", out])
    return out


def build_dropin():
    """The same host layer compiled against the reference's own BIEF_DEF and
    linked in front of the reference BIEF archive (oracle/_ref), so that the
    reference's callers reach libbiefhip through SOLVE/MATVEC: the drop-in
    test library oracle/_ref/libbief_dropin.so (test infrastructure)."""
    ref = os.path.join(ROOT, "oracle", "_ref")
    if not (have_flang() and os.path.exists(os.path.join(ref, "libbief_ref.a"))):
        return None
    out = os.path.join(ref, "libbief_dropin.so")
    mod = os.path.join(ref, "dropin_mod")
    os.makedirs(mod, exist_ok=True)
    srcs = [os.path.join(FSRC, "bief_hip_capi.f90"), os.path.join(FSRC, "bief_hip_shim.F90"),
            os.path.join(ROOT, "oracle", "ref_driver.f90")]
    _run([_fc(), "-O2", "-cpp", "-ffixed-line-length-132", "-fPIC", "-shared",
          "-I", os.path.join(ref, "mod"), "-module-dir", mod, *srcs,
          os.path.join(ref, "libbief_ref.a"), "-L", HERE, "-lbiefhip",
          "-Wl,-rpath,$ORIGIN/../../opentelemac_amd", "-o", out])
    return out


def build_oracle():
    _run(["make", "-C", os.path.join(ROOT, "oracle")])
    if os.path.isdir("/root/reference/sources/utils/bief") and have_flang():
        _run(["make", "-C", ROOT, "-f", "oracle/ref.mk", "-j8"])


def build_all(verbose=False):
    build_mesh()
    build_hip(verbose)
    build_shim()
    build_oracle()
    build_dropin()
