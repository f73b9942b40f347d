"""Build the native parts of seisacomod2d_amd in-tree for gfx950.

    python -m seisacomod2d_amd.build        (from the repository root)

Produces
  seisacomod2d_amd/libsam2d.so             the C ABI of include/sam2d.h (HIP kernels + host)
  seisacomod2d_amd/libsam2d_host.so        include/sam2d_modelling.h: job card, geometry,
                                           model files, modelling_module drop-in
  seisacomod2d_amd/SeisAcoMod2D_MI355X.exe the job-card driver (reference CLI equivalent)
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(CSRC, "host")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are kept out of the device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]

# The multi-step forward kernel is instantiated per depth in translation units
# of its own (sam2d_kernelsn_t*.hip), which compile in parallel.
LIB_SOURCES = ["sam2d_kernelsn_t3.hip", "sam2d_kernelsn_t3_ring.hip", "sam2d_kernelsn_t4.hip",
               "sam2d_kernelsn_t5.hip", "sam2d_kernelsn_t2.hip", "sam2d_adjoint_kernels.hip",
               "sam2d_kernels.hip", "sam2d_kernelsn.hip", "sam2d_noalloc.hip", "sam2d_api.cpp",
               "sam2d_adjoint.cpp", "sam2d_setup.cpp"]
HOST_SOURCES = ["jobcard.cpp", "geometry.cpp", "model_io.cpp", "modelling_module.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def _newer(target, sources):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(s) > t for s in sources)


def build(force=True, jobs=8):
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    headers.append(os.path.join(ROOT, "include", "sam2d.h"))
    objs, procs = [], []
    for src in LIB_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        if force or _newer(obj, [path] + headers):
            cmd = [HIPCC, "--offload-arch=" + ARCH] + COMMON + DEVICE_FLAGS + ["-c", path, "-o", obj]
            if src.endswith(".cpp"):
                cmd = [HIPCC] + COMMON + ["-c", path, "-o", obj]
            print("+", " ".join(cmd), flush=True)
            procs.append(subprocess.Popen(cmd))
            if len(procs) >= jobs:
                for p in procs:
                    if p.wait() != 0:
                        raise RuntimeError("compile failed")
                procs = []
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("compile failed")
    lib = os.path.join(PKG, "libsam2d.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, lib)

    host_src = [os.path.join(HOST, s) for s in HOST_SOURCES]
    hlib = os.path.join(PKG, "libsam2d_host.so")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-fPIC", "-shared", "-I" + os.path.join(ROOT, "include"),
          "-o", hlib + ".tmp"] + host_src + ["-L" + PKG, "-lsam2d", "-Wl,-rpath,$ORIGIN"])
    os.replace(hlib + ".tmp", hlib)
    exe = os.path.join(PKG, "SeisAcoMod2D_MI355X.exe")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"), "-o", exe,
          os.path.join(HOST, "main.cpp"), "-L" + PKG, "-lsam2d_host", "-lsam2d",
          "-Wl,-rpath,$ORIGIN"])
    return lib


def build_oracle():
    """Compile the CPU checker (oracle/) and, where the reference tree is
    present, the reference executable and the reference front end linked
    against libsam2d_host.so into oracle/_ref/."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "liboracle.so"])
    if os.path.isdir("/root/reference/src/src_cpu") and shutil.which("make"):
        _run(["make", "-C", odir, "ref"])


if __name__ == "__main__":
    build(force="--incremental" not in sys.argv)
    if "--no-oracle" not in sys.argv:
        build_oracle()
