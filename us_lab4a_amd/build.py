"""Build the mista HIP library (gfx950) and the C++ host driver in-tree.

    python -m us_lab4a_amd.build          # library + host driver
    python -m us_lab4a_amd.build --clean

The library is a single shared object, us_lab4a_amd/libmista_hip.so, built by
hipcc for gfx950 only.  Packed-FP32 instruction selection is disabled for the
device code (-Xclang -target-feature -Xclang -packed-fp32-ops); hipcc's host
pass warns that it does not know the feature, which is harmless.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libmista_hip.so")
HOST_DRIVER = os.path.join(PKG, "mista_sta_cli")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["api.cpp", "prep.hip", "fft4.hip", "das.hip", "das_tile_f32.hip", "das_tile_f16.hip", "das_tile_bf16.hip", "das_tile_rf.hip",
           "synth.hip", "serve.cpp"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    # keep genuine warnings visible, drop the known host-pass feature notice
    noise = "'-packed-fp32-ops' is not a recognized feature for this target"
    for line in (r.stdout + r.stderr).splitlines():
        if noise not in line and line.strip():
            sys.stderr.write(line + "\n")


def build_library(verbose=True):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    build_dir = os.path.join(PKG, "_build")
    os.makedirs(build_dir, exist_ok=True)
    objs, cmds = [], []
    for src in srcs:
        obj = os.path.join(build_dir, os.path.basename(src) + ".o")
        lang = ["-x", "hip"]
        cmds.append([HIPCC] + HIP_FLAGS + lang + ["-I", os.path.join(ROOT, "include"), "-c", src, "-o", obj])
        objs.append(obj)
        if verbose:
            print("[build]", os.path.relpath(src, ROOT), flush=True)
    # translation units compile in parallel (the tiled DAS kernels dominate)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0")) or os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


def build_host_driver(verbose=True):
    """C++ host program above the C ABI (mirror of STAMethod's call sequence)."""
    src = os.path.join(PKG, "host", "sta_cli.cpp")
    if not os.path.exists(src):
        return None
    if verbose:
        print("[build]", os.path.relpath(src, ROOT), flush=True)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(PKG, "include"), src, "-o", HOST_DRIVER,
           "-L", PKG, "-lmista_hip", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN"]
    # HDF5 for the reference's acquisition/image files.  The conda HDF5 is
    # reached through a private directory holding only its library, so the
    # driver never picks up conda's other runtime libraries.
    h5_inc, h5_lib = "/opt/conda/include/hdf5.h", "/opt/conda/lib/libhdf5.so.103"
    if os.path.exists(h5_inc) and os.path.exists(h5_lib):
        libdir = os.path.join(PKG, "host", "lib")
        os.makedirs(libdir, exist_ok=True)
        link = os.path.join(libdir, "libhdf5.so.103")
        if os.path.islink(link) or os.path.exists(link):
            os.remove(link)
        os.symlink(h5_lib, link)
        cmd += ["-DMISTA_HAVE_HDF5", "-I", os.path.dirname(h5_inc), link, "-Wl,-rpath,$ORIGIN/host/lib"]
    _run(cmd)
    return HOST_DRIVER


def build_oracle(verbose=True):
    """Test-infrastructure oracle (CPU), oracle/liboracle.so."""
    odir = os.path.join(ROOT, "oracle")
    if verbose:
        print("[build] oracle/liboracle.so", flush=True)
    _run(["make", "-s", "-C", odir, "-B", "liboracle.so"])


def clean():
    shutil.rmtree(os.path.join(PKG, "_build"), ignore_errors=True)
    for p in (LIB, HOST_DRIVER):
        if os.path.exists(p):
            os.remove(p)


if __name__ == "__main__":
    if "--clean" in sys.argv:
        clean()
    else:
        build_library()
        build_host_driver()
        build_oracle()
