"""Build recipes for libvoxelcore.so (gfx950) and the CPU oracle.

Everything is compiled in-tree so the built files travel with the source tree.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libvoxelcore.so")

HIP_SOURCES = ["vc_core.hip", "vc_sched.cpp", "vc_edt.hip", "vc_thin.hip", "vc_batch.cpp",
               "vc_boundary.hip", "vc_synth.hip", "vc_ccthin.cpp", "vc_unet.hip",
               "vc_voro.hip", "vc_voro.cpp", "vc_octree.hip", "vc_export.cpp",
               "vc_ccthin_dev.hip", "vc_voro_merge.hip", "vc_noalloc.hip", "vc_vdiag.hip"]
HOST = os.path.join(PKG, "host")
# the drop-in C++ layer (reference class names) linked into the same library
HOST_SOURCES = ["vc_device.cpp", "geomalgo.cpp", "cellcomplex.cpp", "ccthin.cpp",
                "surfacing.cpp", "voxelthin.cpp", "highlevelalgo.cpp", "voroinfo.cpp", "octree.cpp",
                "exporters.cpp"]
CLI = os.path.join(PKG, "voroUtility")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Wall", "-Wno-unused-parameter", "-Wno-unused-result",
    # no packed FP32 instructions: some of their forms are unreliable here
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r.stdout


def build_library(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    srcs += [os.path.join(HOST, s) for s in HOST_SOURCES]
    tmp = LIB + ".tmp"
    out = _run([HIPCC] + HIP_FLAGS + srcs + ["-o", tmp])
    os.replace(tmp, LIB)
    if verbose and out.strip():
        print(out)
    build_cli()
    return LIB


# The host-only sources (no device code): also built with AddressSanitizer
# and UndefinedBehaviorSanitizer into a library of their entry points, which
# the sanitizer test loads in place of libvoxelcore.so (VC_LIBRARY) with the
# ASan runtime preloaded
HOST_ONLY_SOURCES = ["vc_ccthin.cpp", "vc_voro.cpp", "vc_export.cpp"]
ASAN_DIR = os.path.join(PKG, "_asan")
ASAN_LIB = os.path.join(ASAN_DIR, "libvchost_asan.so")


def build_host_sanitized():
    os.makedirs(ASAN_DIR, exist_ok=True)
    tmp = ASAN_LIB + ".tmp"
    _run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fPIC", "-shared",
          "-fsanitize=address,undefined", "-Wall", "-D__HIP_PLATFORM_AMD__",
          "-I", "/opt/rocm/include", "-I", os.path.join(ROOT, "include")]
         + [os.path.join(CSRC, s) for s in HOST_ONLY_SOURCES] + ["-o", tmp, "-lpthread"])
    os.replace(tmp, ASAN_LIB)
    return ASAN_LIB


def build_cli():
    """voroUtility, linked against libvoxelcore.so (rpath = its directory)."""
    tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
          "-I", "/opt/rocm/include", "-I", os.path.join(ROOT, "include"),
          os.path.join(HOST, "voroUtility.cpp"), "-o", tmp,
          "-L", PKG, "-lvoxelcore", "-Wl,-rpath,$ORIGIN",
          "-L", "/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, CLI)
    return CLI


def build_oracle():
    _run(["make", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "_build", "liboracle.so")


def build_oracle_sanitized():
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize"])
    return os.path.join(ROOT, "oracle", "_build", "asan", "liboracle.so")


if __name__ == "__main__":
    build_oracle()
    print(build_library(verbose=True))
