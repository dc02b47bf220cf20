"""Builds the native libraries of the package, in-tree.

  libsw_gpu.so   HIP kernels + C ABI (include/sw_gpu.h) + frame-serving scheduler
                 (include/sw_server.h), gfx950 code objects, built by hipcc
  libsw_host.so  C++ renderer front-end (map/texture/lightmap generation, visibility,
                 per-polygon surface and draw setup) that emits the command buffers

Flags that matter for parity with the reference's numerics:
  -ffp-contract=off                          no implicit FMA (Rust never contracts)
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE f32 divide / sqrt
  -fno-gpu-flush-denormals-to-zero           x86 keeps f32 denormals
  -fwrapv                                    Rust release-build integer wrap-around
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed-FP32 VALU forms
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)

GPU_LIB = os.path.join(HERE, "libsw_gpu.so")
HOST_LIB = os.path.join(HERE, "libsw_host.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SW_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fwrapv",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fwrapv", "-Wall",
              "-Wno-unused-function", "-pthread"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_gpu():
    srcs = [os.path.join(CSRC, "sw_kernels.hip"), os.path.join(CSRC, "sw_runtime.hip"), os.path.join(CSRC, "sw_models.hip"), os.path.join(CSRC, "sw_textures.hip"), os.path.join(CSRC, "sw_depth.hip"),
            os.path.join(CSRC, "sw_server.cpp")]
    tmp = GPU_LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, *srcs])
    os.replace(tmp, GPU_LIB)


def build_host():
    if not os.path.isdir(os.path.join(CSRC, "host")):
        return
    srcs = sorted(os.path.join(CSRC, "host", f) for f in os.listdir(os.path.join(CSRC, "host")) if f.endswith(".cpp"))
    tmp = HOST_LIB + ".tmp"
    _run(["g++", *HOST_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, *srcs])
    os.replace(tmp, HOST_LIB)


def build_all():
    build_host()
    build_gpu()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "gpu": build_gpu, "host": build_host}[w]()
