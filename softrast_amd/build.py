"""Build the native libraries in-tree.

libsrgpu.so   HIP kernels + C ABI (include/srgpu.h), compiled for gfx950 only.
libsrscene.so host C++ scene / texture generators used for synthetic inputs.

Device code is compiled with IEEE fp32 semantics: no FMA contraction
(-ffp-contract=off; explicit fmaf is kept), correctly rounded divide and
sqrt, denormals preserved, and packed-FP32 instructions disabled.
"""
import os
import pathlib
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # atomics on one address are aggregated over the wave by hand where it
    # pays; the compiler's own rewrite adds a wave scan and a broadcast that
    # waits for the atomic's return at once
    "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def _run(cmd):
    res = subprocess.run(cmd, cwd=str(ROOT), capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    return res


def build_gpu(extra=()):
    out = PKG / "libsrgpu.so"
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, *extra, "-o", str(tmp), str(CSRC / "srgpu.hip")])
    os.replace(tmp, out)
    return out


def build_scene():
    out = PKG / "libsrscene.so"
    tmp = out.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o", str(tmp), str(CSRC / "scene.cpp")])
    os.replace(tmp, out)
    return out


def build_copy_calib():
    """Stand-alone copy kernel of known size (counter calibration for profiles)."""
    out = PKG / "copy_calib"
    _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-o", str(out), str(CSRC / "copy_calib.hip")])
    return out


def build_all():
    build_scene()
    build_gpu()
    build_copy_calib()


if __name__ == "__main__":
    build_all()
    print("built", PKG / "libsrgpu.so", PKG / "libsrscene.so")
