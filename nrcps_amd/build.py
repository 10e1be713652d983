"""Builds libnrcps_hip.so (the HIP product library, gfx950) in-tree.

    python -m nrcps_amd.build          # or nrcps_amd.build.build()
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libnrcps_hip.so")
LIB_CHECK = os.path.join(PKG, "libnrcps_hip_check.so")  # the -DNRC_CHECK build of the ABI
SOURCES = ["fft_passes.hip", "fft_axis.hip", "slice_fused.hip", "slice_ext.hip", "lu_rng.hip", "overlap_mfma.hip", "proposal.hip",
           "api.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NRC_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled for all device code of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed: " + " ".join(cmd))
    return r


HOST_LIB = os.path.join(PKG, "libnrcps.so")
HOST_SOURCES = ["host/util.cpp", "host/physics.cpp"]
DRIVER = os.path.join(PKG, "drivers", "slater_det")
DRIVER3 = os.path.join(PKG, "drivers", "slater_det3")
DRIVER2 = os.path.join(PKG, "drivers", "slater_det2")
DRIVER_RW = os.path.join(PKG, "drivers", "one_body_reweight")
CXX = os.environ.get("CXX", "g++")
CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
            f"-I{os.path.join(ROOT, 'include')}"]


def build(jobs: int = 4) -> str:
    """Compiles every HIP source to an object (in parallel) and links libnrcps_hip.so, then
    the host class layer libnrcps.so and the slater_det driver on top of it."""
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
        if len(procs) >= jobs:
            _wait(procs)
    # the check build: api.hip again with -DNRC_CHECK (the kernels are the same objects)
    api_check = os.path.join(CSRC, "api_check.o")
    cmd = [HIPCC, *FLAGS, "-DNRC_CHECK", "-c", os.path.join(CSRC, "api.hip"), "-o", api_check]
    procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                        text=True)))
    _wait(procs)
    tmp = LIB + ".tmp"
    _run([HIPCC, *FLAGS, "-shared", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    tmp = LIB_CHECK + ".tmp"
    _run([HIPCC, *FLAGS, "-shared", "-o", tmp,
          *[o for o in objs if not o.endswith("api.o")], api_check])
    os.replace(tmp, LIB_CHECK)
    build_host()
    return LIB


def build_host() -> None:
    """The C++ class layer (reference class API over the C ABI) and its driver."""
    tmp = HOST_LIB + ".tmp"
    _run([CXX, *CXXFLAGS, "-shared", "-o", tmp, *[os.path.join(PKG, s) for s in HOST_SOURCES],
          f"-L{PKG}", "-lnrcps_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_LIB)
    tmp = DRIVER + ".tmp"
    _run([CXX, *CXXFLAGS, f"-I{os.path.join(ROOT, 'include', 'nrcps')}", "-o", tmp,
          os.path.join(PKG, "drivers", "slater_det.cpp"), f"-L{PKG}", "-lnrcps", "-lnrcps_hip",
          "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, DRIVER)
    for drv, src in ((DRIVER3, "slater_det3.cpp"), (DRIVER2, "slater_det2.cpp"),
                     (DRIVER_RW, "one_body_reweight.cpp")):
        tmp = drv + ".tmp"
        _run([CXX, *CXXFLAGS, f"-I{os.path.join(ROOT, 'include', 'nrcps')}", "-o", tmp,
              os.path.join(PKG, "drivers", src), f"-L{PKG}", "-lnrcps", "-lnrcps_hip",
              "-Wl,-rpath,$ORIGIN/.."])
        os.replace(tmp, drv)


def _wait(procs):
    while procs:
        cmd, p = procs.pop(0)
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + err)
            raise RuntimeError("hipcc failed: " + " ".join(cmd))


if __name__ == "__main__":
    print(build())
