"""In-tree build of the native libraries (no JIT cache, no site-packages).

Outputs under gpuattract_2.0_amd/lib/:
  libattract_gen.so   seeded workload generator (host C++)
  libattract_hip.so   HIP kernels for gfx950 + the C ABI of include/attract_hip.h
  libattract_service.so / service_test
                      C++ Server/Service drop-in (MI355EnergyService6D) over the C ABI
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "lib")
CSRC = os.path.join(PKG, "csrc")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed-FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off", "-Wall", "-Wno-unused-function",
    "-Wno-unknown-warning-option",
]


def _run(cmd):
    r = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r.stdout


def build_gen():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libattract_gen.so")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", INC,
          os.path.join(CSRC, "gen", "attract_gen.cpp"), "-o", out, "-pthread"])
    return out


def hip_sources():
    d = os.path.join(CSRC, "hip")
    return sorted(os.path.join(d, f) for f in os.listdir(d) if f.endswith((".hip", ".cpp")))


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libattract_hip.so")
    srcs = hip_sources()
    if not srcs:
        return None
    objs = []
    for s in srcs:
        o = os.path.join(LIB, os.path.basename(s) + ".o")
        _run([HIPCC] + HIP_FLAGS + ["-I", INC, "-c", s, "-o", o])
        objs.append(o)
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", out] + objs)
    for o in objs:
        os.remove(o)
    return out


def build_service():
    """The C++ drop-in layer (csrc/service -> lib/libattract_service.so) and its
    test programs (tests/cpp -> tests/cpp/_bin)."""
    d = os.path.join(CSRC, "service")
    srcs = sorted(os.path.join(d, f) for f in os.listdir(d) if f.endswith(".cpp"))
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libattract_service.so")
    cxx = ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-I", INC, "-pthread"]
    _run(cxx + ["-shared"] + srcs + ["-o", out, "-L", LIB, "-lattract_hip", "-Wl,-rpath,$ORIGIN"])
    tdir = os.path.join(ROOT, "tests", "cpp")
    if os.path.isdir(tdir):
        bin_dir = os.path.join(tdir, "_bin")
        os.makedirs(bin_dir, exist_ok=True)
        # relative run path: the tree is relocated on other machines
        link = ["-L", LIB, "-lattract_service", "-lattract_hip",
                "-Wl,-rpath,$ORIGIN/../../../gpuattract_2.0_amd/lib"]
        _run(cxx + ["-I", d, os.path.join(tdir, "server_semantics.cpp"), "-o",
                    os.path.join(bin_dir, "server_semantics")])
        _run(cxx + ["-shared", os.path.join(tdir, "service_driver.cpp"), "-o",
                    os.path.join(bin_dir, "libservice_driver.so")] + link)
        # the device minimizer's state machine, compiled for the host
        _run([HIPCC] + HIP_FLAGS + ["-shared", "-I", os.path.join(CSRC, "hip"),
                                    os.path.join(tdir, "va13_host.cpp"), "-o",
                                    os.path.join(bin_dir, "libva13_host.so")])
        # the device sine/cosine, compiled for the host beside the C library
        _run([HIPCC] + HIP_FLAGS + ["-shared", "-I", os.path.join(CSRC, "hip"),
                                    os.path.join(tdir, "trig_host.cpp"), "-o",
                                    os.path.join(bin_dir, "libtrig_host.so"), "-pthread"])
    return out


def build_oracle():
    """Test-only checkers (oracle/liboracle.so and, with the reference tree, oracle/_ref)."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"])


def build_all():
    build_gen()
    build_hip()
    build_service()
    build_oracle()


if __name__ == "__main__":
    build_all()
