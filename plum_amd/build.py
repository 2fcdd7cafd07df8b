"""Compiles the HIP engine (gfx950) into plum_amd/libplum_gpu.so and the C++
host layer into plum_amd/libplum_host.so.  Always recompiles."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
ARCH = os.environ.get("PLUM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: never emit v_pk_{add,mul,fma}_f32
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-I", INC, "-I", CSRC]

ENGINE_SRC = ["engine.hip"]
HOST_SRC = ["host/runin.cc", "host/force_field.cc", "host/cbmc.cc", "host/simulation.cc",
            "host/host_abi.cc", "host/synth.cc", "host/slit_pressure.cc"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_engine():
    out = os.path.join(HERE, "libplum_gpu.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, s) for s in ENGINE_SRC]
    run([HIPCC] + HIP_FLAGS + ["-shared", "-o", tmp] + srcs)
    os.replace(tmp, out)
    return out


def build_host():
    srcs = [os.path.join(CSRC, s) for s in HOST_SRC]
    if not all(os.path.exists(s) for s in srcs):
        return None
    out = os.path.join(HERE, "libplum_host.so")
    tmp = out + ".tmp"
    run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", INC,
         "-I", os.path.join(CSRC, "host"), "-o", tmp] + srcs +
        ["-L" + HERE, "-lplum_gpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def main():
    build_engine()
    build_host()


if __name__ == "__main__":
    sys.exit(main())
