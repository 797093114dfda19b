"""Build recipes for the in-tree native artefacts (no JIT cache; the .so files ship with the tree).

  libdsx.so            the product: gfx950 HIP kernels + C ABI (hipcc)
  oracle/liboracle.so  CPU oracle for tests and the bench's CPU baseline (g++), test infrastructure
  oracle/_ref/*        the reference's own CPU tools, built from /root/reference when present
"""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no FMA contraction, results bit-identical to the reference's CPU arithmetic.
# -packed-fp32-ops off: no v_pk_{add,mul,fma}_f32 anywhere in the device code.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-shared", "-ffp-contract=off",
             "-fno-gpu-flush-denormals-to-zero", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]
HIP_SOURCES = ["csrc/gate.hip", "csrc/tables.hip", "csrc/climate.hip", "csrc/floods.hip", "csrc/flood_lds.hip",
               "csrc/dsx_api.hip",
               "csrc/scan_api.hip", "csrc/floods_api.hip", "host/search.cpp", "host/serve.cpp"]
HOST_TOOLS = ["host/dsearch_desert.cpp"]


def build_lib(verbose=False):
    out = os.path.join(HERE, "libdsx.so")
    tmp = out + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + [os.path.join(HERE, s) for s in HIP_SOURCES] + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose and r.stderr.strip():
        print(r.stderr)
    os.replace(tmp, out)
    for src in HOST_TOOLS:
        exe = os.path.join(HERE, os.path.splitext(os.path.basename(src))[0])
        r = subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(HERE, src), "-L" + HERE, "-ldsx",
                            "-Wl,-rpath,$ORIGIN", "-o", exe + ".tmp"], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("g++ failed:\n" + r.stdout + r.stderr)
        os.replace(exe + ".tmp", exe)
    return out


def build_oracle():
    subprocess.run(["make", "-s", "-B", "-f", os.path.join(REPO, "oracle", "Makefile")], check=True)
    return os.path.join(REPO, "oracle", "liboracle.so")


def build_reference_tools():
    """oracle/_ref from the reference's sources, only where /root/reference exists."""
    if not os.path.isdir("/root/reference/src"):
        return None
    if shutil.which("g++") is None:
        return None
    subprocess.run(["make", "-s", "-B", "-f", os.path.join(REPO, "oracle", "Makefile.ref")], check=True)
    return os.path.join(REPO, "oracle", "_ref")


def build_all(verbose=False):
    lib = build_lib(verbose)
    orc = build_oracle()
    ref = build_reference_tools()
    return lib, orc, ref


if __name__ == "__main__":
    print(build_all(verbose=True))
