"""Build recipes for the in-tree native artefacts (no JIT cache; the .so files ship with the tree).

  libdsx.so            the product: gfx950 HIP kernels + C ABI (hipcc)
  oracle/liboracle.so  CPU oracle for tests and the bench's CPU baseline (g++), test infrastructure
  oracle/_ref/*        the reference's own CPU tools, built from /root/reference when present
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no FMA contraction, results bit-identical to the reference's CPU arithmetic.
# -packed-fp32-ops off: no v_pk_{add,mul,fma}_f32 anywhere in the device code.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-shared", "-ffp-contract=off",
             "-fno-gpu-flush-denormals-to-zero", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]
HIP_SOURCES = ["csrc/gate.hip", "csrc/tables.hip", "csrc/climate.hip", "csrc/floods.hip", "csrc/flood_lds.hip",
               "csrc/dsx_api.hip", "csrc/gate_noalloc.hip",
               "csrc/scan_api.hip", "csrc/floods_api.hip", "host/search.cpp", "host/serve.cpp"]
HOST_TOOLS = ["host/dsearch_desert.cpp", "host/tgate.cpp", "host/biomemap.cpp", "host/dfunc.cpp"]


def _compile(src, obj):
    """One translation unit to an object (every kernel is launched from the unit that defines it, so
    the units link without relocatable device code)."""
    flags = [f for f in HIP_FLAGS if f != "-shared"]
    r = subprocess.run([HIPCC] + flags + ["-c", os.path.join(HERE, src), "-o", obj], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s%s" % (src, r.stdout, r.stderr))
    return r.stderr


def build_lib(verbose=False):
    out = os.path.join(HERE, "libdsx.so")
    tmp = out + ".tmp"
    odir = os.path.join(HERE, "_build")
    os.makedirs(odir, exist_ok=True)
    objs = [os.path.join(odir, os.path.basename(s) + ".o") for s in HIP_SOURCES]
    jobs = max(1, min(len(HIP_SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        logs = list(ex.map(_compile, HIP_SOURCES, objs))
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc link failed:\n" + r.stdout + r.stderr)
    if verbose:
        for log in logs:
            if log.strip():
                print(log)
    os.replace(tmp, out)
    for src in HOST_TOOLS:
        exe = os.path.join(HERE, os.path.splitext(os.path.basename(src))[0])
        # host code only; the HIP headers are included for the __host__ __device__ seed arithmetic
        r = subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
                            "-I/opt/rocm/include", os.path.join(HERE, src), "-L" + HERE, "-ldsx",
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
