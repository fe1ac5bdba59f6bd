"""Build recipe for the HIP library (gfx950) and the C++ interface shim.

Invoked by __graft_entry__.build(); compiles every time it is called.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "tan_amd", "csrc")
LIBDIR = os.path.join(ROOT, "tan_amd", "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# packed FP32 instructions stay off (only the compiler's scalar/vector FP32)
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]

SOURCES = ["tan_kernels.hip", "tan_room.hip", "tan_noalloc.hip", "tan_engine.cpp", "tan_api.cpp", "tan_wav.cpp"]


def build_library(verbose=False, jobs=None):
    """Compiles every source to an object (in parallel) and links the library."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    out = os.path.join(LIBDIR, "libtan_mi355x.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    objs = [os.path.join(objdir, os.path.basename(s) + ".o") for s in srcs]

    def compile_one(pair):
        src, obj = pair
        cmd = [HIPCC] + HIP_FLAGS + ["-c", "-o", obj, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)

    workers = jobs or min(len(srcs), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    cmd = [HIPCC] + HIP_FLAGS + ["-shared", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, out)
    return out


def build_harness(verbose=False):
    """C++ driver restating TALibTestConvolution against include/TrueAudioNext.h
    (test infrastructure; links the library and the oracle)."""
    out_dir = os.path.join(ROOT, "tests", "cpp", "build")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "talib_harness")
    # host C++ (g++) with HIP's runtime API for the caller-created streams
    cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "oracle"), "-I/opt/rocm/include",
           os.path.join(ROOT, "tests", "cpp", "talib_harness.cpp"), "-o", out,
           "-L" + LIBDIR, "-ltan_mi355x", "-L" + os.path.join(ROOT, "oracle", "build"), "-ltan_oracle",
           "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/../../../tan_amd/lib:$ORIGIN/../../../oracle/build:/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return out


def build_oracle(verbose=False):
    cmd = ["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


if __name__ == "__main__":
    build_library(verbose=True)
    build_oracle(verbose=True)
    build_harness(verbose=True)
