"""Builds the native libraries of the package in-tree.

libtal_bsw.so   : HIP kernels + C ABI (include/tal_bsw.h), gfx950 only.
libtal_synth.so : deterministic synthetic workloads (csrc/bsw_synth.c).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TAL_BSW_ARCH", "gfx950")

HIP_SOURCES = ["bsw_kernels.hip", "bsw_scalar.hip", "tal_bsw.hip", "bsw_server.hip",
               "bsw_pairfile.cpp"]
SW_SOURCES = ["sw_kernels.hip", "tal_sw.hip"]
# per-source flags: the SW cell update is four independent packed-16-bit
# chains per step; the ILP scheduler interleaves them, where the default one
# keeps each chain together and pays a wait state between dependent packed
# instructions
SOURCE_FLAGS = {"sw_kernels.hip": ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"],
                # the chaining gap cost divides and multiplies as IEEE float and
                # double, as the reference's source is written
                "chain_kernels.hip": ["-fhip-fp32-correctly-rounded-divide-sqrt",
                                      "-ffp-contract=off"]}
CHAIN_SOURCES = ["chain_kernels.hip", "tal_chain.hip"]
EXTEND_SOURCES = ["extend_kernels.hip", "tal_extend.hip"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{os.path.join(ROOT, 'include')}",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_hip(out=None, extra=(), sources=None, libs=()):
    out = out or os.path.join(HERE, "libtal_bsw.so")
    objs = []
    for src in sources or HIP_SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, *HIP_FLAGS, *SOURCE_FLAGS.get(src, []), *extra, "-c",
              os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", out, "-lpthread",
          *libs])
    for o in objs:
        os.remove(o)
    return out


def build_cli(lib=None, out=None):
    """tal-bench-bsw: bench-bsw's command line on the engine (C++ shim)."""
    out = out or os.path.join(HERE, "tal-bench-bsw")
    _run(["g++", "-O2", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}",
          os.path.join(CSRC, "tal_bench_bsw.cpp"), "-o", out, f"-L{HERE}", "-ltal_bsw",
          "-Wl,-rpath,$ORIGIN"])
    return out


def build_synth(out=None):
    out = out or os.path.join(HERE, "libtal_synth.so")
    _run(["gcc", "-O3", "-fPIC", "-fopenmp", "-shared", "-std=gnu11",
          os.path.join(CSRC, "bsw_synth.c"), "-o", out, "-lm"])
    return out


def build_sw_cli(out=None):
    """tal-bench-sw: bench-smith-waterman's command line on the engine."""
    out = out or os.path.join(HERE, "tal-bench-sw")
    _run(["g++", "-O2", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}",
          os.path.join(CSRC, "tal_bench_sw.cpp"), "-o", out, f"-L{HERE}", "-ltal_sw",
          "-Wl,-rpath,$ORIGIN"])
    return out


def build_chain_cli(out=None):
    """tal-bench-dp-chaining: bench-dp-chaining's input and output on the engine."""
    out = out or os.path.join(HERE, "tal-bench-dp-chaining")
    _run(["g++", "-O2", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}",
          os.path.join(CSRC, "tal_bench_dp_chaining.cpp"), "-o", out, f"-L{HERE}", "-ltal_chain",
          "-Wl,-rpath,$ORIGIN"])
    return out


def build_sw(out=None):
    """libtal_sw.so: the full inter-task SW engine (include/tal_sw.h)."""
    return build_hip(out or os.path.join(HERE, "libtal_sw.so"), sources=SW_SOURCES)


def build_chain(out=None):
    """libtal_chain.so: the DP chaining engine (include/tal_chain.h)."""
    return build_hip(out or os.path.join(HERE, "libtal_chain.so"), sources=CHAIN_SOURCES)


def build_extend(out=None):
    """libtal_extend.so: the extension-pair producer and pipeline
    (include/tal_extend.h), on libtal_chain.so and libtal_bsw.so."""
    return build_hip(out or os.path.join(HERE, "libtal_extend.so"), sources=EXTEND_SOURCES,
                     libs=(f"-L{HERE}", "-ltal_chain", "-ltal_bsw", "-Wl,-rpath,$ORIGIN"))


# Host-code sanitizers for the GPU library: the C ABI, the dispatcher, the
# serving loop and the file I/O run under AddressSanitizer and UBSan, the
# device code is built as usual.  The command-line driver links the
# sanitizer runtime (clang's, as the library was built by clang).
ASAN_DIR = os.path.join(HERE, "_asan")
HOST_SANITIZE = ("-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined "
                 "-Xarch_host -fno-omit-frame-pointer").split()


def build_asan():
    """_asan/libtal_bsw.so and _asan/tal-bench-bsw with sanitized host code."""
    os.makedirs(ASAN_DIR, exist_ok=True)
    lib = build_hip(os.path.join(ASAN_DIR, "libtal_bsw.so"), extra=HOST_SANITIZE)
    cli = os.path.join(ASAN_DIR, "tal-bench-bsw")
    clangxx = os.path.join(os.path.dirname(os.path.realpath(HIPCC)), "..", "llvm", "bin",
                           "clang++")
    san = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer"]
    _run([clangxx, "-O1", "-g", "-std=c++17", *san, f"-I{os.path.join(ROOT, 'include')}",
          os.path.join(CSRC, "tal_bench_bsw.cpp"), "-o", cli, f"-L{ASAN_DIR}", "-ltal_bsw",
          "-Wl,-rpath,$ORIGIN"])
    return lib, cli


def build_all():
    build_synth()
    build_hip()
    build_sw()
    build_chain()
    build_extend()
    build_cli()
    build_sw_cli()
    build_chain_cli()
    build_asan()


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "synth":
        build_synth()
    else:
        build_all()
