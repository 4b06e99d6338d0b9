"""Build the gfx950 library ``cupoly_amd/libcupoly_amd.so`` with hipcc.

Each ``csrc/*.hip`` file is compiled to an object in parallel, then linked
into one shared library exporting the C ABI of ``include/cupoly_amd.h``.

Flags:
  --offload-arch=gfx950            MI355X only, no other targets
  -ffp-contract=off                no FMA contraction: the HPS FP64 steps are
                                   reproduced operation for operation
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                   keep the compiler from emitting packed FP32
                                   instructions (the host pass warns that it
                                   does not know the feature; harmless)

Run:  python -m cupoly_amd.build   (or build() from __graft_entry__)
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcupoly_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s%s" % (src, r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_facade_test(verbose)
    return LIB


def build_facade_test(verbose=False):
    """tests/cpp/facade_test: a C++ caller of include/cupoly_amd.hpp."""
    repo = os.path.dirname(HERE)
    src = os.path.join(repo, "tests", "cpp", "facade_test.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(repo, "tests", "cpp", "facade_test")
    cmd = [HIPCC, "-std=c++17", "-O2", src, "-o", out, "-L" + HERE, "-lcupoly_amd",
           "-Wl,-rpath,$ORIGIN/../../cupoly_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("facade test build failed:\n%s%s" % (r.stdout, r.stderr))
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
