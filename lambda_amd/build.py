"""Builds the native libraries in-tree.

* lambda_amd/lib/liblambda_host.so  - host front end (g++, C++20)
* lambda_amd/lib/liblambda_hip.so   - gfx950 integrator (hipcc)
* lambda_amd/lib/lambda             - command-line renderer (main.cpp equivalent)
* oracle/lib/liblambda_oracle.so    - CPU oracle (test infrastructure)
* oracle/_ref/ref_harness           - reference harness, only where the
                                      reference sources are present

Device code is compiled with -ffp-contract=off (ray/box and ray/primitive
tests must round like the reference's FMA-free CPU build) and without packed
FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lambda_amd")
LIB = os.path.join(PKG, "lib")
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
REFERENCE_SRC = "/root/reference/src"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HOST_SOURCES = ["scene_model.cpp", "parser.cpp", "host_api.cpp"]
HIP_SOURCES = ["lambda_hip.hip", "k_passes.hip", "k_shade.hip", "k_persist.hip", "k_persist_occ.hip", "k_persist_vote.hip", "k_persist_guided.hip",
               "k_persist_table.hip", "k_serve.hip", "k_query.hip", "guide_net.hip", "lbvh.hip"]

# No -march: the host front end must round exactly like the reference's
# FMA-free x86-64 build.
HOST_FLAGS = ["-O3", "-DNDEBUG", "-std=gnu++20", "-ffp-contract=off", "-fPIC", "-fopenmp", "-Wall", "-Wextra"]
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=ROOT):
    print("+ " + " ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build step failed: %s" % cmd[0])
    noisy = [ln for ln in (res.stderr or "").splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if noisy:
        sys.stderr.write("\n".join(noisy[:40]) + "\n")


def build_host():
    os.makedirs(LIB, exist_ok=True)
    srcs = [os.path.join(HOST, s) for s in HOST_SOURCES]
    out = os.path.join(LIB, "liblambda_host.so")
    _run([CXX, *HOST_FLAGS, "-shared", "-I" + INC, *srcs, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_hip(jobs: int = 0):
    """Each translation unit to an object in parallel (the kernels are
    spread over several files so that their device code compiles
    concurrently), then one shared library."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(LIB, exist_ok=True)
    objdir = os.path.join(ROOT, "_build", "hip")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in HIP_SOURCES]
    cmds = [[HIPCC, *HIP_FLAGS, "-I" + INC, "-I" + CSRC, "-c", os.path.join(CSRC, s), "-o", o]
            for s, o in zip(HIP_SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=jobs or min(len(cmds), os.cpu_count() or 1)) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    out = os.path.join(LIB, "liblambda_hip.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", *objs, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_cli():
    out = os.path.join(LIB, "lambda")
    _run([CXX, *HOST_FLAGS, "-I" + INC, os.path.join(HOST, "lambda_main.cpp"), "-L" + LIB,
          "-llambda_host", "-llambda_hip", "-Wl,-rpath,$ORIGIN", "-lpthread", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_oracle(with_reference: bool = True):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"])
    if with_reference and os.path.isdir(REFERENCE_SRC):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", "REF=" + REFERENCE_SRC])


def build_all():
    build_host()
    build_hip()
    build_cli()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "host": build_host, "hip": build_hip, "cli": build_cli,
         "oracle": build_oracle}[w]()
