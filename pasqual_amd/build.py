"""Build the device library in-tree: pasqual_amd/lib/libpasqual_amd_{u32,u64}.so.

Device code (csrc/*.hip) is compiled once by hipcc for gfx950; the host C files
by gcc; the reference-layout entry points (csrc/host/pq_dropin.c) twice, once
per index width, so that each library matches one reference build variant.

    python -m pasqual_amd.build            # build (always recompiles)
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(LIB, "obj")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")
ARCH = os.environ.get("PQ_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["pq_prims.hip", "pq_sa.hip", "pq_search.hip", "pq_api.hip", "pq_learned.hip",
               "pq_dedup.hip"]
HOST_SOURCES = ["host/pq_synth.c", "host/pq_import.c"]

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed fp32 instructions stay off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I" + INCLUDE,
]
C_FLAGS = ["-O3", "-std=gnu99", "-fPIC", "-fopenmp", "-Wall", "-Wextra",
           "-Wno-unused-parameter", "-I" + INCLUDE]

EXPORTS = """{
  global:
    pq_*;
    construct_index; construct_string_graph; destroy_graph; free_index;
    time_importreads;
  local: *;
};
"""


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    hip_objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        hip_objs.append(obj)
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    host_objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(OBJ, os.path.basename(src).replace(".c", ".o"))
        host_objs.append(obj)
        jobs.append([CC, *C_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    dropin = {}
    for width in ("u32", "u64"):
        obj = os.path.join(OBJ, f"pq_dropin_{width}.o")
        dropin[width] = obj
        jobs.append([CC, *C_FLAGS, f"-D__INDEX_{width.upper()}__", "-c",
                     os.path.join(CSRC, "host/pq_dropin.c"), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr, file=sys.stderr)
    vs = os.path.join(OBJ, "exports.map")
    with open(vs, "w") as f:
        f.write(EXPORTS)
    outs = []
    for width in ("u32", "u64"):
        out = os.path.join(LIB, f"libpasqual_amd_{width}.so")
        tmp = out + ".tmp"
        _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *hip_objs, *host_objs,
              dropin[width], "-Wl,--version-script=" + vs, "-lgomp", "-lm"])
        os.replace(tmp, out)
        outs.append(out)
    return outs


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)
