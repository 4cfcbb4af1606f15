"""Build recipe for liblettuce_hip.so (gfx950) from csrc/*.hip with hipcc, and for the probes
bench.py runs (libissue_probe.so and libaccum_probe.so, csrc/probe).

Every HIP translation unit is compiled with packed-FP32 instruction selection disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``) and IEEE arithmetic (no fast-math): the
Morton quantisation must round exactly as the reference CPU backend does.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "liblettuce_hip.so")
# The probes bench.py runs beside the force path (their own libraries, not linked into it): the
# issue-rate probe and the accumulation probe.
PROBES = [(os.path.join(CSRC, "probe", "issue_probe.hip"), os.path.join(HERE, "libissue_probe.so")),
          (os.path.join(CSRC, "probe", "accum_probe.hip"), os.path.join(HERE, "libaccum_probe.so"))]
BUILD = os.path.join(HERE, "_build")
SOURCES = ["kernels_state.hip", "kernels_direct.hip", "kernels_tree.hip", "kernels_build.hip",
           "kernels_walk.hip", "kernels_let.hip", "kernels_adjoint.hip", "kernels_treeadj.hip", "adjoint.hip",
           "treesim.hip", "direct_noalloc.hip", "capi.hip", "scenarios.cpp", "snapshot.cpp"]
ARCH = os.environ.get("LHIP_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++23", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wall", "-Wno-unused-command-line-argument", "-Wno-unused-function"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if warn:
        print(warn, file=sys.stderr)
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    for src, out in PROBES:
        tmp = out + ".tmp"
        r = subprocess.run([HIPCC, *FLAGS, "-shared", "-o", tmp, src], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {os.path.basename(src)}:\n{r.stderr}")
        os.replace(tmp, out)
    if verbose:
        print(f"built {OUT} and the probes {', '.join(os.path.basename(o) for _, o in PROBES)}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
