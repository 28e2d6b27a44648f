"""Builds libpyrticle_mi355x.so in-tree with hipcc for gfx950.

    python -m pyrticle_amd.build            # compile (always)

The library is one shared object from the HIP sources in csrc/; each source
compiles to an object in parallel, then hipcc links them.  Flags:

* ``--offload-arch=gfx950``: MI355X only;
* ``-ffp-contract=off``: no implicit FMA contraction; kernels that want fused
  multiply-adds call ``fma()`` explicitly, the FP64 localization path keeps the
  reference's two-rounding arithmetic;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 VALU
  instructions;
* ``-mllvm -amdgpu-mfma-vgpr-form``: matrix-core results and accumulators in
  VGPRs (no AGPR copies around the short MFMA chains of the FP32 gather).
"""

from __future__ import annotations

import glob
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libpyrticle_mi355x.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
ARCH = "gfx950"
FLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-mllvm",
         "-amdgpu-mfma-vgpr-form", "-Wno-unused-result"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose: bool = False, extra_flags=()) -> str:
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    inc = ["-I", os.path.join(HERE, "..", "include")]
    out_tmp = LIB_PATH + ".tmp"
    with tempfile.TemporaryDirectory(prefix="pyr_build_") as tmp:
        objs = [os.path.join(tmp, os.path.basename(s) + ".o") for s in srcs]
        cmds = [[hipcc(), *FLAGS, *inc, *extra_flags, "-c", "-o", o, s]
                for s, o in zip(srcs, objs)]
        jobs = max(1, min(len(cmds), os.cpu_count() or 1, 16))
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            results = list(ex.map(lambda c: _run(c, verbose), cmds))
        for res in results:
            if res.returncode != 0:
                sys.stderr.write(res.stdout + res.stderr)
                raise RuntimeError("hipcc failed building %s" % LIB_NAME)
        link = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out_tmp, *objs,
                "-lpthread"]
        res = _run(link, verbose)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("hipcc failed linking %s" % LIB_NAME)
    os.replace(out_tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
