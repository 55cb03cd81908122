"""Build the in-tree native libraries.

    python -m suntans_amd.build            # product libraries + checkers
    python -m suntans_amd.build --product  # product libraries only

Products (``suntans_amd/lib``):
  libsunhip.so     gfx950 HIP kernels + the thin C ABI (include/sunhip.h)
  libsunstate.so   host state construction (include/sunstate.h), gcc
  libsuntans.so    reference-signature drop-ins (include/suntans_hip.h), gcc,
                   linked against libsunhip.so
Checkers (test infrastructure, ``oracle/``): liboracle.so and, where the
reference sources are present, oracle/_ref/libsunref.so.

Every HIP source is compiled for gfx950 only, without FMA contraction (the
element-wise arithmetic keeps the reference's rounding) and with packed FP32
instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(ROOT, "_build", "obj")

HIP_SOURCES = ["sunhip_grid.hip", "sunhip_q.hip", "sunhip_h.hip", "sunhip_tvd.hip", "sunhip_batch.hip",
               "sunhip_step.hip", "sunhip_upred.hip", "sunhip_vgrid.hip", "sunhip_wpred.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-Wno-unused-value", f"-I{INC}", f"-I{CSRC}"]
CC = os.environ.get("CC", "gcc")


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")


def _hip_object(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", out])
    return out


def build_product() -> None:
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_hip_object, srcs))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", os.path.join(LIB, "libsunhip.so")])
    host = os.path.join(CSRC, "host")
    _run([CC, "-O2", "-fPIC", "-shared", "-Wall", f"-I{INC}", os.path.join(host, "sunstate.c"),
          "-o", os.path.join(LIB, "libsunstate.so"), "-lm", "-pthread"])
    _run([CC, "-O2", "-fPIC", "-shared", "-Wall", f"-I{INC}", os.path.join(host, "dropin.c"),
          f"-L{LIB}", "-lsunhip", "-Wl,-rpath,$ORIGIN", "-o", os.path.join(LIB, "libsuntans.so"), "-lm", "-pthread"])


def build_checkers() -> None:
    make = shutil.which("make")
    if make is None:
        raise RuntimeError("make not found")
    _run([make, "-C", os.path.join(ROOT, "oracle"), "-j8"])


def build(product_only: bool = False) -> None:
    build_product()
    if not product_only:
        build_checkers()


if __name__ == "__main__":
    build(product_only="--product" in sys.argv)
    print("built:", ", ".join(sorted(os.listdir(LIB))))
