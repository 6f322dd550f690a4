"""Builds tests/native/_build/host_check: the library's host code (vd_api.cpp) and the check
driver compiled with AddressSanitizer and UndefinedBehaviorSanitizer on the host side only, linked
with the library's kernel objects from the normal build (vorothree_amd/_build, made first if
missing). Every -fsanitize= is given to the host pass alone (-Xarch_host): device code is built
as in the product.

Usage: ``python tests/native/build.py``.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from vorothree_amd import build as vb  # noqa: E402

OUT = os.path.join(HERE, "_build")
EXE = os.path.join(OUT, "host_check")
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]
KERNEL_OBJS = ["vd_grid.o", "vd_shell.o", "vd_cells.o"]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"command failed: {cmd[0]} ... {cmd[-1]}")


def build() -> str:
    objdir = os.path.join(os.path.dirname(vb.LIB), "_build")
    if not all(os.path.exists(os.path.join(objdir, o)) for o in KERNEL_OBJS):
        vb.build()
    os.makedirs(OUT, exist_ok=True)
    cc = vb.hipcc()
    api = os.path.join(OUT, "vd_api_san.o")
    drv = os.path.join(OUT, "host_check.o")
    run([cc, "-x", "hip", *vb.flags(), *SAN, "-c", os.path.join(vb.CSRC, "vd_api.cpp"), "-o", api])
    run([cc, "-x", "hip", *vb.flags(), *SAN, "-c", os.path.join(HERE, "host_check.cpp"), "-o", drv])
    run([cc, f"--offload-arch={vb.ARCH}", *SAN, "-pthread", "-o", EXE, api, drv,
         *[os.path.join(objdir, o) for o in KERNEL_OBJS]])
    return EXE


if __name__ == "__main__":
    print(build())
