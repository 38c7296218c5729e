"""Build the HIP library in-tree with hipcc for gfx950.

`libdgswe.so` is the product, built with `-ffp-contract=off`: its arithmetic
is the reference's IEEE operation sequence (the reference's CPU lanes do not
fuse multiply-adds), so a run can be compared with a CPU restatement bit for
bit, wet/dry switches and time steps included. `libdgswe_fma.so` is the same
source with multiply-adds contracted into FMAs (faster, equal to within
rounding per evaluation; on wet/dry runs the time step follows near-dry
velocities, which rounding changes, so whole runs drift apart).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = [os.path.join(CSRC, "dgswe.hip")]
DEPS = SOURCES + [os.path.join(CSRC, "swe_physics.hpp"), os.path.join(HERE, "..", "include", "dgswe.h")]

LIBS = {
    "libdgswe.so": ["-ffp-contract=off"],
    "libdgswe_fma.so": ["-ffp-contract=fast"],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def command(out: str, extra: list[str]) -> list[str]:
    return [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics", *extra, *SOURCES, "-o", out]


def build(verbose: bool = False) -> list[str]:
    """Compile every variant (always, also when a built library is present)."""
    outs = []
    for name, extra in LIBS.items():
        out = os.path.join(HERE, name)
        tmp = out + ".tmp"
        cmd = command(tmp, extra)
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed for {name}")
        os.replace(tmp, out)
        outs.append(out)
    return outs


if __name__ == "__main__":
    for p in build(verbose=True):
        print(p)
