"""Builds the gfx950 code objects (the step, the adjoint) and libfluidsx.so in-tree (fluids3_amd/).

Run as `python -m fluids3_amd.build`.  Always recompiles.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")


def build(verbose=False):
    jobs = str(min(8, os.cpu_count() or 1))
    cmd = ["make", "-B", "-j", jobs, "-C", CSRC]
    r = subprocess.run(cmd, capture_output=not verbose, text=True)
    if r.returncode != 0:
        sys.stderr.write((r.stdout or "") + (r.stderr or ""))
        raise RuntimeError("building fluids3_amd failed")
    for f in ("particles_gfx950.hsaco", "adjoint_gfx950.hsaco", "libfluidsx.so"):
        if not os.path.exists(os.path.join(HERE, f)):
            raise RuntimeError(f"build did not produce {f}")


if __name__ == "__main__":
    build(verbose=True)
