"""Build the HIP library (libtb3.so) and the tb3 driver in-tree for gfx950."""
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")


def build(force=True, jobs=4, verbose=False):
    """Compile every HIP source of the package (make -B rebuilds everything)."""
    cmd = ["make", "-C", CSRC, f"-j{jobs}"]
    if force:
        cmd.insert(1, "-B")
    out = subprocess.run(cmd, capture_output=True, text=True)
    if verbose or out.returncode != 0:
        sys.stdout.write(out.stdout)
        sys.stderr.write(out.stderr)
    if out.returncode != 0:
        raise RuntimeError(f"HIP build failed (exit {out.returncode})")


if __name__ == "__main__":
    build(verbose=True)
