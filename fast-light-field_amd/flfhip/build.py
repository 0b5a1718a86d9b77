"""Build libflf_hip.so in-tree with hipcc for gfx950 (python -m flfhip.build)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")


def build(verbose=False, force=True):
    """Compile every HIP source into flfhip/libflf_hip.so.  Always recompiles
    when force=True (a build request always runs the compiler)."""
    cmd = ["make", "-C", CSRC, "-j4"]
    if force:
        cmd.append("-B")
    res = subprocess.run(cmd, capture_output=True, text=True)
    if verbose or res.returncode != 0:
        sys.stdout.write(res.stdout)
        sys.stderr.write("\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l))
    if res.returncode != 0:
        raise RuntimeError("building libflf_hip.so failed")
    return os.path.join(HERE, "libflf_hip.so")


if __name__ == "__main__":
    print(build(verbose=True))
