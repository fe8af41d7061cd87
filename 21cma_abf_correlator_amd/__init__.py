"""MI355X-native FX correlator hot path: a HIP drop-in for the corr-cuda library.

The product is the C-ABI shared library lib/libcorr_cuda.so (built from csrc/ for
gfx950); this package only locates, builds and binds it (corr.py). There is no
CPU compute path in the package.
"""

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(_HERE, "csrc")
LIB_DIR = os.path.join(_HERE, "lib")


def build(jobs=8, always=True):
    """Compile every HIP/C++ source for gfx950 into lib/ (hipcc cross-compiles
    without a GPU); `always` recompiles even what is up to date."""
    cmd = ["make", "-s", f"-j{jobs}", "-C", CSRC]
    if always:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)


from . import corr  # noqa: E402,F401
