"""MI355X-native lighted volume renderer (the Beams hot path of goodbadwolf/pilot).

The compute path is the HIP library ``pilot_amd/lib/libbeams_mi355x.so`` behind the
C ABI in ``include/beams_mi355x.h``; this package is a thin ctypes binding used by
tests and the benchmark.  Importing :mod:`pilot_amd.beams` fails loudly when the
library has not been built.
"""

import os

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_ROOT = os.path.dirname(PACKAGE_DIR)
LIB_PATH = os.path.join(PACKAGE_DIR, "lib", "libbeams_mi355x.so")
HEADER_PATH = os.path.join(REPO_ROOT, "include", "beams_mi355x.h")

__all__ = ["PACKAGE_DIR", "REPO_ROOT", "LIB_PATH", "HEADER_PATH"]
