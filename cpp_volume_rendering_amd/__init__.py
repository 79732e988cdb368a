"""MI355X-native single-pass structured volume ray casters (HIP for gfx950).

The product is the native library ``libcvr.so`` (C ABI in ``include/cvr.h``) and the
headless C++ plugin layer compiled into it.  This Python package only loads the
library (``native``), builds it (``build``) and defines the seeded workloads
(``workloads``) that tests and the benchmark use.
"""
from . import native  # noqa: F401
