"""desert-search on MI355X: gfx950 HIP kernels for desert-search's per-seed pipeline behind a C ABI.

The directory name is not a Python identifier; load it with ``_load.package()`` (tests and the
bench do) or import the modules by path.
"""
from . import dsx  # noqa: F401
