"""foundation_amd — MI355X path-tracing pass behind a C ABI (include/fnd_pt.h).

The compute path is libfnd_pt.so (hand-written HIP kernels for gfx950 plus the
host encoder); this package only binds it.  Importing does not load the
library; ``foundation_amd.load()`` does, and raises if it is not built.
"""
from ._abi import LIB_PATH, load  # noqa: F401

__all__ = ["load", "LIB_PATH"]
