"""disptools_amd -- MI355X (gfx950) implementation of disptools' displacement-generation path.

The product path is the HIP library ``libdisptools_hip.so`` (kernels and C ABI,
``include/disptools_hip.h``) and the ``_disptools`` extension built on it, with
the reference's method table.  Both are built in-tree by
``python -m disptools_amd.build``; importing the extension raises if they are
missing -- there is no CPU fallback.

``disptools_amd.displacements`` mirrors ``disptools.displacements.displacement``
and ``jacobian`` (reference ``disptools/displacements.py:55-92,238-485``).
"""

import importlib
import logging
import sys

import numpy as np

__version__ = "0.4.1"


def load_extension():
    """Import the in-tree ``_disptools`` extension (raises ImportError if unbuilt)."""
    from . import _hip_runtime
    _hip_runtime.preload()   # one HIP runtime per process, shared with torch
    try:
        return importlib.import_module("disptools_amd._disptools")
    except ImportError as e:
        raise ImportError("disptools_amd._disptools is not built; run "
                          "`python -m disptools_amd.build` (%s)" % e) from e


def float_type():
    """numpy float type of the C library (reference ``disptools/__init__.py:26-39``)."""
    size = load_extension().get_float_type_size()
    if size == 32:
        return np.float32
    if size == 64:
        return np.float64
    logging.critical("Unsupported floating point size (%d bit)" % size)
    raise Exception("Unsupported floating point type size")
