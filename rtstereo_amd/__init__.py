"""rtstereo_amd: dense stereo matching on MI355X (gfx950).

The compute path is librtstereo.so (hand-written HIP kernels behind the C ABI
in include/rtstereo.h); librtstereo_seq.so exports the reference's seq/ stage
functions with their C++ signatures on top of it.  This package holds the
Python side: ctypes bindings, the seeded pair generator, resident pipelines
and the serving loop.

Importing the package does not load the HIP library; ``rtstereo_amd.lib()``
(or any submodule that needs it) does, and fails loudly when it is not built.
"""
from __future__ import annotations

import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librtstereo.so")
SEQ_LIB_PATH = os.path.join(HERE, "librtstereo_seq.so")


def lib():
    from . import _lib
    return _lib.lib


def seq_lib():
    """The drop-in seq/ library (SeqLib binding)."""
    from .seqabi import SeqLib
    lib()
    return SeqLib(SEQ_LIB_PATH)
