"""CPU checks of the built libraries: they load, and they export every entry
point the public headers declare (no compute calls: no GPU here)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

INCLUDE = os.path.join(ROOT, "include")
PKG = os.path.join(ROOT, "rtstereo_amd")


def _declared_c_functions():
    src = open(os.path.join(INCLUDE, "rtstereo.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rts_[a-z0-9_]+)\s*\(", src)))


def test_c_abi_exports_every_declared_function():
    lib = C.CDLL(os.path.join(PKG, "librtstereo.so"))
    names = _declared_c_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_c_abi_binding_covers_header():
    from rtstereo_amd import _lib, serve, batch
    names = _declared_c_functions()
    bound = set(_lib._SIGS) | set(serve._SIGS) | set(batch._SIGS)
    assert set(names) <= bound, set(names) - bound


def _declared_seq_functions():
    src = open(os.path.join(INCLUDE, "rtstereo_seq.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return re.findall(r"^(?:void|float)\s+([A-Za-z0-9_]+)\s*\(", src, flags=re.M)


def test_seq_dropin_exports_every_declared_function():
    out = subprocess.run(["nm", "-D", "--defined-only", "-C", os.path.join(PKG, "librtstereo_seq.so")],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    defined = set(re.findall(r" T ([A-Za-z0-9_]+)\(", out))
    names = set(_declared_seq_functions())
    assert len(names) >= 29
    assert names <= defined, names - defined
    # both spellings of the R2L colour cost, two overloads each
    assert out.count("pixelBasedR2L_Color(") == 2 and out.count("pixelBasedCostR2L_Color(") == 2


def test_seq_mangled_names_match_reference_signatures():
    from rtstereo_amd.seqabi import SeqLib, SYMBOLS
    lib = SeqLib(os.path.join(PKG, "librtstereo_seq.so"))
    assert sorted(lib.fn) == sorted(SYMBOLS)


def test_error_path_without_gpu_work():
    from rtstereo_amd import _lib
    # argument validation happens before any device call
    st = _lib.lib.rts_hprefix(None, 0, 5, 5, None)
    assert st == _lib.RTS_ERR_INVALID
    assert b"must be positive" in _lib.lib.rts_last_error()
    d = _lib.PipelineDesc()
    _lib.lib.rts_pipeline_defaults(C.byref(d), 64, 48, 16)
    assert (d.census_hr, d.census_wr, d.cross_L, d.cross_tao, d.aw_r) == (4, 3, 17, 20, 16)
    assert _lib.lib.rts_pipeline_bytes(C.byref(d)) >= 2 * 64 * 48 * 16 * 4


def test_generator_is_deterministic():
    from rtstereo_amd import gen
    a = gen.make_pair(64, 40, 16, 5, 2.0, seed=3)
    b = gen.make_pair(64, 40, 16, 5, 2.0, seed=3)
    assert a.digest() == b.digest()
    assert a.left.dtype == np.uint8 and a.left.shape == (40, 64, 3)
    assert a.gt.max() < 16
    assert gen.plain_384x288().digest() == "185b1fe43b45fbd7"
