"""The C-ABI library loads and exports every symbol include/smccat_hip.h declares,
and its host-only entry points (no GPU needed) agree with the golden vectors."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from smccat import _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(REPO, "include", "smccat_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(smc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(lib, s), s
    # and the Python binding declares a prototype for each
    assert set(syms) == set(_lib.SIGNATURES)


def test_product_lg_table_matches_fixture():
    from smccat import lg_base_matrix
    want = json.load(open(os.path.join(REPO, "tests", "golden", "lg_exchangeabilities.json")))["values"]
    np.testing.assert_array_equal(lg_base_matrix(), np.array(want))


def test_product_discretize_gamma_vs_scipy():
    from smccat import discretize_gamma
    d = json.load(open(os.path.join(REPO, "tests", "golden", "gamma_vectors.json")))
    for g in d["discretize"]:
        np.testing.assert_allclose(discretize_gamma(g["shape"], g["k"]), g["rates"], rtol=1e-10)


def test_discretize_gamma_rejects_bad_args():
    lib = _lib.load()
    out = np.empty(4)
    p = out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.smc_discretize_gamma(-1.0, 4, p) == _lib.SMC_ERR_ARG
    assert lib.smc_discretize_gamma(1.0, 0, p) == _lib.SMC_ERR_ARG


def test_engine_create_rejects_bad_config():
    lib = _lib.load()
    h = ctypes.c_void_p()
    cfg = _lib.EngineConfig(0, 2, 10, 1, 1, 1, 0, 0, 0, 1, 0)      # taxa < 3
    assert lib.smc_engine_create(ctypes.byref(cfg), ctypes.byref(h)) == _lib.SMC_ERR_ARG
    cfg = _lib.EngineConfig(0, 8, 10, 1, 1, 1, 7, 0, 0, 1, 0)      # bad precision
    assert lib.smc_engine_create(ctypes.byref(cfg), ctypes.byref(h)) == _lib.SMC_ERR_ARG
