"""smc_compute_next_step_noalloc through the C-ABI: the same results as
smc_compute_next_step, bit for bit, on two calls in a row (the second with a
smaller population, so stale staging from the first call would show), and a
clear error above its fixed maximum."""
import ctypes
import sys

import numpy as np
import pytest

from smccat import Engine, _lib

pytestmark = pytest.mark.gpu

MAX_N = 65536


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _noalloc(eng, w, l, target, last):
    w = np.array(w, dtype=np.float64)
    l = np.ascontiguousarray(l, dtype=np.float64)
    norm = np.full_like(w, np.nan)
    ess, temp = ctypes.c_double(), ctypes.c_double()
    rc = eng.lib.smc_compute_next_step_noalloc(eng.h, len(w), _dp(w), _dp(l), _dp(norm), ctypes.byref(ess),
                                               int(target), ctypes.byref(temp), float(last))
    return rc, w, norm, ess.value, temp.value


def _case(rng, n):
    w = np.full(n, -np.log(n)) + rng.normal(0.0, 0.3, n)
    l = rng.normal(-5000.0, 40.0, n)
    return w, l


def test_noalloc_next_step_equals_allocating_entry(gpu):
    rng = np.random.default_rng(5)
    with Engine(8, 16, 1, 4, 1) as eng:
        for step, (n, last) in enumerate(((3000, 0.0), (700, 0.25))):
            _say("case no-alloc next step: call %d, %d particles" % (step, n))
            w, l = _case(rng, n)
            target = int(0.6 * n)
            want = eng.compute_next_step(w, l, target, last)
            rc, got_w, got_norm, got_ess, got_temp = _noalloc(eng, w, l, target, last)
            assert rc == _lib.SMC_OK, eng.lib.smc_engine_last_error(eng.h)
            assert np.array_equal(got_w, want[0])
            assert np.array_equal(got_norm, want[1])
            assert got_ess == want[2] and got_temp == want[3]
            assert last < got_temp <= 1.0 and 1.0 <= got_ess <= n
        # beyond the fixed maximum: refused on the host, before any device work
        w, l = np.zeros(MAX_N + 1), np.zeros(MAX_N + 1)
        rc = _noalloc(eng, w, l, 10, 0.0)[0]
        assert rc == _lib.SMC_ERR_ARG
        assert b"exceed" in eng.lib.smc_engine_last_error(eng.h)
