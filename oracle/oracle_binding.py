"""ctypes binding of the CPU FP64 oracle (TEST INFRASTRUCTURE ONLY).

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
always as the checker or the CPU baseline, never as part of the product path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsmccat_oracle.so")

_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int32)
_CP = ctypes.POINTER(ctypes.c_int8)


class OrcParticle(ctypes.Structure):
    _fields_ = [("taxa", ctypes.c_int), ("sites", ctypes.c_int), ("num_rates", ctypes.c_int),
                ("codes", _CP), ("children", _IP), ("branch_len", _DP), ("base", _DP),
                ("num_classes", ctypes.c_int), ("logits", _DP), ("assignment", _IP),
                ("pinv", ctypes.c_double), ("rates", _DP)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        lib = ctypes.CDLL(LIB_PATH)
        lib.orc_particle_lnl.restype = ctypes.c_double
        lib.orc_particle_lnl.argtypes = [ctypes.POINTER(OrcParticle)]
        lib.orc_particle_lnl_sites.restype = ctypes.c_double
        lib.orc_particle_lnl_sites.argtypes = [ctypes.POINTER(OrcParticle), _DP]
        lib.orc_particle_lnl_sites_mode.restype = ctypes.c_double
        lib.orc_particle_lnl_sites_mode.argtypes = [ctypes.POINTER(OrcParticle), _DP, ctypes.c_int]
        lib.orc_lnl_batch.restype = ctypes.c_int
        lib.orc_lnl_batch.argtypes = [ctypes.POINTER(OrcParticle), ctypes.c_int, _DP, ctypes.c_int]
        lib.orc_ln_prior.restype = ctypes.c_double
        lib.orc_ln_prior.argtypes = [ctypes.POINTER(OrcParticle), ctypes.c_double, ctypes.c_int,
                                     ctypes.c_double]
        lib.orc_eigen.restype = None
        lib.orc_eigen.argtypes = [_DP, _DP, _DP, _DP, _DP, _DP]
        lib.orc_pmatrix.restype = None
        lib.orc_pmatrix.argtypes = [_DP, _DP, _DP, ctypes.c_double, ctypes.c_double, _DP]
        lib.orc_build_q.restype = None
        lib.orc_build_q.argtypes = [_DP, _DP, _DP, _DP]
        lib.orc_gamma_p.restype = ctypes.c_double
        lib.orc_gamma_p.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.orc_gamma_p_inv.restype = ctypes.c_double
        lib.orc_gamma_p_inv.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.orc_discretize_gamma.restype = None
        lib.orc_discretize_gamma.argtypes = [ctypes.c_double, ctypes.c_int, _DP]
        lib.orc_compute_next_step.restype = None
        lib.orc_compute_next_step.argtypes = [ctypes.c_int, _DP, _DP, _DP, _DP, ctypes.c_int, _DP,
                                              ctypes.c_double]
        lib.orc_systematic_resample.restype = None
        lib.orc_systematic_resample.argtypes = [ctypes.c_int, _DP, ctypes.c_double, _IP]
        lib.orc_site_class_lnl.restype = None
        lib.orc_site_class_lnl.argtypes = [ctypes.POINTER(OrcParticle), ctypes.c_int, ctypes.c_int,
                                           _DP, ctypes.c_double, _DP]
        lib.orc_lower_triangle_coord.restype = None
        lib.orc_lower_triangle_coord.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                 ctypes.POINTER(ctypes.c_int)]
        _lib = lib
    return _lib


def _d(a):
    return a.ctypes.data_as(_DP)


class _Keep:
    """Holds the numpy arrays an OrcParticle points into."""


def make_particle(codes, p, num_rates, rates=None):
    """codes int8 [T,S]; p: dict with children, branch_len, base, num_classes, logits,
    assignment, pinv, shape (as ParticleBatch.particle())."""
    lib = load()
    keep = _Keep()
    T, S = codes.shape
    keep.codes = np.ascontiguousarray(codes, dtype=np.int8)
    keep.children = np.ascontiguousarray(p["children"], dtype=np.int32).reshape(-1)
    keep.blen = np.ascontiguousarray(p["branch_len"], dtype=np.float64)
    keep.base = np.ascontiguousarray(p["base"], dtype=np.float64)
    k = int(p["num_classes"])
    keep.logits = np.ascontiguousarray(np.asarray(p["logits"], dtype=np.float64)[:k].reshape(-1))
    keep.asg = np.ascontiguousarray(p["assignment"], dtype=np.int32)
    if rates is None:
        rates = discretize_gamma(p["shape"], num_rates) if num_rates > 1 else np.ones(1)
    keep.rates = np.ascontiguousarray(rates, dtype=np.float64)
    op = OrcParticle(T, S, num_rates, keep.codes.ctypes.data_as(_CP), keep.children.ctypes.data_as(_IP),
                     _d(keep.blen), _d(keep.base), k, _d(keep.logits), keep.asg.ctypes.data_as(_IP),
                     float(p["pinv"]), _d(keep.rates))
    op._keep = keep
    return op


def particle_lnl(codes, p, num_rates, rates=None):
    return load().orc_particle_lnl(ctypes.byref(make_particle(codes, p, num_rates, rates)))


def particle_site_lnl(codes, p, num_rates, rates=None):
    op = make_particle(codes, p, num_rates, rates)
    out = np.empty(codes.shape[1])
    tot = load().orc_particle_lnl_sites(ctypes.byref(op), _d(out))
    return tot, out


def particle_lnl_mode(codes, p, num_rates, logspace, rates=None):
    op = make_particle(codes, p, num_rates, rates)
    out = np.empty(codes.shape[1])
    tot = load().orc_particle_lnl_sites_mode(ctypes.byref(op), _d(out), int(logspace))
    return tot, out


def lnl_batch(codes, particles, num_rates, threads=0):
    ops = [make_particle(codes, p, num_rates) for p in particles]
    arr = (OrcParticle * len(ops))(*ops)
    out = np.empty(len(ops))
    used = load().orc_lnl_batch(arr, len(ops), _d(out), int(threads))
    return out, used


def ln_prior(codes, p, num_rates, using_lg, dpp_alpha=0.01):
    op = make_particle(codes, p, num_rates)
    return load().orc_ln_prior(ctypes.byref(op), float(p["shape"]), int(using_lg), float(dpp_alpha))


def eigen(base, logits):
    A = np.empty(400); lam = np.empty(20); B = np.empty(400); pi = np.empty(20)
    load().orc_eigen(_d(np.ascontiguousarray(base, dtype=np.float64)),
                     _d(np.ascontiguousarray(logits, dtype=np.float64)), _d(A), _d(lam), _d(B), _d(pi))
    return A.reshape(20, 20), lam, B.reshape(20, 20), pi


def pmatrix(A, lam, B, t, r):
    P = np.empty(400)
    load().orc_pmatrix(_d(np.ascontiguousarray(A, dtype=np.float64)), _d(np.ascontiguousarray(lam)),
                       _d(np.ascontiguousarray(B, dtype=np.float64)), float(t), float(r), _d(P))
    return P.reshape(20, 20)


def build_q(base, logits):
    pi = np.empty(20); Q = np.empty(400)
    load().orc_build_q(_d(np.ascontiguousarray(base, dtype=np.float64)),
                       _d(np.ascontiguousarray(logits, dtype=np.float64)), _d(pi), _d(Q))
    return Q.reshape(20, 20), pi


def gamma_p(a, x):
    return load().orc_gamma_p(float(a), float(x))


def gamma_p_inv(a, p):
    return load().orc_gamma_p_inv(float(a), float(p))


def discretize_gamma(shape, num):
    out = np.empty(num)
    load().orc_discretize_gamma(float(shape), int(num), _d(out))
    return out


def compute_next_step(raw_weights, raw_lnl, target_ess, last_temp):
    w = np.array(raw_weights, dtype=np.float64)
    l = np.ascontiguousarray(raw_lnl, dtype=np.float64)
    norm = np.empty_like(w)
    ess = ctypes.c_double(); temp = ctypes.c_double()
    load().orc_compute_next_step(len(w), _d(w), _d(l), _d(norm), ctypes.byref(ess), int(target_ess),
                                 ctypes.byref(temp), float(last_temp))
    return w, norm, ess.value, temp.value


def systematic_resample(normalized, u):
    w = np.ascontiguousarray(normalized, dtype=np.float64)
    out = np.empty(len(w), dtype=np.int32)
    load().orc_systematic_resample(len(w), _d(w), float(u), out.ctypes.data_as(_IP))
    return out


def site_class_lnl(codes, p, num_rates, site, cand_logits, tempering):
    op = make_particle(codes, p, num_rates)
    cl = np.ascontiguousarray(cand_logits, dtype=np.float64)
    out = np.empty(cl.shape[0])
    load().orc_site_class_lnl(ctypes.byref(op), int(site), cl.shape[0], _d(cl), float(tempering), _d(out))
    return out


def lower_triangle_coord(c):
    i = ctypes.c_int(); j = ctypes.c_int()
    load().orc_lower_triangle_coord(int(c), ctypes.byref(i), ctypes.byref(j))
    return i.value, j.value
