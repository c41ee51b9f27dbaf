"""Writes the golden fixtures under tests/golden/ (run in the development container only).

Fixtures written here are data (inputs and expected outputs), produced by
independent means -- scipy/numpy routines and a text parse of a data table --
never by running the reference program:

* lg_exchangeabilities.json -- the 190 LG exchangeabilities (Le & Gascuel 2008)
  in the 190-vector order of RateMatrices.hpp:97-111 (i = 1..19, j < i), parsed
  as text from RateMatrices.hpp:7-95 (no reference code is executed), plus a
  checksum.
* expm_vectors.json -- Q matrices built as TransitionProbabilityClass.cpp:50-75
  describes, and scipy.linalg.expm(Q*t*r) for several (t, r).
* gamma_vectors.json -- scipy.special.gammaincinv / scipy.stats.gamma.ppf
  quantiles and the discretizeGamma rates (Miscellaneous.hpp:22-37) they imply.
* prior_vectors.json -- scipy log-pdfs used by Particle::lnPrior (Particle.cpp:484-517).

Usage:  python tests/golden/make_golden.py [/path/to/reference]
"""
import hashlib
import json
import os
import re
import sys

import numpy as np
import scipy.linalg
import scipy.special
import scipy.stats

HERE = os.path.dirname(os.path.abspath(__file__))


def coords190():
    out = []
    for i in range(1, 20):
        for j in range(i):
            out.append((i, j))
    return out


def parse_lg(ref_root):
    path = os.path.join(ref_root, "src", "analysis", "RateMatrices.hpp")
    text = open(path).read()
    m = np.full((20, 20), np.nan)
    for i, j, v in re.findall(r"m\(\s*(\d+),\s*(\d+)\)\s*=\s*([0-9.]+);", text):
        m[int(i), int(j)] = float(v)
    assert not np.isnan(m).any(), "LG table incomplete"
    return m


def build_q(base, logits):
    pi = np.exp(logits)
    pi = pi / pi.sum()
    R = np.zeros((20, 20))
    for c, (i, j) in enumerate(coords190()):
        R[i, j] = R[j, i] = np.exp(base[c])
    Q = R * pi[None, :]
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(axis=1))
    mu = float(np.sum(-np.diag(Q) * pi))
    return Q / mu, pi


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    lg = parse_lg(ref_root)
    lg190 = [float(lg[i, j]) for (i, j) in coords190()]
    blob = ",".join("%.7f" % v for v in lg190).encode()
    json.dump({"source": "RateMatrices.hpp:7-95 parsed as text; order RateMatrices.hpp:97-111",
               "values": lg190,
               "symmetric": bool(np.allclose(lg, lg.T)),
               "sha256_of_7dp_csv": hashlib.sha256(blob).hexdigest()},
              open(os.path.join(HERE, "lg_exchangeabilities.json"), "w"), indent=1)

    rng = np.random.default_rng(20240607)
    cases = []
    for case in range(6):
        if case == 0:
            base = np.array(lg190)          # "LG" uses exp(LG) (Particle.cpp:84-93)
            logits = np.zeros(20)
        else:
            base = rng.normal(size=190)
            logits = np.concatenate([rng.normal(size=19), [0.0]])
        Q, pi = build_q(base, logits)
        ps = []
        for (t, r) in [(0.0, 1.0), (1e-4, 1.0), (0.05, 0.3), (0.1, 1.0), (0.7, 2.5), (3.0, 1.7)]:
            ps.append({"t": t, "r": r, "P": scipy.linalg.expm(Q * t * r).ravel().tolist()})
        cases.append({"base": base.tolist(), "logits": logits.tolist(), "pi": pi.tolist(),
                      "Q": Q.ravel().tolist(), "P": ps})
    json.dump({"generator": "scipy.linalg.expm %s" % scipy.__version__, "cases": cases},
              open(os.path.join(HERE, "expm_vectors.json"), "w"))

    gam = []
    for a in [0.05, 0.1, 0.3, 0.5, 0.8, 1.0, 2.0, 5.0, 20.0, 50.0]:
        for q in [1e-6, 0.01, 0.125, 0.375, 0.5, 0.625, 0.875, 0.99]:
            gam.append({"a": a, "q": q, "x": float(scipy.special.gammaincinv(a, q))})
    disc = []
    for a in [0.05, 0.2, 0.5, 0.8, 1.0, 3.0, 10.0, 50.0]:
        for k in [1, 2, 4, 8]:
            qs = [(2 * i + 1) / (2.0 * k) for i in range(k)]
            v = np.array([scipy.stats.gamma.ppf(qq, a, scale=1.0 / a) for qq in qs])
            disc.append({"shape": a, "k": k, "rates": (v * (k / v.sum())).tolist()})
    json.dump({"generator": "scipy.special.gammaincinv / scipy.stats.gamma.ppf %s" % scipy.__version__,
               "gammaincinv": gam, "discretize": disc},
              open(os.path.join(HERE, "gamma_vectors.json"), "w"), indent=0)

    pri = {"gamma_logpdf": [], "norm_logpdf": [], "gammaln": []}
    for (k, th, x) in [(125.0, 0.1, 12.3), (30.0, 0.1, 2.5), (3.0, 2.0, 0.8), (3.0, 2.0, 7.0)]:
        pri["gamma_logpdf"].append({"k": k, "theta": th, "x": x,
                                    "v": float(scipy.stats.gamma.logpdf(x, k, scale=th))})
    for z in [-3.0, -0.5, 0.0, 0.25, 2.0]:
        pri["norm_logpdf"].append({"z": z, "v": float(scipy.stats.norm.logpdf(z))})
    for v in [0.01, 1.0, 2.0, 17.0, 200.01, 5000.0]:
        pri["gammaln"].append({"x": v, "v": float(scipy.special.gammaln(v))})
    json.dump(pri, open(os.path.join(HERE, "prior_vectors.json"), "w"), indent=0)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
