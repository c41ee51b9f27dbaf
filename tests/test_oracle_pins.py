"""Pins of the CPU oracle against golden vectors, closed forms and brute force.

Each test names the reference function it pins.  These run on the CPU only.
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg
import scipy.special
import scipy.stats

import bruteforce
import oracle_binding as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


# ---------------------------------------------------------------- a1
def test_lower_triangle_order():
    # RateMatrices.hpp:97-111: i = 1..19, j = 0..i-1
    expect = [(i, j) for i in range(1, 20) for j in range(i)]
    got = [orc.lower_triangle_coord(c) for c in range(190)]
    assert got == expect


def test_lg_fixture_invariants():
    d = _load("lg_exchangeabilities.json")
    v = np.array(d["values"])
    assert v.shape == (190,) and d["symmetric"]
    assert np.all(v > 0)
    # spot values of the published LG table (Le & Gascuel 2008): R<->A, W<->Y? use A-R and V-I
    assert v[0] == pytest.approx(0.425093)      # (1,0) = R-A
    i, j = 19, 9
    assert v[i * (i - 1) // 2 + j] == pytest.approx(10.649107)  # V-I


# ---------------------------------------------------------------- a3 a4
@pytest.mark.parametrize("case", range(6))
def test_q_invariants_and_fixture(case):
    c = _load("expm_vectors.json")["cases"][case]
    Q, pi = orc.build_q(np.array(c["base"]), np.array(c["logits"]))
    np.testing.assert_allclose(pi, c["pi"], rtol=1e-14)
    np.testing.assert_allclose(Q.ravel(), c["Q"], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(Q.sum(axis=1), 0.0, atol=1e-13)            # rows sum to 0
    assert float(np.sum(pi * -np.diag(Q))) == pytest.approx(1.0, abs=1e-14)  # mean rate 1
    F = pi[:, None] * Q                                                       # detailed balance
    np.testing.assert_allclose(F, F.T, atol=1e-14)


# ---------------------------------------------------------------- a5
@pytest.mark.parametrize("case", range(6))
def test_pmatrix_vs_scipy_expm(case):
    c = _load("expm_vectors.json")["cases"][case]
    A, lam, B, pi = orc.eigen(np.array(c["base"]), np.array(c["logits"]))
    for pv in c["P"]:
        P = orc.pmatrix(A, lam, B, pv["t"], pv["r"])
        np.testing.assert_allclose(P.ravel(), pv["P"], rtol=0, atol=2e-13)
    # invariants: P(0) = I, rows sum to 1, Chapman-Kolmogorov, reversibility
    np.testing.assert_allclose(orc.pmatrix(A, lam, B, 0.0, 1.0), np.eye(20), atol=1e-13)
    P1 = orc.pmatrix(A, lam, B, 0.3, 1.0)
    P2 = orc.pmatrix(A, lam, B, 0.5, 1.0)
    np.testing.assert_allclose(P1.sum(axis=1), 1.0, atol=1e-13)
    np.testing.assert_allclose(P1 @ P2, orc.pmatrix(A, lam, B, 0.8, 1.0), atol=1e-13)
    np.testing.assert_allclose(pi[:, None] * P1, (pi[:, None] * P1).T, atol=1e-14)
    slowest = np.sort(np.abs(lam))[1]          # smallest non-zero |eigenvalue|
    np.testing.assert_allclose(orc.pmatrix(A, lam, B, 40.0 / slowest, 1.0), np.tile(pi, (20, 1)),
                               atol=1e-12)


# ---------------------------------------------------------------- a6
def test_gamma_quantiles_vs_scipy():
    d = _load("gamma_vectors.json")
    for g in d["gammaincinv"]:
        x = orc.gamma_p_inv(g["a"], g["q"])
        assert x == pytest.approx(g["x"], rel=1e-11, abs=1e-300), g
    for g in d["discretize"]:
        r = orc.discretize_gamma(g["shape"], g["k"])
        np.testing.assert_allclose(r, g["rates"], rtol=1e-10)


def test_gamma_closed_forms():
    # alpha = 1: exponential quantile -ln(1-q); K = 1: exactly 1 (Miscellaneous.hpp:33-36)
    for q in [0.01, 0.3, 0.5, 0.9]:
        assert orc.gamma_p_inv(1.0, q) == pytest.approx(-math.log1p(-q), rel=1e-13)
    assert orc.discretize_gamma(0.37, 1)[0] == pytest.approx(1.0, abs=1e-15)
    for a in [0.2, 1.0, 7.0]:
        assert np.mean(orc.discretize_gamma(a, 4)) == pytest.approx(1.0, rel=1e-14)


# ---------------------------------------------------------------- a7-a10
def _tiny_state(rng, T, S, K=1, R=1, gaps=True, pinv=0.0):
    import smccat.simulate as sim
    ch, bl, _ = sim.random_tree(rng, T)
    codes = rng.integers(0, 20, size=(T, S)).astype(np.int8)
    if gaps:
        codes[0, 0] = 20            # '-'
        codes[1, 1 % S] = 21        # 'X'
        codes[:, 2 % S] = 22        # all-'?' column: invariant with character 0 (Q5)
        codes[:, 3 % S] = 7         # invariant column
    base = rng.normal(size=190)
    logits = np.zeros((K, 20))
    logits[:, :19] = rng.normal(size=(K, 19))
    asg = rng.integers(0, K, size=S).astype(np.int32)
    asg[:K] = np.arange(K)
    p = dict(children=ch, branch_len=bl, base=base, num_classes=K, logits=logits, assignment=asg,
             pinv=pinv, shape=0.7)
    return codes, p


@pytest.mark.parametrize("T,S,K,R,pinv", [(3, 6, 1, 1, 0.0), (3, 5, 2, 2, 0.2), (4, 4, 1, 4, 0.1),
                                          (4, 5, 3, 2, 0.0)])
def test_oracle_lnl_vs_bruteforce(T, S, K, R, pinv):
    rng = np.random.default_rng(100 + T * 10 + K)
    codes, p = _tiny_state(rng, T, S, K, R, pinv=pinv)
    rates = orc.discretize_gamma(p["shape"], R) if R > 1 else np.ones(1)
    got = orc.particle_lnl(codes, p, R, rates)
    want = bruteforce.lnl(codes, p["children"], p["branch_len"], p["base"], p["logits"],
                          p["assignment"], pinv, rates)
    assert got == pytest.approx(want, rel=1e-12)


def test_oracle_pulley_and_permutation_invariance():
    rng = np.random.default_rng(7)
    codes, p = _tiny_state(rng, 6, 30, K=2, R=2, gaps=True, pinv=0.05)
    base = orc.particle_lnl(codes, p, 2)
    # pulley principle: moving the root along the root's two edges leaves lnL unchanged
    ch = p["children"]
    T = codes.shape[0]
    par = {int(c): T + k for k in range(T - 1) for c in ch[k]}
    root = [n for n in range(T, 2 * T - 1) if n not in par][0]
    a, b = (int(x) for x in ch[root - T])
    q = dict(p)
    bl = p["branch_len"].copy()
    tot = bl[a] + bl[b]
    bl[a], bl[b] = 0.3 * tot, 0.7 * tot
    q["branch_len"] = bl
    assert orc.particle_lnl(codes, q, 2) == pytest.approx(base, rel=1e-12)
    # site permutation (with assignment permuted alongside)
    perm = rng.permutation(codes.shape[1])
    q = dict(p)
    q["assignment"] = p["assignment"][perm]
    assert orc.particle_lnl(codes[:, perm], q, 2) == pytest.approx(base, rel=1e-12)
    # class relabelling
    q = dict(p)
    q["logits"] = p["logits"][::-1].copy()
    q["assignment"] = (1 - p["assignment"]).astype(np.int32)
    assert orc.particle_lnl(codes, q, 2) == pytest.approx(base, rel=1e-12)


def test_oracle_nan_on_invalid_pinv():
    # pInvar > 1 makes finalSiteLike negative for variable sites: log -> NaN (Particle.cpp:667-670)
    rng = np.random.default_rng(3)
    codes, p = _tiny_state(rng, 4, 6, pinv=1.5)
    assert math.isnan(orc.particle_lnl(codes, p, 1))


# ---------------------------------------------------------------- a13
def test_ln_prior_vs_scipy():
    rng = np.random.default_rng(11)
    codes, p = _tiny_state(rng, 5, 12, K=3, R=4)
    T = 5
    par = {int(c) for row in p["children"] for c in row}
    root = [n for n in range(T, 2 * T - 1) if n not in par][0]
    L = sum(p["branch_len"][n] for n in range(2 * T - 1) if n != root)
    nn = 2 * T - 1
    want = scipy.stats.gamma.logpdf(L, nn - 1, scale=0.1)
    want += scipy.stats.gamma.logpdf(p["shape"], 3.0, scale=2.0)
    want += scipy.stats.norm.logpdf(p["base"]).sum()
    alpha = 0.01
    want += math.log(alpha) * 3
    sizes = np.bincount(p["assignment"], minlength=3)
    for k in range(3):
        want += scipy.stats.norm.logpdf(p["logits"][k]).sum() + scipy.special.gammaln(sizes[k])
    want += scipy.special.gammaln(alpha) - scipy.special.gammaln(alpha + 12)
    got = orc.ln_prior(codes, p, 4, using_lg=False, dpp_alpha=alpha)
    assert got == pytest.approx(want, rel=1e-12)


# ---------------------------------------------------------------- a16
def test_next_step_equal_lnl_gives_full_ess():
    n = 64
    w = np.full(n, -math.log(n))
    w2, norm, ess, temp = orc.compute_next_step(w, np.full(n, -123.4), int(0.6 * n), 0.0)
    assert ess == pytest.approx(n, rel=1e-12)
    assert temp == 1.0
    np.testing.assert_allclose(norm, 1.0 / n, rtol=1e-12)


def test_next_step_two_point_closed_form():
    # m particles with lnL = 0, n-m with lnL = -d.  With uniform start weights,
    # ESS(T) = (m + (n-m) x)^2 / (m + (n-m) x^2), x = exp(-d T): solve ESS = target.
    n, m, d = 100, 30, 50.0
    target = 60
    l = np.where(np.arange(n) < m, 0.0, -d)
    w = np.full(n, -math.log(n))
    _, _, ess, temp = orc.compute_next_step(w, l, target, 0.0)

    def ess_of(T):
        x = math.exp(-d * T)
        return (m + (n - m) * x) ** 2 / (m + (n - m) * x * x)
    lo, hi = 0.0, 1.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if ess_of(mid) > target:
            lo = mid
        else:
            hi = mid
    assert abs(temp - hi) <= 1e-5
    assert ess == pytest.approx(ess_of(temp), rel=1e-12)


# ---------------------------------------------------------------- a17
def test_systematic_resample_hand_computed():
    w = np.array([0.1, 0.2, 0.3, 0.4])
    # u = 0.5: ruler 0.125, 0.375, 0.625, 0.875 against cumulative 0.1, 0.3, 0.6, 1.0
    assert orc.systematic_resample(w, 0.5).tolist() == [1, 2, 3, 3]
    assert orc.systematic_resample(w, 0.0).tolist() == [0, 1, 2, 3]


def test_systematic_resample_counts():
    rng = np.random.default_rng(5)
    for n in [7, 64, 1000]:
        w = rng.random(n) ** 4
        w /= w.sum()
        a = orc.systematic_resample(w, rng.random())
        counts = np.bincount(a, minlength=n)
        assert counts.sum() == n
        assert np.all(counts >= np.floor(n * w) - 1) and np.all(counts <= np.ceil(n * w) + 1)
        assert np.all(np.diff(a) >= 0)


# ---------------------------------------------------------------- a15
def test_site_class_lnl_vs_bruteforce():
    rng = np.random.default_rng(21)
    codes, p = _tiny_state(rng, 4, 6, K=1, R=2, pinv=0.1)
    cands = np.zeros((3, 20))
    cands[:, :19] = rng.normal(size=(3, 19))
    tau = 0.37
    got = orc.site_class_lnl(codes, p, 2, 4, cands, tau)
    rates = orc.discretize_gamma(p["shape"], 2)
    for c in range(3):
        want = bruteforce.lnl(codes[:, 4:5], p["children"], p["branch_len"], p["base"], cands[c:c + 1],
                              np.zeros(1, dtype=int), 0.1, rates)
        assert got[c] == pytest.approx(tau * want, rel=1e-12)


def test_logspace_mixture_pins():
    """logspace mode equals the reference formula where that is finite, and equals a
    log-domain brute-force enumeration where exp(logSiteLike) underflows."""
    rng = np.random.default_rng(12)
    codes, p = _tiny_state(rng, 4, 5, K=1, R=2, pinv=0.15)
    ref, _ = orc.particle_lnl_mode(codes, p, 2, 0)
    log, _ = orc.particle_lnl_mode(codes, p, 2, 1)
    assert log == pytest.approx(ref, rel=1e-13)
    # extreme stationary frequencies: pi_i down to e^-300, so variable sites fall below -745
    q = dict(p)
    q["logits"] = np.zeros((1, 20))
    q["logits"][0, :10] = -300.0
    codes2 = codes.copy()
    codes2[:, 0] = [0, 1, 2, 3]          # rare states everywhere: deep underflow
    codes2[:, 1] = 4                     # invariant rare column
    codes2[:, 2] = 15                    # invariant common column
    ref2, sites_ref = orc.particle_lnl_mode(codes2, q, 2, 0)
    log2, sites_log = orc.particle_lnl_mode(codes2, q, 2, 1)
    assert np.isneginf(sites_ref[0]) and np.isfinite(sites_log[0]) and sites_log[0] < -745
    rates = orc.discretize_gamma(q["shape"], 2)
    want = bruteforce.lnl_logdomain(codes2, q["children"], q["branch_len"], q["base"], q["logits"],
                                    q["assignment"], q["pinv"], rates)
    # transition probabilities into states with pi ~ e^-300 carry the absolute (not
    # relative) accuracy of expm / the eigen back-transform, hence 1e-9 here
    assert log2 == pytest.approx(want, rel=1e-9)
    finite = np.isfinite(sites_ref)
    np.testing.assert_allclose(sites_log[finite], sites_ref[finite], rtol=1e-12)


def _underflow_case():
    d = _load("underflow_rate_case.json")
    codes = np.array(d["column"], dtype=np.int8)[:, None]
    p = dict(children=np.array(d["children"]), branch_len=np.array(d["branch_len"]), base=np.array(d["base"]),
             num_classes=1, logits=np.array(d["logits"])[None, :], assignment=np.zeros(1, dtype=np.int32),
             pinv=d["pinv"], shape=d["shape"])
    return codes, p, d["num_rates"]


def test_logspace_mixture_drops_underflowed_rate():
    """A variable column under a gamma shape of 0.0576 (lowest rate 8e-15): the lowest
    rate's partials underflow and the reference formula is NaN (Particle.cpp:653-670
    divides a zero partial by its zero max).  Log space adds 0 for that rate, so the site
    equals the log-sum-exp of the other rates' terms, each evaluated on its own."""
    codes, p, R = _underflow_case()
    assert len(set(codes[:, 0].tolist())) > 1           # variable: no invariant term
    rates = orc.discretize_gamma(p["shape"], R)
    assert rates[0] < 1e-14
    ref, _ = orc.particle_lnl_mode(codes, p, R, 0)
    assert math.isnan(ref)
    terms = []
    for r in rates:
        # R copies of one rate: log(R e^{l_r} (1 - p)) in either mode
        v, _ = orc.particle_lnl_mode(codes, p, R, 1, rates=np.full(R, r))
        terms.append(v - math.log(R) - math.log1p(-p["pinv"]))
    assert not np.isfinite(terms[0]) and np.all(np.isfinite(terms[1:]))
    want = scipy.special.logsumexp(terms[1:]) + math.log1p(-p["pinv"])
    got, _ = orc.particle_lnl_mode(codes, p, R, 1)
    assert got == pytest.approx(want, rel=1e-13)
