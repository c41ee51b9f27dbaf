"""Parity of the HIP engine (through the C-ABI) against the CPU FP64 oracle.

Bar: per-particle log-likelihood within 1e-10 relative on the plain FP64 path
(the tolerance the plain drop-in must meet); resampling assignments bit-exact;
computeNextStep temperature within the bisection tolerance and ESS within 1e-12.
"""
import math
import sys

import numpy as np
import pytest

import oracle_binding as orc
from smccat import Engine, lg_base_matrix
from smccat import simulate as sim

pytestmark = pytest.mark.gpu

RTOL_FP64 = 1e-10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


VARIANTS = ["mfma", "valu"]
RTOL_FP32 = 1e-5   # fast path: FP32 partials and P, FP64 eigen/back-transform and reductions


def _engine_lnl(codes, batch, R, Kmax, precision="fp64", variant="mfma"):
    T, S = codes.shape
    with Engine(T, S, R, batch.n, Kmax, precision) as eng:
        eng.set_prune_variant(variant)
        eng.set_alignment(codes)
        eng.set_particles(batch)
        return eng.refresh_likelihood()


def _oracle_lnl(codes, batch, R):
    return np.array([orc.particle_lnl(codes, batch.particle(i), R) for i in range(batch.n)])


def _with_gaps(rng, codes, frac=0.03):
    c = codes.copy()
    m = rng.random(c.shape) < frac
    c[m] = rng.choice([20, 21, 22], size=int(m.sum()))
    c[:, 0] = 20                    # an all-gap column: invariant, character 0 (Q5)
    return c


@pytest.mark.parametrize("variant", VARIANTS)
def test_c1_full_config(gpu, variant):
    _say("case c1: 16 taxa x 200 sites, LG+G4, 64 particles, %s" % variant)
    codes, cfg = sim.simulate_config("c1")
    rng = np.random.default_rng(101)
    batch = sim.prior_particles(rng, 64, 16, 200, num_rates=4, lg=True,
                                lg_table=sim_lg())
    lnl, nf = _engine_lnl(codes, batch, 4, 1, variant=variant)
    want = _oracle_lnl(codes, batch, 4)
    assert not nf.any()
    np.testing.assert_allclose(lnl, want, rtol=RTOL_FP64)


def sim_lg():
    from smccat import lg_base_matrix
    return lg_base_matrix()


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("taxa,sites,R,invar,seed", [(64, 300, 4, True, 1), (9, 130, 1, True, 2),
                                                     (33, 77, 8, False, 3), (3, 1, 2, True, 4)])
def test_gtr_gamma_invariant(gpu, taxa, sites, R, invar, seed, variant):
    _say("case gtr: taxa=%d sites=%d R=%d invar=%s %s" % (taxa, sites, R, invar, variant))
    rng = np.random.default_rng(seed)
    ch, bl, root = sim.random_tree(rng, taxa)
    codes = rng.integers(0, 20, size=(taxa, sites)).astype(np.int8)
    if sites > 4:
        codes[:, 1] = 5                 # an invariant column
        codes = _with_gaps(rng, codes)
    batch = sim.prior_particles(rng, 12, taxa, sites, num_rates=R, invar=invar)
    lnl, nf = _engine_lnl(codes, batch, R, 1, variant=variant)
    want = _oracle_lnl(codes, batch, R)
    assert not nf.any()
    np.testing.assert_allclose(lnl, want, rtol=RTOL_FP64)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("taxa,sites,R,krange", [(20, 400, 2, (2, 12)), (12, 150, 4, (5, 40)),
                                                 (7, 64, 1, (64, 64))])
def test_cat_classes_fp64(gpu, taxa, sites, R, krange, variant):
    _say("case cat: taxa=%d sites=%d R=%d K=%s %s" % (taxa, sites, R, krange, variant))
    rng = np.random.default_rng(taxa + sites)
    codes = _with_gaps(rng, rng.integers(0, 20, size=(taxa, sites)).astype(np.int8))
    batch = sim.prior_particles(rng, 10, taxa, sites, num_rates=R, invar=True, k_range=krange)
    lnl, nf = _engine_lnl(codes, batch, R, krange[1], variant=variant)
    want = _oracle_lnl(codes, batch, R)
    assert not nf.any()
    np.testing.assert_allclose(lnl, want, rtol=RTOL_FP64)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_p_kernel_forms_bit_identical(gpu, precision):
    """The P kernel's forms (one entry per lane staged through LDS for launches over
    many particles, stored straight from the lanes for a few) give the same lnL bit
    for bit: a refresh of all 16 particles against refreshes of 4 at a time."""
    _say("case P forms: 16 particles at once against 4 at a time, %s" % precision)
    rng = np.random.default_rng(31)
    T, S, R = 30, 300, 4
    codes = _with_gaps(rng, rng.integers(0, 20, size=(T, S)).astype(np.int8))
    batch = sim.prior_particles(rng, 16, T, S, num_rates=R, invar=True, k_range=(3, 20))
    with Engine(T, S, R, 16, 20, precision) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        whole, _ = eng.refresh_likelihood()
        parts = np.concatenate([eng.refresh_likelihood(f, 4)[0] for f in range(0, 16, 4)])
    assert np.array_equal(whole, parts)
    if precision == "fp64":
        np.testing.assert_allclose(whole, _oracle_lnl(codes, batch, R), rtol=RTOL_FP64)


def test_threaded_upload_matches_one_at_a_time(gpu):
    """An upload of many particles builds their slot orders and tiles on host threads;
    it gives the same lnL bit for bit as uploads of one particle at a time, and an
    invalid particle is reported by the lowest index, as an in-order upload would."""
    from smccat import EngineError
    _say("case threaded upload: 96 particles at once against one at a time")
    rng = np.random.default_rng(47)
    T, S, R, N = 24, 400, 4, 96
    codes = _with_gaps(rng, rng.integers(0, 20, size=(T, S)).astype(np.int8))
    batch = sim.prior_particles(rng, N, T, S, num_rates=R, invar=True, k_range=(2, 12))
    with Engine(T, S, R, N, 12, "fp64") as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        whole, _ = eng.refresh_likelihood()
        for i in range(N):
            eng.set_particles(batch.take([i]), first=i)
        single, _ = eng.refresh_likelihood()
        assert np.array_equal(whole, single)
        bad = batch.take(np.arange(N))
        bad.num_classes[70] = 13
        bad.assignment[90, 5] = -1
        with pytest.raises(EngineError, match="particle 70: num_classes out of range"):
            eng.set_particles(bad)
    np.testing.assert_allclose(whole, _oracle_lnl(codes, batch, R), rtol=RTOL_FP64)


@pytest.mark.parametrize("variant", VARIANTS)
def test_nonfinite_flag(gpu, variant):
    _say("case nonfinite: pinv = 1.5, %s" % variant)
    rng = np.random.default_rng(9)
    codes = rng.integers(0, 20, size=(6, 40)).astype(np.int8)
    batch = sim.prior_particles(rng, 3, 6, 40, num_rates=2, invar=True)
    batch.pinv[1] = 1.5          # log of a negative site likelihood, as the reference
    lnl, nf = _engine_lnl(codes, batch, 2, 1, variant=variant)
    assert nf.tolist() == [0, 1, 0]
    assert math.isnan(lnl[1]) and math.isnan(orc.particle_lnl(codes, batch.particle(1), 2))


def test_partial_refresh_and_reupload(gpu):
    _say("case partial refresh: slots [3, 7) after re-upload")
    rng = np.random.default_rng(17)
    T, S, R = 14, 90, 4
    codes = rng.integers(0, 20, size=(T, S)).astype(np.int8)
    b1 = sim.prior_particles(rng, 10, T, S, num_rates=R, invar=True, k_range=(1, 3))
    b2 = sim.prior_particles(rng, 4, T, S, num_rates=R, invar=True, k_range=(1, 3))
    with Engine(T, S, R, 10, 3) as eng:
        eng.set_alignment(codes)
        eng.set_particles(b1)
        full, _ = eng.refresh_likelihood()
        eng.set_particles(b2, first=3)
        part, _ = eng.refresh_likelihood(3, 4)
        again, _ = eng.refresh_likelihood()
    np.testing.assert_allclose(full, _oracle_lnl(codes, b1, R), rtol=RTOL_FP64)
    np.testing.assert_allclose(part, _oracle_lnl(codes, b2, R), rtol=RTOL_FP64)
    np.testing.assert_allclose(again[3:7], part, rtol=0, atol=0)
    np.testing.assert_allclose(again[:3], full[:3], rtol=0, atol=0)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("taxa,sites,R,krange", [(16, 200, 4, (1, 1)), (30, 500, 4, (3, 20)),
                                                 (64, 300, 2, (1, 1))])
def test_fp32_fast_path_error(gpu, taxa, sites, R, krange, variant):
    _say("case fp32: taxa=%d sites=%d R=%d K=%s %s" % (taxa, sites, R, krange, variant))
    rng = np.random.default_rng(taxa * 7 + sites)
    codes = _with_gaps(rng, rng.integers(0, 20, size=(taxa, sites)).astype(np.int8))
    batch = sim.prior_particles(rng, 8, taxa, sites, num_rates=R, invar=True, k_range=krange)
    lnl32, nf = _engine_lnl(codes, batch, R, krange[1], "fp32", variant)
    want = _oracle_lnl(codes, batch, R)
    assert not nf.any()
    rel = np.abs(lnl32 - want) / np.abs(want)
    _say("fp32 max rel err %.3e" % rel.max())
    assert rel.max() <= RTOL_FP32


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_waves_of_particles_and_classes(gpu, precision):
    """Partials for 3 particles and P for 14 classes at a time: 11 CAT particles
    run in several waves and must give the single-wave results."""
    _say("case waves: 11 particles, partial_slots=3, pmat_classes=14, %s" % precision)
    rng = np.random.default_rng(77)
    T, S, R = 13, 210, 4
    codes = _with_gaps(rng, rng.integers(0, 20, size=(T, S)).astype(np.int8))
    batch = sim.prior_particles(rng, 11, T, S, num_rates=R, invar=True, k_range=(2, 9))
    with Engine(T, S, R, 11, 9, precision, partial_slots=3, pmat_classes=14) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        waved, nf = eng.refresh_likelihood()
        part, _ = eng.refresh_likelihood(4, 5)
    one, _ = _engine_lnl(codes, batch, R, 9, precision)
    assert not nf.any()
    np.testing.assert_allclose(waved, one, rtol=0, atol=0)
    np.testing.assert_allclose(part, one[4:9], rtol=0, atol=0)
    if precision == "fp64":
        np.testing.assert_allclose(waved, _oracle_lnl(codes, batch, R), rtol=RTOL_FP64)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("variant", VARIANTS)
def test_logspace_mixture(gpu, variant, precision):
    """200 taxa on long branches push most sites below log L = -745 (P stays well
    conditioned, every tip costs ~log pi): the reference formula gives -inf there, the
    log-space mixture a finite value equal to the oracle's."""
    _say("case logspace mixture: %s %s" % (variant, precision))
    rng = np.random.default_rng(55)
    T, S, R = 200, 96, 4
    codes = _with_gaps(rng, rng.integers(0, 20, size=(T, S)).astype(np.int8))
    codes[:, 5] = 3                      # invariant columns keep finite reference terms
    batch = sim.prior_particles(rng, 4, T, S, num_rates=R, invar=True, k_range=(1, 4))
    batch.branch_len *= 3.0
    batch.logits *= 2.0
    with Engine(T, S, R, batch.n, 4, precision) as eng:
        eng.set_prune_variant(variant)
        eng.set_alignment(codes)
        eng.set_particles(batch)
        ref_mode, _ = eng.refresh_likelihood()
        eng.set_mixture("logspace")
        log_mode, nf = eng.refresh_likelihood()
    want = np.array([orc.particle_lnl_mode(codes, batch.particle(i), R, 1)[0] for i in range(batch.n)])
    assert np.all(np.isneginf(ref_mode)) and not nf.any()
    np.testing.assert_allclose(log_mode, want, rtol=RTOL_FP64 if precision == "fp64" else RTOL_FP32)



def _underflow_case():
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "underflow_rate_case.json")) as f:
        d = json.load(f)
    T = d["taxa"]
    codes = np.array(d["column"], dtype=np.int8)[:, None]
    logits = np.zeros((1, 1, 20))
    logits[0, 0] = d["logits"]
    batch = sim.ParticleBatch(np.array(d["children"], dtype=np.int32)[None], np.array(d["branch_len"])[None],
                              np.array(d["base"])[None], np.ones(1, dtype=np.int32), logits,
                              np.zeros((1, 1), dtype=np.int32), np.array([d["pinv"]]), np.array([d["shape"]]))
    return codes, batch, d["num_rates"], T


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("variant", VARIANTS)
def test_logspace_underflowed_rate(gpu, variant, precision):
    """Gamma shape 0.0576 on a variable 200-taxon column: the lowest rate's partials
    underflow.  The reference formula is NaN (flagged); log space drops that rate as the
    oracle does, in the pruning kernels and in the single-site class kernel."""
    _say("case logspace underflowed rate: %s %s" % (variant, precision))
    codes, batch, R, T = _underflow_case()
    want, _ = orc.particle_lnl_mode(codes, batch.particle(0), R, 1)
    with Engine(T, 1, R, 1, 1, precision) as eng:
        eng.set_prune_variant(variant)
        eng.set_alignment(codes)
        eng.set_particles(batch)
        ref_mode, nf_ref = eng.refresh_likelihood()
        eng.set_mixture("logspace")
        log_mode, nf = eng.refresh_likelihood()
        site = eng.site_class_lnl(np.zeros(1), np.zeros(1), batch.logits[:, :1])
    _say("  reference %s, logspace %.9f, oracle %.9f, site class %.9f" % (ref_mode[0], log_mode[0], want, site[0, 0]))
    assert math.isnan(ref_mode[0]) and nf_ref[0]
    assert np.isfinite(log_mode[0]) and not nf[0]
    tol = RTOL_FP64 if precision == "fp64" else RTOL_FP32
    assert log_mode[0] == pytest.approx(want, rel=tol)
    assert site[0, 0] == pytest.approx(want, rel=RTOL_FP64)

def test_c2_full_size_properties(gpu):
    _say("case c2 full size: 64 taxa x 2000 sites, GTR+G4+I, 1024 particles")
    codes, cfg = sim.simulate_config("c2")
    rng = np.random.default_rng(202)
    batch = sim.prior_particles(rng, 1024, 64, 2000, num_rates=4, invar=True)
    lnl, nf = _engine_lnl(codes, batch, 4, 1)
    assert not nf.any() and np.all(np.isfinite(lnl)) and np.all(lnl < 0)
    for i in [0, 511, 1023]:
        want = orc.particle_lnl(codes, batch.particle(i), 4)
        assert lnl[i] == pytest.approx(want, rel=RTOL_FP64)


@pytest.mark.parametrize("config,n,check", [("c3", 512, (0, 255, 511)), ("c4", 12, (0, 11))])
def test_cat_full_size_fp32_against_fp64_oracle(gpu, config, n, check):
    """The FP32 fast path at the CAT configurations' full alignment size (c3: 100 taxa x
    5,000 sites, 512 particles, K 5-60; c4: 200 taxa x 10,000 sites, K 5-120, a dozen
    particles of the 4,096), log-space mixture: every lnL finite, and the checked
    particles within the fast path's bound of the FP64 oracle."""
    _say("case %s full size fp32: %d particles" % (config, n))
    codes, cfg = sim.simulate_config(config)
    T, S, R, kr = cfg["taxa"], cfg["sites"], cfg["rates"], cfg["k_range"]
    batch = sim.prior_particles(np.random.default_rng(303), n, T, S, num_rates=R, invar=cfg["invar"],
                                lg=cfg["lg"], lg_table=lg_base_matrix() if cfg["lg"] else None, k_range=kr)
    with Engine(T, S, R, n, kr[1], "fp32", partial_slots=min(n, 64)) as eng:
        eng.set_mixture("logspace")
        eng.set_alignment(codes)
        eng.set_particles(batch)
        lnl, nf = eng.refresh_likelihood()
    assert not nf.any() and np.all(np.isfinite(lnl))
    for i in check:
        want = orc.particle_lnl_mode(codes, batch.particle(i), R, 1)[0]
        _say("  particle %d: fp32 %.6f oracle %.6f" % (i, lnl[i], want))
        assert lnl[i] == pytest.approx(want, rel=RTOL_FP32)


def test_next_step_vs_oracle(gpu):
    _say("case computeNextStep")
    rng = np.random.default_rng(31)
    with Engine(4, 8, 1, 1, 1) as eng:
        for n, last in [(64, 0.0), (1024, 0.3), (4096, 0.95), (3, 0.0)]:
            l = -1000.0 + 30.0 * rng.standard_normal(n)
            w = np.log(rng.random(n) + 0.1)
            w -= np.log(np.exp(w).sum())
            target = int(0.6 * n)
            gw, gn, gess, gt = eng.compute_next_step(w, l, target, last)
            ow, on, oess, ot = orc.compute_next_step(w, l, target, last)
            assert abs(gt - ot) <= 1e-5
            if gt == ot:
                assert gess == pytest.approx(oess, rel=1e-12)
                np.testing.assert_allclose(gn, on, rtol=1e-12, atol=1e-300)
                np.testing.assert_allclose(gw, ow, rtol=1e-15)


def test_systematic_resample_bit_exact(gpu):
    _say("case systematic resampling")
    rng = np.random.default_rng(41)
    with Engine(4, 8, 1, 1, 1) as eng:
        # up to 4096 the kernel searches the two running sums, beyond it it walks
        for n in [1, 2, 7, 64, 1000, 4096, 4097, 6000]:
            for rep in range(5):
                w = rng.random(n) ** 3
                if rep == 1:
                    w[rng.random(n) < 0.7] = 0.0      # runs of equal cumulative weights
                w /= w.sum()
                u = 1.0 - 1e-12 if rep == 2 else rng.random()
                assert eng.systematic_resample(w, u).tolist() == orc.systematic_resample(w, u).tolist(), (n, rep)
        w = np.array([0.1, 0.2, 0.3, 0.4])
        assert eng.systematic_resample(w, 0.5).tolist() == [1, 2, 3, 3]
