"""Single-site x candidate-profile scoring (the Gibbs conditional of
Particle.cpp:857-941, Q15 corrected) through the C-ABI, against the oracle's
orc_site_class_lnl.  FP64 throughout: 1e-10 relative."""
import sys

import numpy as np
import pytest

import oracle_binding as orc
from smccat import Engine
from smccat import simulate as sim

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("taxa,sites,R,krange,tau", [(3, 5, 1, (1, 1), 1.0), (12, 60, 4, (2, 6), 0.37),
                                                     (64, 40, 4, (1, 1), 1.0), (200, 20, 2, (3, 8), 0.05)])
def test_site_class_vs_oracle(gpu, taxa, sites, R, krange, tau):
    _say("case site-class: taxa=%d sites=%d R=%d K=%s tau=%g" % (taxa, sites, R, krange, tau))
    rng = np.random.default_rng(taxa + sites)
    codes = rng.integers(0, 20, size=(taxa, sites)).astype(np.int8)
    if sites > 4:
        codes[rng.random(codes.shape) < 0.05] = 20
        codes[:, 1] = 6
        codes[:, 2] = 22
    batch = sim.prior_particles(rng, 5, taxa, sites, num_rates=R, invar=True, k_range=krange)
    n_entries, n_cand = 9, 7
    parts = rng.integers(0, batch.n, size=n_entries).astype(np.int32)
    cols = rng.integers(0, sites, size=n_entries).astype(np.int32)
    ncand = rng.integers(1, n_cand + 1, size=n_entries).astype(np.int32)
    cand = np.zeros((n_entries, n_cand, 20))
    cand[:, :, :19] = rng.normal(size=(n_entries, n_cand, 19))
    with Engine(taxa, sites, R, batch.n, krange[1]) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        got = eng.site_class_lnl(parts, cols, cand, ncand, tau)
    for i in range(n_entries):
        want = orc.site_class_lnl(codes, batch.particle(int(parts[i])), R, int(cols[i]),
                                  cand[i, :ncand[i]], tau)
        np.testing.assert_allclose(got[i, :ncand[i]], want, rtol=1e-10)
        assert np.all(np.isnan(got[i, ncand[i]:]))


def test_site_class_matches_full_sweep(gpu):
    """Scoring every site of a K=1 particle under its own profile reproduces the
    per-particle likelihood of the sweep kernels."""
    _say("case site-class vs sweep")
    rng = np.random.default_rng(3)
    T, S, R = 30, 50, 4
    codes = rng.integers(0, 20, size=(T, S)).astype(np.int8)
    batch = sim.prior_particles(rng, 2, T, S, num_rates=R, invar=True)
    with Engine(T, S, R, 2, 1) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        lnl, _ = eng.refresh_likelihood()
        for p in range(2):
            cand = np.broadcast_to(batch.logits[p, :1], (S, 1, 20)).copy()
            per_site = eng.site_class_lnl(np.full(S, p), np.arange(S), cand)
            assert per_site.sum() == pytest.approx(lnl[p], rel=1e-11)
