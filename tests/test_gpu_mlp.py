"""The learned CAT-allocation proposal on the GPU.

Kernels: loss, gradient, log q and one Adam step of the hand-written HIP
kernels against the float64 oracle (oracle/mlp_oracle.py, itself pinned by
finite differences).  Data: the Gibbs move's recorded draws carry the
features of the oracle's definition.  Move: the Metropolis-Hastings move keeps
the DP-CAT posterior whatever the network -- at tempering 0 with a proposal
tempered at 1 the class count follows the exact CRP law, and at tempering 1 it
agrees with the Gibbs move's.  Training: the trained network lowers the loss
below the CRP-only proposal and raises the move's acceptance rate."""
import sys

import numpy as np
import pytest

import mlp_oracle as mo
from smccat import Engine, ProposalNetwork, lg_base_matrix
from smccat import simulate as sim

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _random_examples(rng, n, kmax):
    hdr = np.zeros((n, mo.HDR), dtype=np.float32)
    hdr[:, :20] = rng.dirichlet(np.ones(20), n)
    hdr[:, 20:40] = rng.dirichlet(np.ones(20), n)
    hdr[:, 40] = rng.uniform(0.1, 1.0, n)
    hdr[:, 41] = np.log(rng.uniform(0.05, 2.0, n))
    K = rng.integers(1, kmax + 1, n).astype(np.int32)
    K[0] = kmax
    cls, lab = [], []
    for i in range(n):
        c = np.zeros((K[i], mo.CLS_STRIDE), dtype=np.float32)
        lg = rng.normal(size=(K[i], 20))
        c[:, :20] = lg - np.log(np.exp(lg).sum(1, keepdims=True))
        c[:, 20] = np.log(rng.integers(1, 30, K[i]))
        cls.append(c)
        lab.append(int(rng.integers(0, K[i] + 1)))
    lab[1] = int(K[1])  # a "new" label
    return hdr, K, np.array(lab, dtype=np.int32), cls


def _engine_small(T=6, S=20, K=4, N=1):
    eng = Engine(T, S, 1, N, K, workspaces=1)
    eng.set_alignment(np.random.default_rng(0).integers(0, 20, (T, S)).astype(np.int8))
    return eng


def _oracle_batch(hdr, K, lab, cls):
    return (hdr[:, :40].astype(np.float64), hdr[:, 40].astype(np.float64),
            hdr[:, 41].astype(np.float64), [c.astype(np.float64) for c in cls], list(lab))


def test_loss_grad_and_logq_match_oracle(gpu):
    _say("case mlp kernels: 70 examples, K up to 200, random theta")
    rng = np.random.default_rng(1)
    hdr, K, lab, cls = _random_examples(rng, 70, 200)
    theta = (rng.normal(0, 0.15, mo.NPARAM)).astype(np.float32)
    with _engine_small() as eng, ProposalNetwork(eng) as net:
        net.params = theta
        net.add_examples(hdr, K, lab, cls)
        loss, g = net.loss_grad(0, len(K))
        logq = net.infer(0, len(K), 256)
    x, tau, la, cl, lb = _oracle_batch(hdr, K, lab, cls)
    th64 = theta.astype(np.float64)
    want_loss, want_g = mo.loss_grad(th64, x, tau, la, cl, lb)
    assert abs(loss - want_loss) <= 1e-5 * abs(want_loss)
    blocks = [(mo.W1, mo.B1), (mo.B1, mo.W2), (mo.W2, mo.B2), (mo.B2, mo.U), (mo.U, mo.UB),
              (mo.UB, mo.V), (mo.V, mo.BN), (mo.BN, mo.NPARAM)]
    for a, b in blocks:
        err = np.abs(g[a:b] - want_g[a:b]).max()
        scale = np.abs(want_g[a:b]).max()
        assert err <= 2e-4 * scale + 1e-9, (a, err, scale)
    logits, _ = mo.forward(th64, x, tau, la, cl)
    for i, lg in enumerate(logits):
        np.testing.assert_allclose(logq[i, :K[i] + 1], mo.log_softmax(lg), atol=2e-4, rtol=0)


def test_one_adam_step_matches_oracle(gpu):
    _say("case mlp adam: one full-batch step")
    rng = np.random.default_rng(2)
    hdr, K, lab, cls = _random_examples(rng, 40, 30)
    theta = (rng.normal(0, 0.15, mo.NPARAM)).astype(np.float32)
    lr = 1e-2
    with _engine_small() as eng, ProposalNetwork(eng) as net:
        net.params = theta
        net.add_examples(hdr, K, lab, cls)
        losses = net.train(1, batch=len(K), lr=lr, seed=1)
        after = net.params
    x, tau, la, cl, lb = _oracle_batch(hdr, K, lab, cls)
    th64 = theta.astype(np.float64)
    want_loss, g = mo.loss_grad(th64, x, tau, la, cl, lb)
    want, _, _ = mo.adam(th64, np.zeros_like(th64), np.zeros_like(th64), g, 1, lr)
    assert abs(losses[0] - want_loss) <= 1e-5 * want_loss
    # Adam's first step is lr * g / (|g| + eps): compare where |g| is clearly resolved
    sel = np.abs(g) > 1e-3 * np.abs(g).max()
    np.testing.assert_allclose(after[sel] - theta[sel], (want - th64)[sel], atol=1e-3 * lr, rtol=0)
    assert np.all(np.abs(after - theta) <= lr * (1 + 1e-5))


def test_recorded_gibbs_examples_have_oracle_features(gpu):
    _say("case mlp collect: Gibbs draws on 8 taxa x 60 sites")
    rng = np.random.default_rng(3)
    codes, _ = sim.simulate_config("c3", taxa=8, sites=60)
    batch = sim.prior_particles(rng, 4, 8, 60, num_rates=4, lg=True, lg_table=lg_base_matrix(),
                                k_range=(2, 5), max_classes=12)
    with Engine(8, 60, 4, 4, 12) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        with ProposalNetwork(eng) as net:
            st = net.collect_gibbs(range(4), 0.7, seed=5, num_iter=10, dpp_alpha=0.8)
            assert net.num_examples == 40
            for i in (0, 17, 39):
                hdr, K, lab, cl = net.example(i)
                assert 0 <= lab <= K
                site_freq = hdr[:20]
                assert abs(site_freq.sum() - 1.0) < 1e-5
                # the class sizes (without the site) add up to S - 1
                assert abs(np.exp(cl[:, 20]).sum() - 59) < 1e-3
                mean = (np.exp(cl[:, 20])[:, None] * np.exp(cl[:, :20])).sum(0) / 59
                np.testing.assert_allclose(hdr[20:40], mean, atol=1e-5)
                assert abs(hdr[40] - 0.7) < 1e-7 and abs(hdr[41] - np.log(0.8)) < 1e-6
            # every recorded frequency vector is a column of the alignment
            cols = {tuple(np.round(np.bincount(codes[:, s], minlength=20) / 8.0, 5)) for s in range(60)}
            assert tuple(np.round(net.example(5)[0][:20].astype(np.float64), 5)) in cols


def _crp_k_law(S, alpha):
    # P(K = k) = alpha^k |s(S, k)| / (alpha)_S  (unsigned Stirling numbers of the first kind)
    s = np.zeros((S + 1, S + 1))
    s[0, 0] = 1
    for n in range(1, S + 1):
        for k in range(1, n + 1):
            s[n, k] = s[n - 1, k - 1] + (n - 1) * s[n - 1, k]
    w = np.array([alpha ** k * s[S, k] for k in range(S + 1)])
    return w / w.sum()


def test_learned_move_keeps_crp_law_at_zero_tempering(gpu):
    _say("case learned move at T=0: 1500 particles, 6 sites, proposal tempered at 1")
    rng = np.random.default_rng(4)
    T, S, N, alpha = 5, 6, 1500, 1.3
    codes = rng.integers(0, 20, (T, S)).astype(np.int8)
    batch = sim.prior_particles(rng, N, T, S, num_rates=1, k_range=(1, 1), max_classes=S)
    with Engine(T, S, 1, N, S) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        with ProposalNetwork(eng, seed=9) as net:
            th = net.params
            th[mo.BN] = 2.0  # a proposal far from the target
            net.params = th
            st = net.move(range(N), 0.0, seed=11, num_iter=120, dpp_alpha=alpha, proposal_tempering=1.0)
        Ks = np.array([eng.particle_state(i)[0] for i in range(N)])
        lg = np.concatenate([eng.particle_state(i)[1][:Ks[i], :19].ravel() for i in range(0, N, 5)])
    _say("  stats %s" % st)
    assert 0 < st[2] < st[1]
    law = _crp_k_law(S, alpha)
    obs = np.bincount(Ks, minlength=S + 1)
    expct = law * N
    keep = expct > 5
    chi2 = (((obs - expct) ** 2 / np.maximum(expct, 1e-12))[keep]).sum()
    # chi-square with <= 5 degrees of freedom: 99.99% quantile is about 25.7
    assert chi2 < 25.7, (obs, expct)
    # class profiles keep the logit prior N(0, 1)
    assert abs(lg.mean()) < 0.05 and abs(lg.var() - 1.0) < 0.08


def test_learned_move_agrees_with_gibbs_at_full_tempering(gpu):
    _say("case learned vs Gibbs at T=1: 600 particles, 6 taxa x 10 sites")
    rng = np.random.default_rng(5)
    T, S, N, alpha = 6, 10, 600, 1.0
    codes, _ = sim.simulate_config("c3", taxa=T, sites=S)
    base = sim.prior_particles(rng, 1, T, S, num_rates=1, lg=True, lg_table=lg_base_matrix())
    batch = base.take(np.zeros(N, dtype=int))
    res = {}
    for how in ("gibbs", "learned"):
        # room for the Gibbs move's 5 auxiliary classes at every K
        with Engine(T, S, 1, N, S + 5) as eng:
            eng.set_alignment(codes)
            eng.set_particles(batch)
            if how == "gibbs":
                for k in range(6):
                    eng.gibbs_partition_move(range(N), 1.0, seed=20 + k, num_iter=20, dpp_alpha=alpha)
            else:
                with ProposalNetwork(eng, seed=3) as net:
                    for k in range(6):
                        net.move(range(N), 1.0, seed=40 + k, num_iter=60, dpp_alpha=alpha)
            res[how] = np.array([eng.particle_state(i)[0] for i in range(N)], dtype=float)
    a, b = res["gibbs"], res["learned"]
    _say("  mean K gibbs %.3f learned %.3f" % (a.mean(), b.mean()))
    se = np.sqrt(a.var() / N + b.var() / N)
    assert abs(a.mean() - b.mean()) < 4.5 * se + 1e-9


def test_training_lowers_loss_and_raises_acceptance(gpu):
    _say("case learned training: Gibbs draws on 16 taxa x 300 sites, 64 particles")
    rng = np.random.default_rng(6)
    T, S, N = 16, 300, 64
    codes, _ = sim.simulate_config("c3", taxa=T, sites=S)
    batch = sim.prior_particles(rng, N, T, S, num_rates=4, lg=True, lg_table=lg_base_matrix(),
                                k_range=(5, 20), max_classes=40)
    with Engine(T, S, 4, N, 40) as eng:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        with ProposalNetwork(eng, seed=1) as net:
            for k in range(3):
                net.collect_gibbs(range(N), 1.0, seed=100 + k, dpp_alpha=1.0)
            n = net.num_examples
            hold = n // 10
            crp = ProposalNetwork(eng, seed=2)
            crp.params = np.zeros(net.num_params, dtype=np.float32)
            crp.add_examples(*_examples_of(net, n - hold, n))
            crp_loss, _ = crp.loss_grad(0, hold)
            before, _ = net.loss_grad(n - hold, hold)
            losses = net.train(6, batch=1024, lr=2e-3, seed=3)
            after, _ = net.loss_grad(n - hold, hold)
            _say("  examples %d held-out loss: CRP-only %.3f untrained %.3f trained %.3f; train %s"
                 % (n, crp_loss, before, after, np.round(losses, 3)))
            assert after < 0.8 * crp_loss and after < before
            st_t = net.move(range(N), 1.0, seed=7, num_iter=100, dpp_alpha=1.0)
            st_c = crp.move(range(N), 1.0, seed=7, num_iter=100, dpp_alpha=1.0)
            crp.close()
    acc_t = st_t[2] / max(st_t[1], 1)
    acc_c = st_c[2] / max(st_c[1], 1)
    _say("  acceptance trained %.3f (moves %d) CRP-only %.3f (moves %d)" % (acc_t, st_t[2], acc_c, st_c[2]))
    # measured on this case: 0.46 vs 0.25 acceptance, 2,259 vs 1,406 accepted moves
    assert acc_t > 1.4 * acc_c and st_t[2] > 1.3 * st_c[2]


def _examples_of(net, first, last):
    hdr, K, lab, cls = [], [], [], []
    for i in range(first, last):
        h, k, l, c = net.example(i)
        hdr.append(h)
        K.append(k)
        lab.append(l)
        cls.append(c)
    return np.array(hdr), np.array(K), np.array(lab), cls
