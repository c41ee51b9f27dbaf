"""Benchmark of the particle-likelihood hot path on one MI355X.

One step = one pass of the hot path over the whole particle population held
in HBM: eigen-decomposition of every class rate matrix, P(t) for every
(node, rate), level-ordered pruning with the root mixture fused, the
per-particle lnL reduction (Particle::refreshLikelihood(true) for every
particle, Particle.cpp:530-675), then computeNextStep (Miscellaneous.hpp:42-109)
and systematic resampling (main.cpp:179-195) on the resulting lnLs.

Default workload: 64 taxa x 2,000 sites, GTR+G4+I, 1,024 particles, plain FP64
(synthetic alignment simulated under GTR+G4+I, particle states drawn from the
prior).  Prints one JSON line.
"""
import argparse
import json
import os
import re
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "smc-cat_amd"))

import numpy as np  # noqa: E402

METRIC = "pruned node·pattern·category updates/s; SMC ESS/sec vs plain FP64, 1 MI355X"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_column(T, word):
    """Persisted-partials traffic per (pattern, rate) column of one full sweep:
    T-2 internal non-root children read 20 words, T tip children read a 1-byte
    code, T-1 internal nodes write 20 words + an 8-byte log-scale."""
    return (T - 2) * 20 * word + T * 1 + (T - 1) * (20 * word + 8)


def profiled_value(name, config, precision, variant, particles, default_particles):
    """A figure measured with rocprofv3 for this exact workload (machine-readable
    lines of the committed profile summary), or None: traffic_bytes_per_launch (HBM
    bytes per pruning launch, FETCH_SIZE/WRITE_SIZE calibrated against a plain copy),
    mfma_busy (MfmaUtil over the pruning dispatches), mfma_tflops (MfmaFlops / time)."""
    if config != "c2" or variant != "mfma" or particles != default_particles:
        return None
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "prune_mfma_c2_64taxa_2000sites_1024particles.md")
    try:
        with open(path) as f:
            for line in f:
                m = re.match(r"(\w+)\[(\w+)\] = ([0-9.e+]+)", line.strip())
                if m and m.group(1) == name and m.group(2) == precision:
                    return float(m.group(3))
    except OSError:
        return None
    return None


def profiled_traffic(config, precision, variant, particles, default_particles):
    return profiled_value("traffic_bytes_per_launch", config, precision, variant, particles, default_particles)


def cpu_baseline(codes, batch, R, threads, budget_s=10.0):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_binding as orc
    T, S = codes.shape
    ps = [batch.particle(i) for i in range(batch.n)]
    done, t0 = 0, time.perf_counter()
    while True:
        chunk = [ps[(done + k) % len(ps)] for k in range(threads)]
        orc.lnl_batch(codes, chunk, R, threads=threads)
        done += len(chunk)
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": done * (T - 1) * S * R / dt, "unit": "updates/s", "cores": threads,
            "kind": "port",
            "sample": "%d particles (%d taxa x %d sites, K_r=%d) through the CPU FP64 oracle in "
                      "reference loop order, one particle per OpenMP thread, %.1f s" %
                      (done, T, S, R, dt)}


def progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def ess_run(codes, batch, R, precision, batch_size, learned, iterations, seed, dpp_alpha, kmax,
            alg8=0.1, warmup=1, epochs=8):
    """One SMC run (main.cpp:143-336) on the engine; returns its summary."""
    from smccat import Engine, ProposalNetwork
    T, S = codes.shape
    N = batch.n
    eng = Engine(T, S, R, N, kmax, precision, workspaces=2)
    net = None
    try:
        eng.set_alignment(codes)
        eng.set_particles(batch)
        if learned:
            net = ProposalNetwork(eng, seed=seed)
        t0 = time.perf_counter()
        last = [t0]

        def tick(it, temp, ess):
            now = time.perf_counter()
            if now - last[0] > 20.0:
                progress("  iteration %d T=%.5f ESS %.1f (%.0f s)" % (it, temp, ess, now - t0))
                last[0] = now
        res, lnl, _ = eng.run_sampler(N, iterations, seed=seed, using_lg=True, invariant=False, cat=True,
                                      alg8_probability=alg8, batch=batch_size, dpp_alpha=dpp_alpha,
                                      proposal=net, proposal_warmup=warmup, train_epochs=epochs,
                                      on_iteration=tick)
        phases = {k: round(v[0] / 1e3, 3) for k, v in eng.phase_stats().items()}
    finally:
        if net is not None:
            net.close()
        eng.close()
    return {"seconds": res["seconds"], "final_ess": res["final_ess"], "host_phase_seconds": phases,
            "ess_per_sec": res["final_ess"] / res["seconds"],
            "mean_ess": res["ess_sum"] / max(res["iterations"], 1),
            "iterations": res["iterations"], "resamples": res["resamples"],
            "rejuvenation_seconds": res["rejuvenation_seconds"],
            "training_seconds": res["training_seconds"],
            "cat_site_updates": int(res["cat_stats"][0]), "cat_site_moves": int(res["cat_stats"][1]),
            "training_examples": int(res["cat_stats"][2]),
            "lnl_median": float(np.median(lnl))}


def ess_compare(args):
    """SMC ESS/sec of the plain FP64 path (one particle at a time, as the
    reference's per-thread loop; Gibbs CAT move) against the FP32 + batched +
    learned-proposal path on the same CAT data and initial population."""
    from smccat import lg_base_matrix
    from smccat import simulate as sim
    T, S, N = args.ess_taxa, args.ess_sites, args.ess_particles
    codes, cfg = sim.simulate_config("c3", taxa=T, sites=S)
    R = cfg["rates"]
    kr = (5, 60)
    batch = sim.prior_particles(np.random.default_rng(11), N, T, S, num_rates=R, lg=True,
                                lg_table=lg_base_matrix(), k_range=kr)
    out = {"workload": "CAT data as c3 (DP-CAT, 20 profiles ~ Dirichlet(0.5), LG+G4) at %d taxa x %d "
                       "sites; %d particles, K ~ U{%d..%d}; %d rejuvenation iterations; CAT move "
                       "probability 0.1; DP alpha %.1f" % (T, S, N, kr[0], kr[1], args.ess_iterations,
                                                           1.0),
           "metric": "ESS at T = 1 / wall seconds of the SMC loop (initial likelihoods included)"}
    runs = {"plain_fp64": ("fp64", 1, False), "fp64_batched": ("fp64", 0, False),
            "fp32_batched_learned": ("fp32", 0, True)}
    for name, (prec, bs, learned) in runs.items():
        if name == "fp64_batched" and not args.ess_attribution:
            continue
        vals = []
        for seed in range(1, args.ess_seeds + 1):
            vals.append(ess_run(codes, batch, R, prec, bs, learned, args.ess_iterations, seed, 1.0, kr[1]))
            progress("ess %s seed %d: %.1f s, final ESS %.1f, %d iterations" %
                     (name, seed, vals[-1]["seconds"], vals[-1]["final_ess"], vals[-1]["iterations"]))
        out[name] = {k: float(np.mean([v[k] for v in vals])) for k in vals[0] if k != "host_phase_seconds"}
        out[name]["host_phase_seconds"] = vals[0]["host_phase_seconds"]
        out[name]["seeds"] = args.ess_seeds
    out["ratio_ess_per_sec"] = out["fp32_batched_learned"]["ess_per_sec"] / out["plain_fp64"]["ess_per_sec"]
    return out


def c5_bench(args):
    """Learned CAT-allocation proposal on c3 data: Gibbs draws recorded as
    training pairs, GPU training, and the learned move against the Gibbs move."""
    from smccat import Engine, ProposalNetwork, lg_base_matrix
    from smccat import simulate as sim
    codes, cfg = sim.simulate_config("c3", taxa=args.c5_taxa or None, sites=args.c5_sites or None)
    T, S = codes.shape
    R, N, kr = cfg["rates"], args.c5_particles or cfg["particles"], cfg["k_range"]
    batch = sim.prior_particles(np.random.default_rng(5), N, T, S, num_rates=R, lg=True,
                                lg_table=lg_base_matrix(), k_range=kr)
    out = {"workload": "c3 data (%d taxa x %d sites, LG+G4, DP-CAT), %d particles K ~ U{%d..%d}, "
                       "DP alpha 1.0, tempering 1" % (T, S, N, kr[0], kr[1])}
    eng = Engine(T, S, R, N, kr[1] + 8, "fp32", partial_slots=8)
    eng.set_alignment(codes)
    eng.set_particles(batch)
    net = ProposalNetwork(eng, seed=1)
    crp = ProposalNetwork(eng, seed=2)
    crp.params = np.zeros(net.num_params, dtype=np.float32)
    t0 = time.perf_counter()
    rounds = 0
    while net.num_examples < args.c5_examples:
        net.collect_gibbs(range(N), 1.0, seed=100 + rounds, dpp_alpha=1.0)
        rounds += 1
        progress("c5 collected %d examples in %.1f s" % (net.num_examples, time.perf_counter() - t0))
    t_collect = time.perf_counter() - t0
    n = net.num_examples
    hold = max(1, n // 20)
    n_train = n - hold
    # held-out examples: the last 5% (the latest Gibbs iterations)
    before, _ = net.loss_grad(n_train, hold)
    # train on the first 95%: move the held-out block to a second network's store
    held = [net.example(i) for i in range(n_train, n)]
    ev = ProposalNetwork(eng, seed=3)
    ev.add_examples(np.array([h[0] for h in held]), [h[1] for h in held], [h[2] for h in held],
                    [h[3] for h in held])
    crp.add_examples(np.array([h[0] for h in held]), [h[1] for h in held], [h[2] for h in held],
                     [h[3] for h in held])
    crp_loss, _ = crp.loss_grad(0, hold)
    t0 = time.perf_counter()
    losses = net.train(args.c5_epochs, batch=4096, lr=2e-3, seed=1)
    t_train = time.perf_counter() - t0
    progress("c5 trained %d epochs in %.1f s" % (args.c5_epochs, t_train))
    ev.params = net.params
    after, _ = ev.loss_grad(0, hold)
    out["examples"] = n
    out["collect_seconds"] = t_collect
    out["train"] = {"epochs": args.c5_epochs, "seconds": t_train,
                    "examples_per_sec": n * args.c5_epochs / t_train,
                    "epoch_loss": [float(x) for x in losses]}
    out["heldout_loss"] = {"crp_only": crp_loss, "untrained": before, "trained": after}
    # the moves on the same (Gibbs-updated) particles: site updates per second
    iters = args.c5_move_iters
    moves = {}
    for name, fn in (("gibbs", lambda: eng.gibbs_partition_move(range(N), 1.0, seed=7, num_iter=iters,
                                                                   dpp_alpha=1.0)),
                     ("learned", lambda: net.move(range(N), 1.0, seed=7, num_iter=iters, dpp_alpha=1.0)),
                     ("crp_proposal", lambda: crp.move(range(N), 1.0, seed=7, num_iter=iters,
                                                       dpp_alpha=1.0))):
        t0 = time.perf_counter()
        st = fn()
        dt = time.perf_counter() - t0
        progress("c5 move %s: %.1f s" % (name, dt))
        moves[name] = {"seconds": dt, "site_updates_per_sec": N * iters / dt,
                       "site_moves": int(st[0] if name == "gibbs" else st[2])}
        if name != "gibbs":
            moves[name]["acceptance"] = float(st[2] / max(st[1], 1))
    out["moves"] = moves
    out["speedup_site_updates"] = moves["learned"]["site_updates_per_sec"] / moves["gibbs"]["site_updates_per_sec"]
    for x in (net, crp, ev):
        x.close()
    eng.close()
    return out


def rejuv_bench(args):
    """Cost of one rejuvenation step per particle on c4-shaped data (200 taxa x
    10,000 sites, GTR+G4+I, CAT K ~ U{5..120}): the plain FP64 path runs one
    particle at a time with the Gibbs CAT move, as the reference's per-thread
    loop does; the fast path runs FP32, all resident particles in lockstep, with
    the learned CAT move.  Both run the same moves on the same target, so at
    equal ESS per SMC step the ratio of these costs is the ratio of ESS/sec."""
    from smccat import Engine, ProposalNetwork
    from smccat import simulate as sim
    codes, cfg = sim.simulate_config("c4", taxa=args.rj_taxa or None, sites=args.rj_sites or None)
    T, S = codes.shape
    R, kr = cfg["rates"], cfg["k_range"]
    kmax = kr[1] + 8
    nf, npl, it, tau = args.rj_particles, args.rj_plain, args.rj_iterations, args.rj_temp
    batch = sim.prior_particles(np.random.default_rng(4), nf, T, S, num_rates=R, invar=True, k_range=kr,
                                max_classes=kmax)
    out = {"workload": "c4 data (%d taxa x %d sites, GTR+G4+I, DP-CAT), K ~ U{%d..%d}, %d MCMC iterations, "
                       "CAT move probability 0.1, DP alpha 1.0, tempering %.2f" % (T, S, kr[0], kr[1], it, tau)}
    # fast path: FP32, lockstep batch, learned CAT move (trained beforehand on Gibbs draws)
    eng = Engine(T, S, R, nf, kmax, "fp32", workspaces=2)
    eng.set_graphs(not args.no_graphs)
    if args.chain_max >= 0:
        eng.set_chain_max_nodes(args.chain_max)
    eng.set_mixture("logspace")
    eng.set_alignment(codes)
    eng.set_particles(batch)
    net = ProposalNetwork(eng, seed=1)
    t0 = time.perf_counter()
    net.collect_gibbs(range(nf), tau, seed=3, num_iter=args.rj_train_iters, dpp_alpha=1.0)
    t_col = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = net.train(args.rj_epochs, batch=4096, lr=2e-3, seed=1)
    t_train = time.perf_counter() - t0
    progress("rejuv: %d training examples in %.1f s, trained in %.1f s" % (net.num_examples, t_col, t_train))
    eng.set_particles(batch)
    eng.phase_stats(reset=True)
    t0 = time.perf_counter()
    lnl, stats = eng.rejuvenate(tau, it, seed=5, invariant=True, gibbs_probability=0.1, dpp_alpha=1.0,
                                proposal=net)
    t_fast = time.perf_counter() - t0
    phases = {k: round(v[0] / 1e3, 3) for k, v in eng.phase_stats().items()}
    progress("rejuv fast: %d particles in %.1f s" % (nf, t_fast))
    out["fp32_batched_learned"] = {"particles": nf, "seconds": t_fast, "seconds_per_particle": t_fast / nf,
                                   "host_phase_seconds": phases, "mh_stats": stats.tolist(),
                                   "collect_seconds": t_col, "train_seconds": t_train,
                                   "final_train_loss": float(losses[-1]) if len(losses) else None}
    # The sampler's shape (smc_run_sampler): MCMC batch by batch, then one CAT lockstep
    # over the chosen particles of the whole population (about 410 of c4's 4,096 at
    # probability 0.1) and one refresh of every particle.  Here all resident particles
    # move in one lockstep and a tenth of that move is charged per particle; the
    # refresh is charged in full.
    eng.set_particles(batch)
    t0 = time.perf_counter()
    eng.rejuvenate(tau, it, seed=5, invariant=True, gibbs_probability=0.0, dpp_alpha=1.0)
    t_mcmc = time.perf_counter() - t0
    steps_fast = eng.rejuvenation_steps(it, 5, np.arange(nf), invariant=True)
    t0 = time.perf_counter()
    net.move(range(nf), tau, seed=6, dpp_alpha=1.0)
    t_move = time.perf_counter() - t0
    t0 = time.perf_counter()
    eng.refresh_likelihood(0, nf)
    t_ref = time.perf_counter() - t0
    per_def = (t_mcmc + 0.1 * t_move + t_ref) / nf
    progress("rejuv fast, sampler shape: %.1f ms MCMC + 0.1 x %.1f ms CAT lockstep of %d + %.1f ms refresh "
             "per particle" % (t_mcmc / nf * 1e3, t_move / nf * 1e3, nf, t_ref / nf * 1e3))
    out["fp32_batched_learned_sampler_shape"] = {
        "particles": nf, "mcmc_seconds_per_particle": t_mcmc / nf,
        "mcmc_seconds_per_step": t_mcmc / float(steps_fast.sum()),
        "cat_lockstep_particles": nf, "cat_move_seconds_per_particle": t_move / nf,
        "refresh_seconds_per_particle": t_ref / nf, "seconds_per_particle": per_def}
    net.close()
    eng.close()
    if args.rj_batches > 1:
        # The sampler's lockstep batches over a larger population: a batch runs as many
        # rounds as its longest chain has MH steps, so the sampler keys the slots'
        # random streams in order of their chains' step counts.  MCMC only, nf
        # particles per batch, keys in slot order against keys sorted by steps.
        nb = args.rj_batches
        popn = nf * nb
        pop = sim.prior_particles(np.random.default_rng(9), popn, T, S, num_rates=R, invar=True, k_range=kr,
                                  max_classes=kmax)
        eng = Engine(T, S, R, popn, kmax, "fp32", workspaces=2, partial_slots=nf,
                     pmat_classes=int(nf * (kr[0] + kr[1]) / 2 * 1.4))
        eng.set_graphs(not args.no_graphs)
        if args.chain_max >= 0:
            eng.set_chain_max_nodes(args.chain_max)
        eng.set_mixture("logspace")
        eng.set_alignment(codes)
        slot_keys = np.arange(popn, dtype=np.int32)
        steps = eng.rejuvenation_steps(it, 5, slot_keys, invariant=True)
        sorted_keys = slot_keys[np.argsort(-steps, kind="stable")]
        res = {"particles": popn, "batch": nf, "mean_steps": float(steps.mean())}
        for name, keys in (("slot_order", slot_keys), ("sorted_by_steps", sorted_keys)):
            eng.set_particles(pop)
            t0 = time.perf_counter()
            for b in range(nb):
                eng.rejuvenate(tau, it, seed=5, first=b * nf, count=nf, invariant=True, gibbs_probability=0.0,
                               stream_keys=keys[b * nf:(b + 1) * nf])
            dt = time.perf_counter() - t0
            rounds = [int(steps[keys[b * nf:(b + 1) * nf]].max()) for b in range(nb)]
            res[name] = {"mcmc_seconds_per_particle": dt / popn, "mcmc_seconds_per_step": dt / float(steps.sum()),
                         "rounds_per_batch": rounds}
            progress("rejuv population of %d in batches of %d, %s keys: %.1f ms MCMC per particle, rounds %s"
                     % (popn, nf, name, dt / popn * 1e3, rounds))
        eng.close()
        out["fp32_batched_population"] = res
    if npl < 1:  # fast path only (profiling)
        return out
    # plain path: FP64, one particle at a time; the Gibbs move is timed on its own
    # (one particle, a tenth of its S/2 iterations, scaled) and weighted by its
    # probability 0.1, so a small sample gives the expected cost
    sub = batch.take(np.arange(npl))
    eng = Engine(T, S, R, npl, kmax, "fp64", workspaces=2)
    eng.set_graphs(not args.no_graphs)
    if args.chain_max >= 0:
        eng.set_chain_max_nodes(args.chain_max)
    eng.set_mixture("logspace")
    eng.set_alignment(codes)
    eng.set_particles(sub)
    t0 = time.perf_counter()
    for i in range(npl):
        eng.rejuvenate(tau, it, seed=5, first=i, count=1, invariant=True, gibbs_probability=0.0)
    t_mcmc = (time.perf_counter() - t0) / npl
    steps_plain = eng.rejuvenation_steps(it, 5, np.arange(npl), invariant=True)
    gi = max(1, S // 20)
    t0 = time.perf_counter()
    eng.gibbs_partition_move([0], tau, seed=7, num_iter=gi, dpp_alpha=1.0)
    eng.refresh_likelihood(0, 1)
    t_gibbs = (time.perf_counter() - t0) * (S / 2.0) / gi
    eng.close()
    per = t_mcmc + 0.1 * t_gibbs
    progress("rejuv plain: %.2f s MCMC + 0.1 x %.1f s Gibbs per particle" % (t_mcmc, t_gibbs))
    s_plain = t_mcmc * npl / float(steps_plain.sum())
    out["plain_fp64"] = {"particles_sampled": npl, "mcmc_seconds_per_particle": t_mcmc,
                         "mcmc_seconds_per_step": s_plain, "mean_steps": float(steps_plain.mean()),
                         "gibbs_seconds_per_move": t_gibbs, "seconds_per_particle": per}
    out["ratio_per_particle"] = per / out["fp32_batched_learned"]["seconds_per_particle"]
    out["ratio_per_particle_sampler_shape"] = per / per_def
    # the same comparison per MH step: a few sampled chains differ in length from the
    # batch's average, so both MCMC parts are priced at the mean step count of the
    # batch's chains (or of the population's, when measured)
    sh = out["fp32_batched_learned_sampler_shape"]
    m = float(steps_fast.mean())
    s_fast = sh["mcmc_seconds_per_step"]
    if "fp32_batched_population" in out:
        m = out["fp32_batched_population"]["mean_steps"]
        s_fast = out["fp32_batched_population"]["sorted_by_steps"]["mcmc_seconds_per_step"]
    plain_m = s_plain * m + 0.1 * t_gibbs
    fast_m = s_fast * m + 0.1 * sh["cat_move_seconds_per_particle"] + sh["refresh_seconds_per_particle"]
    out["per_step_normalised"] = {"mean_steps": m, "plain_seconds_per_particle": plain_m,
                                  "fast_seconds_per_particle": fast_m, "ratio": plain_m / fast_m}
    progress("rejuv per-step normalised at %.0f steps: plain %.3f s, fast %.4f s per particle: %.2fx"
             % (m, plain_m, fast_m, plain_m / fast_m))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4"])
    ap.add_argument("--partial-slots", type=int, default=0,
                    help="particles whose partials are resident at once (0: all, or a "
                         "size that fits HBM for c3/c4)")
    ap.add_argument("--particles", type=int, default=0)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--variant", default="mfma", choices=["mfma", "valu"])
    ap.add_argument("--mixture", default="", choices=["", "reference", "logspace"],
                    help="root mixture: the reference formula (default for c1/c2) or its "
                         "log-space form (default for c3/c4, where exp(logSiteLike) underflows)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--rj-taxa", type=int, default=0)
    ap.add_argument("--rj-sites", type=int, default=0)
    ap.add_argument("--rj-particles", type=int, default=150)
    ap.add_argument("--rj-plain", type=int, default=8)
    ap.add_argument("--rj-iterations", type=int, default=16)
    ap.add_argument("--rj-temp", type=float, default=0.5)
    ap.add_argument("--rj-train-iters", type=int, default=1000)
    ap.add_argument("--rj-epochs", type=int, default=6)
    ap.add_argument("--rj-batches", type=int, default=4,
                    help="lockstep batches of the population measurement (0: skip it)")
    ap.add_argument("--mode", default="sweep", choices=["sweep", "c5", "rejuv"],
                    help="sweep: the updates/s line (with the SMC ESS/sec comparison); c5: the "
                         "learned-proposal benchmark")
    ap.add_argument("--no-ess", action="store_true", help="skip the SMC ESS/sec comparison")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--unroll", default="",
                    help="fixed launch shape instead of autotuning: U0,U1,U2[,U_runtime]")
    ap.add_argument("--blocks-per-cu", type=int, default=32)
    ap.add_argument("--xcd-groups", type=int, default=8)
    ap.add_argument("--no-graphs", action="store_true", help="time eager launches instead of graph replays")
    ap.add_argument("--chain-max", type=int, default=-1,
                    help="incremental refreshes: dirty sets of at most this many stored nodes run as chains "
                         "(-1: the engine default)")
    ap.add_argument("--ess-taxa", type=int, default=16)
    ap.add_argument("--ess-sites", type=int, default=200)
    ap.add_argument("--ess-particles", type=int, default=64)
    ap.add_argument("--ess-iterations", type=int, default=4)
    ap.add_argument("--ess-seeds", type=int, default=3)
    ap.add_argument("--ess-attribution", action="store_true",
                    help="also run FP64 batched with the Gibbs move")
    ap.add_argument("--c5-taxa", type=int, default=0)
    ap.add_argument("--c5-sites", type=int, default=0)
    ap.add_argument("--c5-particles", type=int, default=0)
    ap.add_argument("--c5-examples", type=int, default=1_000_000)
    ap.add_argument("--c5-epochs", type=int, default=6)
    ap.add_argument("--c5-move-iters", type=int, default=500)
    ap.add_argument("--fp32-error", type=int, default=64,
                    help="fp32 runs: particles re-evaluated in FP64 for the error report")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(particles would shard across GPUs by slot)", file=sys.stderr)
        sys.exit(2)
    if args.mode == "rejuv":
        print(json.dumps({"metric": "rejuvenation step cost per particle, plain FP64 vs FP32 + batched + learned",
                          "rejuv": rejuv_bench(args)}), flush=True)
        return
    if args.mode == "c5":
        print(json.dumps({"metric": "learned CAT-allocation proposal: training and move throughput",
                          "c5": c5_bench(args)}), flush=True)
        return

    from smccat import Engine, lg_base_matrix
    from smccat import simulate as sim

    codes, cfg = sim.simulate_config(args.config)
    T, S, R = cfg["taxa"], cfg["sites"], cfg["rates"]
    N = args.particles or cfg["particles"]
    kr = cfg["k_range"]
    rng = np.random.default_rng(1)
    batch = sim.prior_particles(rng, N, T, S, num_rates=R, invar=cfg["invar"], lg=cfg["lg"],
                                lg_table=lg_base_matrix() if cfg["lg"] else None, k_range=kr)
    word = 8 if args.precision == "fp64" else 4
    slots = args.partial_slots
    if slots == 0:
        # partials of one wave within ~160 GB of HBM
        per_particle = (T - 1) * R * S * (20 * word + 4)
        slots = int(min(N, max(1, 160e9 // per_particle)))
    eng = Engine(T, S, R, N, kr[1], args.precision, partial_slots=slots,
                 pmat_classes=int(slots * (kr[0] + kr[1]) // 2 * 1.25) if kr[1] > 1 else 0)
    eng.set_prune_variant(args.variant)
    mixture = args.mixture or ("logspace" if args.config in ("c3", "c4") else "reference")
    eng.set_mixture(mixture)
    eng.set_alignment(codes)
    eng.set_particles(batch)
    target = int(0.6 * N)

    def step():
        lnl, nf = eng.refresh_likelihood()
        w = np.full(N, -np.log(N))
        w, norm, ess, temp = eng.compute_next_step(w, lnl, target, 0.0)
        eng.systematic_resample(norm, 0.5)
        return lnl, nf, ess, temp

    progress("sweep workload %s: %d particles uploaded" % (args.config, N))
    tune = None
    if args.variant == "mfma" and args.unroll:
        u = [int(x) for x in args.unroll.split(",")]
        eng.set_tuning(args.blocks_per_cu, u[:3], args.xcd_groups, unroll_runtime=u[3] if len(u) > 3 else None)
    elif args.variant == "mfma" and not args.no_autotune:
        tune = eng.autotune(0, N, sweeps=2)
        progress("autotune: %s" % tune)
    for _ in range(args.warmup):
        step()
    # timed steps: each sweep is one replay of its captured graph
    eng.set_graphs(not args.no_graphs)
    if args.chain_max >= 0:
        eng.set_chain_max_nodes(args.chain_max)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lnl, nf, ess, temp = step()
    dt = time.perf_counter() - t0
    progress("sweep: %d steps in %.2f s" % (args.steps, dt))
    # per-kernel durations (HIP events on the engine's stream): the same kernels
    # and launch shapes, launched eagerly so each launch can be timed
    n_timed = max(1, min(args.steps, 3))
    eng.set_timing(True)
    eng.reset_stats()
    for _ in range(n_timed):
        step()
    stats = eng.kernel_stats()
    eng.set_timing(False)

    updates = N * (T - 1) * S * R
    value = updates * args.steps / dt
    bytes_sweep = N * S * R * algorithmic_bytes_per_column(T, word)
    n_prune, ms_prune = stats["prune"]
    launches_per_sweep = n_prune / n_timed
    prune_ms_per_sweep = ms_prune / n_timed
    achieved = bytes_sweep / (prune_ms_per_sweep * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": "%s: %d taxa x %d sites, %s+G%d%s%s, %d particles; step = full "
                               "refreshLikelihood(true) of every particle + computeNextStep + "
                               "systematic resampling" %
                               (args.config, T, S, "LG" if cfg["lg"] else "GTR", R,
                                "+I" if cfg["invar"] else "",
                                " CAT K~U{%d..%d}" % kr if kr[1] > 1 else "", N),
                   "taxa": T, "sites": S, "rates": R, "particles": N, "partial_slots": slots,
                   "precision": args.precision, "prune_kernel": args.variant, "mixture": mixture,
                   "updates_per_step": updates},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic(args.config, args.precision, args.variant, N,
                                                 cfg["particles"]),
                     "mfma": {"busy": profiled_value("mfma_busy", args.config, args.precision, args.variant, N,
                                                     cfg["particles"]),
                              "achieved": profiled_value("mfma_tflops", args.config, args.precision, args.variant,
                                                         N, cfg["particles"]),
                              "peak": 78.6 if word == 8 else 157.3, "unit": "TFLOP/s",
                              "source": "rocprofv3 --pmc MfmaUtil MfmaFlops (committed profile)"},
                     "kernel": "%s<%s>" % ("k_prune_mfma" if args.variant == "mfma" else "k_prune",
                                           "double" if word == 8 else "float"),
                     "launches_per_step": launches_per_sweep,
                     "avg_launch_ms": ms_prune / max(n_prune, 1),
                     "algorithmic_bytes_per_step": bytes_sweep},
        "kernel_ms_per_step": {k: v[1] / n_timed for k, v in stats.items()},
        "launch_shape": tune or eng.tuning(),
        "graphs": eng.graph_stats(),
        "check": {"lnl_finite": bool(np.all(np.isfinite(lnl)) and not nf.any()),
                  "nonfinite_particles": int(np.sum(~np.isfinite(lnl))),
                  "lnl_mean_finite": float(np.mean(lnl[np.isfinite(lnl)])) if np.isfinite(lnl).any()
                  else None, "ess": ess, "temp": temp},
    }
    eng.close()
    if args.precision == "fp32" and args.fp32_error:
        # per-particle error of the fast path against the plain FP64 engine (which
        # matches the FP64 oracle within 1e-10) on a sample of the same particles
        idx = np.arange(0, N, max(1, N // args.fp32_error))[: args.fp32_error]
        sub = batch.take(idx)
        e64 = Engine(T, S, R, len(idx), kr[1], "fp64",
                     partial_slots=min(len(idx), max(1, slots // 2)))
        e64.set_mixture(mixture)
        e64.set_alignment(codes)
        e64.set_particles(sub)
        ref, _ = e64.refresh_likelihood()
        e64.close()
        rel = np.abs(lnl[idx] - ref) / np.abs(ref)
        out["fp32_error_vs_fp64"] = {"particles": int(len(idx)), "max_rel": float(rel.max()),
                                     "median_rel": float(np.median(rel)),
                                     "max_abs": float(np.abs(lnl[idx] - ref).max())}
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        sample = sim.prior_particles(np.random.default_rng(2), 64, T, S, num_rates=R,
                                     invar=cfg["invar"], lg=cfg["lg"],
                                     lg_table=lg_base_matrix() if cfg["lg"] else None, k_range=kr)
        out["cpu_baseline"] = cpu_baseline(codes, sample, R, threads, args.cpu_seconds)
        progress("cpu baseline done")
    else:
        out["cpu_baseline"] = None
    out["smc_ess"] = None if args.no_ess else ess_compare(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
