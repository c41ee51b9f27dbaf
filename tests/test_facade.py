"""The C++ facade (reference call shapes: Alignment, SerializedParticle,
Particle::copyFromSerialized / lnLikelihood / lnPrior / writeToSerialized,
computeNextStep, systematic resampling) driven end to end by its self-test
executable.  CPU part: FASTA parsing (Alignment.cpp:6-104) and the Newick
round trip (Tree.cpp:79-166, 398-437).  GPU part: lnL, lnPrior, batch sweep
and SMC step against the CPU oracle."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_binding as orc
from smccat import build as smbuild
from smccat import simulate as sim

AA = "ARNDCQEGHILKMFPSTWYV-X?"


def newick(children, blen, names, root):
    T = len(names)

    def rec(n):
        if n >= T:
            a, b = children[n - T]
            s = "(" + rec(int(a)) + "," + rec(int(b)) + ")"
            return s if n == root else s + ":%f" % blen[n]
        return names[n] + ":%f" % blen[n]
    return rec(root) + ";"


def _root(children, T):
    kids = {int(c) for row in children for c in row}
    return [n for n in range(T, 2 * T - 1) if n not in kids][0]


def write_inputs(tmp_path, codes, batch, names):
    fa = tmp_path / "aln.fasta"
    with open(fa, "w") as f:
        f.write("# comment line\n")
        for t, nm in enumerate(names):
            seq = "".join(AA[c] for c in codes[t])
            f.write(">%s\n%s\n%s\n" % (nm, seq[: len(seq) // 2].lower(), seq[len(seq) // 2:]))
    pf = tmp_path / "particles.txt"
    T = codes.shape[0]
    newicks = []
    with open(pf, "w") as f:
        for i in range(batch.n):
            root = _root(batch.children[i], T)
            # lengths go through "%f" (6 dp) as the reference's std::to_string (Q11)
            batch.branch_len[i] = np.round(batch.branch_len[i], 6)
            nw = newick(batch.children[i], batch.branch_len[i], names, root)
            newicks.append(nw)
            f.write("newick %s\npinv %.17g\nshape %.17g\n" % (nw, batch.pinv[i], batch.shape[i]))
            f.write("base " + " ".join("%.17g" % v for v in batch.base[i]) + "\n")
            K = int(batch.num_classes[i])
            f.write("classes %d\n" % K)
            for k in range(K):
                f.write("logits " + " ".join("%.17g" % v for v in batch.logits[i, k]) + "\n")
            f.write("assign " + " ".join(str(a) for a in batch.assignment[i]) + "\n")
            f.write("end\n")
    return str(fa), str(pf), newicks


def run_facade(args):
    exe = smbuild.FACADE_EXE
    if not os.path.exists(exe):
        raise RuntimeError("facade_selftest not built")
    r = subprocess.run([exe] + args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        parts = line.split()
        out.setdefault(parts[0], []).append(parts[1:])
    return out


def _case(seed, T=9, S=57, R=4, K=(1, 3)):
    rng = np.random.default_rng(seed)
    codes = rng.integers(0, 23, size=(T, S)).astype(np.int8)
    batch = sim.prior_particles(rng, 5, T, S, num_rates=R, invar=True, k_range=K)
    names = ["taxon_%d" % i for i in range(T)]
    return codes, batch, names


def test_facade_fasta_and_newick_roundtrip_cpu(tmp_path):
    codes, batch, names = _case(1)
    fa, pf, newicks = write_inputs(tmp_path, codes, batch, names)
    out = run_facade([fa, pf, "4", "3", "0", "--cpu"])
    got = [row[1] for row in out["newick"]]
    assert got == newicks


@pytest.mark.gpu
def test_facade_matches_oracle(gpu, tmp_path):
    print("case facade: 9 taxa x 57 sites, K_r=4, K<=3, 5 particles", flush=True)
    codes, batch, names = _case(2)
    fa, pf, newicks = write_inputs(tmp_path, codes, batch, names)
    out = run_facade([fa, pf, "4", "3", "0"])
    lnl = np.array([float(r[1]) for r in out["lnl"]])
    batch_lnl = np.array([float(r[1]) for r in out["batch"]])
    prior = np.array([float(r[1]) for r in out["prior"]])
    want = np.array([orc.particle_lnl(codes, batch.particle(i), 4) for i in range(batch.n)])
    want_prior = np.array([orc.ln_prior(codes, batch.particle(i), 4, using_lg=False)
                           for i in range(batch.n)])
    np.testing.assert_allclose(lnl, want, rtol=1e-10)
    np.testing.assert_array_equal(batch_lnl, lnl)
    np.testing.assert_allclose(prior, want_prior, rtol=1e-12)
    assert [r[1] for r in out["newick"]] == newicks
    temp, ess = (float(v) for v in out["next_step"][0])
    w = np.full(batch.n, -np.log(batch.n))
    _, norm, oess, otemp = orc.compute_next_step(w, want, int(0.6 * batch.n), 0.0)
    assert abs(temp - otemp) <= 1e-5
    assert [int(v) for v in out["resample"][0]] == orc.systematic_resample(norm, 0.5).tolist() or \
        abs(temp - otemp) > 0


def parse_records(path, names):
    """Records written by the self-test (read_particles format) -> ParticleBatch."""
    from smccat import ParticleBatch
    T = len(names)
    recs, cur = [], {}
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            k = parts[0]
            if k == "newick":
                cur = {"newick": parts[1], "logits": []}
            elif k in ("pinv", "shape"):
                cur[k] = float(parts[1])
            elif k == "base":
                cur["base"] = [float(v) for v in parts[1:]]
            elif k == "logits":
                cur["logits"].append([float(v) for v in parts[1:]])
            elif k == "assign":
                cur["assign"] = [int(v) for v in parts[1:]]
            elif k == "end":
                recs.append(cur)
    n, S = len(recs), len(recs[0]["assign"])
    Kmax = max(len(r["logits"]) for r in recs)
    ch = np.zeros((n, T - 1, 2), dtype=np.int32)
    bl = np.zeros((n, 2 * T - 1))
    for i, r in enumerate(recs):
        s, pos, nxt = r["newick"], [0], [T]

        def node():
            if s[pos[0]] == "(":
                pos[0] += 1
                a = node()
                pos[0] += 1  # ','
                b = node()
                pos[0] += 1  # ')'
                me = nxt[0]
                nxt[0] += 1
                ch[i, me - T] = (a, b)
            else:
                j = pos[0]
                while s[pos[0]] not in ":,);":
                    pos[0] += 1
                me = names.index(s[j:pos[0]])
            if s[pos[0]] == ":":
                j = pos[0] + 1
                pos[0] = j
                while s[pos[0]] not in ",);":
                    pos[0] += 1
                bl[i, me] = float(s[j:pos[0]])
            return me
        node()
    lg = np.zeros((n, Kmax, 20))
    for i, r in enumerate(recs):
        lg[i, :len(r["logits"])] = r["logits"]
    return ParticleBatch(ch, bl, np.array([r["base"] for r in recs]),
                         np.array([len(r["logits"]) for r in recs], dtype=np.int32), lg,
                         np.array([r["assign"] for r in recs], dtype=np.int32),
                         np.array([r["pinv"] for r in recs]), np.array([r["shape"] for r in recs]))


@pytest.mark.gpu
def test_facade_moves_roundtrip_oracle(gpu, tmp_path):
    print("case facade moves: Gibbs + rejuvenation through the facade, 9 taxa x 57 sites", flush=True)
    codes, batch, names = _case(3)
    fa, pf, _ = write_inputs(tmp_path, codes, batch, names)
    outf = str(tmp_path / "moved.txt")
    out = run_facade([fa, pf, "4", "3", "0", "--moves", outf])
    assert out["gibbs_return"][0][0] == "inf"
    moved = parse_records(outf, names)
    rj = np.array([[float(v) for v in r[1:]] for r in out["rj"]])
    want = np.array([orc.particle_lnl(codes, moved.particle(i), 4) for i in range(moved.n)])
    want_prior = np.array([orc.ln_prior(codes, moved.particle(i), 4, using_lg=False)
                           for i in range(moved.n)])
    # the re-uploaded records (6-decimal Newick lengths) against the oracle
    np.testing.assert_allclose(rj[:, 1], want, rtol=1e-10)
    np.testing.assert_allclose(rj[:, 2], want_prior, rtol=1e-10)
    # and the engine's own lnL before the text round trip is within its rounding
    np.testing.assert_allclose(rj[:, 0], want, rtol=1e-4)


@pytest.mark.gpu
def test_facade_mcmc_loop_incremental_matches_oracle(gpu, tmp_path):
    """The loop of Mcmc::operator() (Mcmc.cpp:25-73), in its own shape, on facade
    Particles: moves, refreshLikelihood() (incremental), lnPrior, accept() and
    reject().  The incremental lnL of every final state equals a forced refresh
    and the oracle's full recompute (1e-10), on copied records and on one particle
    built by the seeded constructor (Particle.cpp:21-158)."""
    print("case facade mcmc: 9 taxa x 57 sites, K_r=4, 5 records + 1 constructed, 6 iterations", flush=True)
    codes, batch, names = _case(4)
    fa, pf, _ = write_inputs(tmp_path, codes, batch, names)
    out = run_facade([fa, pf, "4", "3", "0", "--mcmc", "6"])
    T, S = codes.shape
    rows = {k: {int(r[0]): r[1:] for r in out[k]} for k in
            ("mcmc_lnl", "mcmc_ch", "mcmc_len", "mcmc_par", "mcmc_base", "mcmc_cls", "mcmc_asg")}
    assert len(rows["mcmc_lnl"]) == batch.n + 1
    acc = rej = 0
    for i in sorted(rows["mcmc_lnl"]):
        inc, full, prior = (float(v) for v in rows["mcmc_lnl"][i][:3])
        acc += int(rows["mcmc_lnl"][i][3])
        rej += int(rows["mcmc_lnl"][i][4])
        cls = [float(v) for v in rows["mcmc_cls"][i]]
        K = int(cls[0])
        q = dict(children=np.array([int(v) for v in rows["mcmc_ch"][i]], dtype=np.int32).reshape(T - 1, 2),
                 branch_len=np.array([float(v) for v in rows["mcmc_len"][i]]),
                 base=np.array([float(v) for v in rows["mcmc_base"][i]]), num_classes=K,
                 logits=np.array(cls[1:]).reshape(K, 20),
                 assignment=np.array([int(v) for v in rows["mcmc_asg"][i]], dtype=np.int32),
                 pinv=float(rows["mcmc_par"][i][0]), shape=float(rows["mcmc_par"][i][1]))
        want = orc.particle_lnl(codes, q, 4)
        assert inc == pytest.approx(full, rel=1e-12), i
        assert inc == pytest.approx(want, rel=1e-10), i
        assert prior == pytest.approx(orc.ln_prior(codes, q, 4, using_lg=False), rel=1e-10), i
    assert acc > 0 and rej > 0, (acc, rej)
