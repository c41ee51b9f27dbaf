// Drives the C++ facade the way the reference's main() drives Particle
// (main.cpp:296-311): copyFromSerialized -> lnLikelihood -> lnPrior ->
// writeToSerialized, for every particle record of a text file, and prints
// one line per result for the test suite to compare with the oracle.
//
// usage: facade_selftest <fasta> <particles.txt> <num_rates> <max_classes> <lg 0|1>
//                        [--cpu | --moves <out.txt>]
// particles.txt: blocks of
//   newick <string> / pinv <v> / shape <v> / base <190 values> / classes <K>
//   logits <20 values> (K lines) / assign <S values> / end
// --cpu: Newick parse/write only (no GPU calls).  --moves: also runs the Gibbs
// move and rejuvenation and writes the resulting records to out.txt.
// --mcmc <iterations>: runs the loop of Mcmc::operator() (Mcmc.cpp:25-73), in its
// own shape, on every record and on one particle built by the seeded constructor,
// through the facade's moves, refreshLikelihood(), accept() and reject(); prints
// each final state with its incremental and forced likelihoods.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <sstream>
#include <tuple>

#include "particle.hpp"

using namespace smccat;

static std::vector<SerializedParticle> read_particles(const std::string &path)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<SerializedParticle> out;
    SerializedParticle cur;
    std::string key;
    size_t next_logit = 0;
    while (f >> key) {
        if (key == "newick") {
            f >> cur.newick;
        } else if (key == "pinv") {
            f >> cur.pInvar;
        } else if (key == "shape") {
            f >> cur.shape;
        } else if (key == "base") {
            for (double &b : cur.baseMatrix) f >> b;
        } else if (key == "classes") {
            int k;
            f >> k;
            cur.stationaries.assign(k, {});
            next_logit = 0;
        } else if (key == "logits") {
            if (next_logit >= cur.stationaries.size()) throw std::runtime_error("too many logits rows");
            for (double &v : cur.stationaries[next_logit]) f >> v;
            next_logit++;
        } else if (key == "assign") {
            std::string line;
            std::getline(f, line);
            std::istringstream ss(line);
            cur.assignments.clear();
            int a;
            while (ss >> a) cur.assignments.push_back(a);
        } else if (key == "end") {
            out.push_back(cur);
            cur = SerializedParticle{};
        }
    }
    return out;
}

// Mcmc::operator() (Mcmc.cpp:25-73) with the move table of main.cpp:54-105,
// written as the reference writes it, on a facade Particle.
struct FacadeMcmc {
    std::vector<std::tuple<double, std::function<int(Particle &)>, std::function<double(Particle &)>>> moves;
    std::vector<double> moveProbs;
    long accepted = 0, rejected = 0;

    void initMoveProbs()
    {
        double tot = 0.0, cum = 0.0;
        for (auto &m : moves) tot += std::get<0>(m);
        for (auto &m : moves) {
            cum += std::get<0>(m);
            moveProbs.push_back(cum / tot);
        }
    }
    void operator()(Particle &model, int iterations, double tempering)
    {
        Rng &rng = model.getRng();
        model.refreshLikelihood();
        double currentLikelihood = model.lnLikelihood() * tempering;
        double currentPrior = model.lnPrior();
        double currentPosterior = currentLikelihood + currentPrior;
        for (int i = 0; i < iterations; i++) {
            const double moveChoice = rng.uniform();
            std::function<double(Particle &)> proposal;
            int numGibbs = 3;
            for (size_t j = 0; j < moveProbs.size(); j++) {
                if (moveChoice < moveProbs[j]) {
                    proposal = std::get<2>(moves[j]);
                    numGibbs = std::get<1>(moves[j])(model);
                    break;
                }
            }
            for (int j = 0; j < numGibbs; j++) {
                const double hastings = proposal(model);
                model.refreshLikelihood();
                const double newPrior = model.lnPrior();
                const double newLikelihood = model.lnLikelihood() * tempering;
                const double newPosterior = newPrior + newLikelihood;
                const double posteriorRatio = newPosterior - currentPosterior + hastings;
                if (std::log(rng.uniform()) <= posteriorRatio) {
                    model.accept();
                    accepted++;
                    currentPosterior = newPosterior;
                    currentPrior = newPrior;
                    currentLikelihood = newLikelihood;
                } else {
                    model.reject();
                    rejected++;
                }
            }
        }
    }
};

static int run_mcmc(const Alignment &aln, std::vector<SerializedParticle> &recs, int R, int K, bool lg,
                    int iterations)
{
    const int T = aln.getNumTaxa();
    const int n = (int)recs.size() + 1;
    ParticleEngine eng(aln, R, n, K, lg, SMC_PREC_FP64, 0, 2);
    const std::map<int, double> splitPosteriorProbabilities;  // adaptive NNI's input (unused)
    FacadeMcmc mcmc;
    mcmc.moves.emplace_back(3.0, [T](Particle &) { return (int)(T * 0.25); },
                            [&](Particle &m) { return m.treeMove(splitPosteriorProbabilities); });
    mcmc.moves.emplace_back(1.0, [](Particle &m) { return m.getNumCategories() * 2; },
                            [](Particle &m) { return m.stationaryMove(); });
    if (R > 1) mcmc.moves.emplace_back(1.0, [](Particle &) { return 2; }, [](Particle &m) { return m.shapeMove(); });
    mcmc.moves.emplace_back(1.0, [](Particle &) { return 2; }, [](Particle &m) { return m.invarMove(); });
    if (!lg) mcmc.moves.emplace_back(1.0, [](Particle &) { return 5; }, [](Particle &m) { return m.rateMatrixMove(); });
    mcmc.initMoveProbs();
    for (int i = 0; i < n; ++i) {
        const bool built = i == (int)recs.size();
        Particle p = built ? Particle(42, eng, i, true, true) : Particle(eng, i);
        if (!built) p.copyFromSerialized(recs[i]);
        const long a0 = mcmc.accepted, r0 = mcmc.rejected;
        mcmc(p, iterations, 0.5);
        const double inc = p.lnLikelihood(), prior = p.lnPrior();
        p.refreshLikelihood(true);
        printf("mcmc_lnl %d %.17g %.17g %.17g %ld %ld\n", i, inc, p.lnLikelihood(), prior, mcmc.accepted - a0,
               mcmc.rejected - r0);
        const ParticleEngine::Slot &s = eng.slots.at(i);
        printf("mcmc_ch %d", i);
        for (int32_t c : s.cur.ch) printf(" %d", c);
        printf("\nmcmc_len %d", i);
        for (double b : s.cur.bl) printf(" %.17g", b);
        printf("\nmcmc_par %d %.17g %.17g\nmcmc_base %d", i, s.cur.pinv, s.cur.shape, i);
        for (double b : s.cur.base) printf(" %.17g", b);
        printf("\nmcmc_cls %d %d", i, p.getNumCategories());
        for (auto &st : p.getCategories())
            for (double v : st) printf(" %.17g", v);
        printf("\nmcmc_asg %d", i);
        for (int a : p.getAssignments()) printf(" %d", a);
        printf("\n");
    }
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 6) {
        fprintf(stderr, "usage: %s fasta particles num_rates max_classes lg [--cpu]\n", argv[0]);
        return 2;
    }
    const bool cpu = argc > 6 && std::strcmp(argv[6], "--cpu") == 0;
    try {
        Alignment aln(argv[1]);
        auto recs = read_particles(argv[2]);
        const int R = std::atoi(argv[3]), K = std::atoi(argv[4]);
        const bool lg = std::atoi(argv[5]) != 0;
        if (argc > 7 && std::strcmp(argv[6], "--mcmc") == 0) return run_mcmc(aln, recs, R, K, lg, std::atoi(argv[7]));
        if (cpu) {
            for (size_t i = 0; i < recs.size(); ++i) {
                TreeTopology t = parseNewick(recs[i].newick, aln.getTaxaNames());
                printf("newick %zu %s\n", i, writeNewick(t).c_str());
            }
            return 0;
        }
        ParticleEngine eng(aln, R, (int)recs.size(), K, lg);
        std::vector<Particle> ps;
        for (size_t i = 0; i < recs.size(); ++i) ps.emplace_back(eng, (int)i);
        for (size_t i = 0; i < recs.size(); ++i) {
            ps[i].copyFromSerialized(recs[i]);
            SerializedParticle back;
            ps[i].writeToSerialized(back);
            printf("lnl %zu %.17g\n", i, ps[i].lnLikelihood());
            printf("prior %zu %.17g\n", i, ps[i].lnPrior());
            printf("newick %zu %s\n", i, back.newick.c_str());
        }
        std::vector<double> all;
        eng.refreshAll((int)recs.size(), all);
        for (size_t i = 0; i < all.size(); ++i) printf("batch %zu %.17g\n", i, all[i]);
        // SMC step on the batch (Miscellaneous.hpp:42-109, main.cpp:179-195)
        std::vector<double> w(all.size(), -std::log((double)all.size())), norm;
        double ess = 0, temp = 0;
        const int target = (int)(0.6 * all.size());
        computeNextStep(eng, w, all, norm, ess, target, temp, 0.0);
        printf("next_step %.17g %.17g\n", temp, ess);
        auto asg = systematicResample(eng, norm, 0.5);
        printf("resample");
        for (int a : asg) printf(" %d", a);
        printf("\n");
        if (argc > 7 && std::strcmp(argv[6], "--moves") == 0) {
            // moves on a second engine with two workspaces: the Gibbs move on slot 0,
            // then rejuvenation of every slot (main.cpp:296-311); the resulting states
            // go through writeToSerialized and back into the first engine
            ParticleEngine eng2(aln, R, (int)recs.size(), K, lg, SMC_PREC_FP64, 0, 2);
            for (size_t i = 0; i < recs.size(); ++i) Particle(eng2, (int)i).copyFromSerialized(recs[i]);
            const double g = Particle(eng2, 0).gibbsPartitionMove(0.8);
            printf("gibbs_return %g\n", g);
            smc_rejuvenation_options o{};
            o.iterations = 3;
            o.using_lg = lg;
            o.invariant = 1;
            o.seed = 11;
            o.scale_delta = o.subtree_scale_delta = o.shape_delta = 1.0;
            o.rate_delta = 0.25;
            o.rate_full_scale = 0.5;
            o.invar_alpha = 100.0;
            o.gibbs_probability = 0.5;
            o.dpp_alpha = 0.01;
            std::vector<double> rl;
            eng2.rejuvenate(0, (int)recs.size(), 0.5, o, rl);
            std::ofstream f(argv[7]);
            for (size_t i = 0; i < recs.size(); ++i) {
                SerializedParticle back;
                Particle(eng2, (int)i).writeToSerialized(back);
                f << "newick " << back.newick << "\npinv ";
                char buf[64];
                snprintf(buf, sizeof buf, "%.17g", back.pInvar);
                f << buf << "\nshape ";
                snprintf(buf, sizeof buf, "%.17g", back.shape);
                f << buf << "\nbase";
                for (double b : back.baseMatrix) {
                    snprintf(buf, sizeof buf, " %.17g", b);
                    f << buf;
                }
                f << "\nclasses " << back.stationaries.size() << "\n";
                for (auto &st : back.stationaries) {
                    f << "logits";
                    for (double v : st) {
                        snprintf(buf, sizeof buf, " %.17g", v);
                        f << buf;
                    }
                    f << "\n";
                }
                f << "assign";
                for (int a : back.assignments) f << " " << a;
                f << "\nend\n";
                Particle q(eng, (int)i);
                q.copyFromSerialized(back);
                printf("rj %zu %.17g %.17g %.17g\n", i, rl[i], q.lnLikelihood(), q.lnPrior());
            }
        }
    } catch (const std::exception &ex) {
        fprintf(stderr, "error: %s\n", ex.what());
        return 1;
    }
    return 0;
}
