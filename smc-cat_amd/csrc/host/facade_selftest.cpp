// Drives the C++ facade the way the reference's main() drives Particle
// (main.cpp:296-311): copyFromSerialized -> lnLikelihood -> lnPrior ->
// writeToSerialized, for every particle record of a text file, and prints
// one line per result for the test suite to compare with the oracle.
//
// usage: facade_selftest <fasta> <particles.txt> <num_rates> <max_classes> <lg 0|1> [--cpu]
// particles.txt: blocks of
//   newick <string> / pinv <v> / shape <v> / base <190 values> / classes <K>
//   logits <20 values> (K lines) / assign <S values> / end
// --cpu: Newick parse/write and lnPrior only (no GPU calls).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

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
    } catch (const std::exception &ex) {
        fprintf(stderr, "error: %s\n", ex.what());
        return 1;
    }
    return 0;
}
