// C++ facade that keeps the reference's particle/likelihood call shapes on top
// of the engine's C-ABI (include/smccat_hip.h).
//
//   reference (Smc-CAT)                               here
//   Alignment(path), operator()(t,s)  Alignment.hpp   smccat::Alignment
//   SerializedParticle                SerializedParticle.hpp:8-17   smccat::SerializedParticle
//   Particle::copyFromSerialized / writeToSerialized / refreshLikelihood /
//   lnLikelihood / lnPrior / getters  Particle.hpp:51-83            smccat::Particle
//   computeNextStep, discretizeGamma  Miscellaneous.hpp:22,42       smccat::computeNextStep, ...
//
// A Particle is a view of one slot of a ParticleEngine, which owns the device
// state of every slot; ParticleEngine::refreshAll replaces the OpenMP loop of
// main.cpp:295-312 by one batched sweep.  Like the reference's Particle, a
// view is confined to one host thread.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "smccat_hip.h"

namespace smccat {

// Alignment.cpp:7-31 code table: 0..19 amino acids, 20 '-', 21 'X', 22 '?'.
int alignment_code(char c);

class Alignment {
public:
    explicit Alignment(const std::string &fasta_path);  // Alignment.cpp:6-104
    Alignment(std::vector<std::string> names, std::vector<int8_t> codes, int num_char);
    int getNumTaxa() const { return (int)taxaNames.size(); }
    int getNumChar() const { return numChar; }
    const std::vector<std::string> &getTaxaNames() const { return taxaNames; }
    int operator()(int taxon, int site) const { return codes[(size_t)taxon * numChar + site]; }
    const std::vector<int8_t> &data() const { return codes; }

private:
    std::vector<std::string> taxaNames;
    std::vector<int8_t> codes;
    int numChar = 0;
};

struct SerializedParticle {  // SerializedParticle.hpp:8-17
    std::string newick;
    std::vector<double> branchLengths;
    std::vector<int> assignments;
    std::vector<int> categorySize;
    std::vector<std::array<double, 20>> stationaries;  // logits
    std::array<double, 190> baseMatrix{};
    double pInvar = 0.0;
    double shape = 1.0;
};

// Parsed tree in the engine's terms (node ids as Tree.cpp:62-67,138-157).
struct TreeTopology {
    std::vector<int32_t> children;  // [T-1][2]
    std::vector<double> branchLen;  // [2T-1]
    std::vector<std::string> names; // tip names by id
    int root = -1;
};

// Newick parse with the semantics of Tree::Tree(newick, names) (Tree.cpp:79-166);
// throws std::runtime_error where the reference prints and exits.
TreeTopology parseNewick(const std::string &newick, const std::vector<std::string> &taxaNames);
// Newick write with the semantics of Tree::generateNewick (Tree.cpp:398-437):
// lengths through std::to_string (6 decimals), no length on the root.
std::string writeNewick(const TreeTopology &t);

class ParticleEngine {
public:
    // usingLG: particles carry the fixed LG table (no GTR prior term), as Particle's `lg`.
    // workspaces = 2 keeps two partial buffers per particle (accept/reject, as the
    // reference's CL double buffer) and enables rejuvenate / runSampler.
    ParticleEngine(const Alignment &aln, int numRates, int maxParticles, int maxClasses,
                   bool usingLG = false, int precision = SMC_PREC_FP64, int device = 0,
                   int workspaces = 1);
    ~ParticleEngine();
    ParticleEngine(const ParticleEngine &) = delete;
    ParticleEngine &operator=(const ParticleEngine &) = delete;

    // Uploads slot `slot` from a serialized record (Particle::copyFromSerialized).
    void upload(int slot, const SerializedParticle &sp);
    // Full likelihood of slots [0, n) in one batched sweep; lnl resized to n.
    void refreshAll(int n, std::vector<double> &lnl);
    // mcmc(p, r, T) + optional gibbsPartitionMove for slots [first, first+count)
    // (main.cpp:296-311) in one batched call; lnl resized to count.
    void rejuvenate(int first, int count, double tempering, const smc_rejuvenation_options &opt,
                    std::vector<double> &lnl);
    // The SMC loop of main.cpp:143-336 over slots [0, opt.num_particles).
    smc_sampler_result runSampler(const smc_sampler_options &opt, std::vector<double> &lnl,
                                  std::vector<double> &rawWeights);
    // Re-reads a slot's state from the engine after moves made there.
    void pull(int slot);
    uint64_t nextSeed() { return ++seedCounter; }
    double lnl(int slot) const { return lnls.at(slot); }

    smc_engine *handle() { return eng; }
    const Alignment &alignment() const { return aln; }
    int numRates() const { return rates; }
    int maxClasses() const { return kmax; }
    bool usingLG() const { return lg; }

    // slot state kept on the host for writeToSerialized / lnPrior
    struct Slot {
        SerializedParticle sp;
        TreeTopology tree;
        bool valid = false;
    };
    std::vector<Slot> slots;
    std::vector<double> lnls;

private:
    const Alignment &aln;
    smc_engine *eng = nullptr;
    int rates, kmax;
    bool lg;
    uint64_t seedCounter = 0;
};

class Particle {
public:
    Particle(ParticleEngine &engine, int slot) : e(engine), slot(slot) {}
    void copyFromSerialized(SerializedParticle &sp);   // Particle.cpp:218-300
    void writeToSerialized(SerializedParticle &sp);    // Particle.cpp:302-320
    void refreshLikelihood(bool forceUpdate = false);  // Particle.cpp:530-675
    double lnLikelihood() { return currentLnLikelihood; }  // Particle.cpp:480-482
    double lnPrior();                                  // Particle.cpp:484-517
    // Particle::gibbsPartitionMove (Particle.cpp:785-995) on this slot; returns
    // +infinity as the reference does (the move is always accepted).
    double gibbsPartitionMove(double tempering);
    int getNumCategories() const;
    int getNumNodes() const { return 2 * e.alignment().getNumTaxa() - 1; }
    int getNumRates() const { return e.numRates(); }
    double getPInvar() const;
    double getShape() const;
    std::vector<double> getRates() const;
    std::vector<int> getAssignments() const;
    std::vector<std::array<double, 20>> getCategories() const;
    std::string getNewick() const;

private:
    ParticleEngine &e;
    int slot;
    double currentLnLikelihood = 0.0;
    double dppAlpha = 0.01;  // Particle.hpp:106
};

// Miscellaneous.hpp:42-44, same argument meaning; runs on the GPU of `engine`.
void computeNextStep(ParticleEngine &engine, std::vector<double> &rawWeights,
                     std::vector<double> &rawLogLikelihoods, std::vector<double> &normalizedWeights,
                     double &ESS, const int &targetESS, double &currentTemp, const double &lastTemp);
// main.cpp:179-195 systematic resampling with the uniform `u`.
std::vector<int> systematicResample(ParticleEngine &engine, const std::vector<double> &normalized,
                                    double u);
// Miscellaneous.hpp:22-37.
void discretizeGamma(std::vector<double> &outVec, double shape, int num);

}  // namespace smccat
