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

#include "moves.hpp"
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
    // Uploads a parsed state: tree (ids as Tree.cpp), classes/assignments/pInvar/shape/base of sp.
    void uploadState(int slot, TreeTopology t, const SerializedParticle &sp);
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

    // slot state kept on the host for writeToSerialized / lnPrior and the MCMC moves
    struct Slot {
        SerializedParticle sp;
        TreeTopology tree;
        bool valid = false;
        // Particle::accept / reject (Particle.cpp:343-478): the current (possibly
        // proposed) parameters and the last committed ones
        MoveParams cur, committed;
        std::vector<uint8_t> flags;  // [2T-1] updateCL | updateTP << 1 since the last refresh
        uint8_t eig = 0;             // the exchangeabilities changed since the last refresh
        bool unsent = false;         // cur is not yet the engine's
        bool pending = false;        // a proposal awaits refreshLikelihood
        bool haveLnl = false;        // the engine holds this slot's likelihood
        double lnlCommitted = 0.0;
        Rng rng{1, 0};               // Particle::getRng (Particle.hpp:64)
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
    // A view of slot `slot` (its state comes from copyFromSerialized or an earlier view).
    Particle(ParticleEngine &engine, int slot) : e(engine), slot(slot) {}
    // Particle::Particle(seed, aln, nR, initInvar, lg, cat) (Particle.cpp:21-158) on an engine
    // slot: the alignment, K_r and lg are the engine's; draws a random tree (Tree.cpp:10-77,
    // Exp(10) lengths), pInvar ~ U(0,1) if initInvar, shape ~ Gamma(3, 2) if K_r > 1, the LG
    // table or 190 N(0,1) log-exchangeabilities, and with cat a CRP(0.01) partition (one class
    // otherwise) of N(0,1)-logit profiles; random numbers from a Philox stream keyed by
    // (seed, slot), where the reference uses Boost's mt19937(seed).
    Particle(int seed, ParticleEngine &engine, int slot, bool initInvar = false, bool cat = true);
    void copyFromSerialized(SerializedParticle &sp);   // Particle.cpp:218-300
    void writeToSerialized(SerializedParticle &sp);    // Particle.cpp:302-320
    // Particle.cpp:530-675: forceUpdate recomputes every node and commits; otherwise only
    // the nodes the proposals since the last refresh made dirty (their other workspace),
    // to be kept by accept() or dropped by reject()
    void refreshLikelihood(bool forceUpdate = false);
    double lnLikelihood() { return e.lnls.at(slot); }  // Particle.cpp:480-482
    double lnPrior();                                  // Particle.cpp:484-517
    void accept();                                     // Particle.cpp:343-403
    void reject();                                     // Particle.cpp:405-478
    // Moves (Particle.cpp:677-783): edit the current state, mark the dirty nodes and
    // return the log Hastings ratio.  treeMove's split posteriors (adaptive NNI) are
    // not used: it proposes a plain NNI in their place.
    double treeMove();
    template <typename SplitPosterior>
    double treeMove(const SplitPosterior &) { return treeMove(); }
    double shapeMove();
    double invarMove();
    double rateMatrixMove();
    double stationaryMove() { return 0.0; }  // a no-op in the reference (Q14)
    Rng &getRng() { return e.slots.at(slot).rng; }
    // tuning fields (Particle.hpp:85-94)
    double shapeDelta = 1.0, scaleDelta = 1.0, subtreeScaleDelta = 1.0, rateMatrixDelta = 0.25;
    double invarAlpha = 100.0;
    std::vector<double> rateMatrixCholesky;  // [190 * 190] row-major lower; empty: 0.5 * identity
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
    double dppAlpha = 0.01;  // Particle.hpp:106
    smc_rejuvenation_options moveOptions() const;
    void sendParams();       // uploads the current parameters if the engine lacks them
    void proposed(ParticleEngine::Slot &s);
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
