// C++ facade over the engine's C-ABI; see particle.hpp for the mapping to the
// reference's classes.
#include "particle.hpp"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <stdexcept>

namespace smccat {

namespace {

void check(smc_engine *e, int rc, const char *what)
{
    if (rc != SMC_OK)
        throw std::runtime_error(std::string(what) + " failed: " +
                                 (e ? smc_engine_last_error(e) : "status " + std::to_string(rc)));
}

}  // namespace

int alignment_code(char c)
{
    static const char *aa = "ARNDCQEGHILKMFPSTWYV";
    const char u = (char)std::toupper((unsigned char)c);
    for (int i = 0; i < 20; ++i)
        if (aa[i] == u) return i;
    if (u == '-') return 20;
    if (u == 'X') return 21;
    if (u == '?') return 22;
    return -1;
}

Alignment::Alignment(const std::string &path)
{
    std::ifstream file(path);
    if (!file.is_open()) throw std::runtime_error("Unable to open alignment file: " + path);
    std::vector<std::vector<int8_t>> seqs;
    std::string line;
    bool in_seq = false;
    while (std::getline(file, line)) {
        if (!line.empty() && line[0] == '#') continue;  // comments
        if (!line.empty() && line[0] == '>') {
            std::string name;
            for (size_t i = 1; i < line.size(); ++i)
                if (std::isprint((unsigned char)line[i])) name += line[i];
            taxaNames.push_back(name);
            seqs.emplace_back();
            in_seq = true;
            continue;
        }
        if (!in_seq) continue;
        for (char c : line) {
            int code = alignment_code(c);
            if (code >= 0)
                seqs.back().push_back((int8_t)code);
            else if (std::isprint((unsigned char)c) && !std::isspace((unsigned char)c))
                std::cerr << "Warning: Detected unknown character in fasta (we are ignoring it): "
                          << c << std::endl;  // Alignment.cpp:62-71 (Q6)
        }
    }
    if (seqs.empty()) throw std::runtime_error("empty alignment: " + path);
    numChar = (int)seqs[0].size();
    for (size_t i = 1; i < seqs.size(); ++i)
        if ((int)seqs[i].size() != numChar)
            throw std::runtime_error("Sequence 1 had length " + std::to_string(numChar) +
                                     " but sequence " + std::to_string(i + 1) + " had length " +
                                     std::to_string(seqs[i].size()));
    codes.reserve(seqs.size() * numChar);
    for (auto &s : seqs) codes.insert(codes.end(), s.begin(), s.end());
}

Alignment::Alignment(std::vector<std::string> names, std::vector<int8_t> c, int num_char)
    : taxaNames(std::move(names)), codes(std::move(c)), numChar(num_char)
{
    if (codes.size() != taxaNames.size() * (size_t)numChar)
        throw std::runtime_error("alignment size mismatch");
}

// ---------------------------------------------------------------- Newick
TreeTopology parseNewick(const std::string &newick, const std::vector<std::string> &taxaNames)
{
    const int T = (int)taxaNames.size();
    // tokenise: ( ) , : ; are single tokens, everything else accumulates (Tree.cpp:377-396)
    std::vector<std::string> toks;
    std::string cur;
    for (char c : newick) {
        if (c == '(' || c == ')' || c == ',' || c == ':' || c == ';') {
            if (!cur.empty()) toks.push_back(cur), cur.clear();
            toks.emplace_back(1, c);
        } else {
            cur += c;
        }
    }
    // build with creation-order node records (Tree.cpp:79-150)
    struct N {
        bool tip;
        int taxon;
        int parent;
        double len;
        std::vector<int> kids;
    };
    std::vector<N> nodes;
    int p = -1, root = -1;
    bool reading_len = false;
    for (std::string tok : toks) {
        if (tok == "(") {
            nodes.push_back(N{false, -1, p, 1.0, {}});
            int id = (int)nodes.size() - 1;
            if (p < 0) {
                root = id;
                nodes[id].len = 0.0;
            } else {
                nodes[p].kids.push_back(id);
            }
            p = id;
        } else if (tok == ")" || tok == ",") {
            if (p < 0 || nodes[p].parent < 0)
                throw std::runtime_error("Poorly formatted Newick! Non-root node has no ancestor!");
            p = nodes[p].parent;
        } else if (tok == ":") {
            reading_len = true;
        } else if (tok == ";") {
            if (p != root) throw std::runtime_error("Poorly formatted Newick! Did not end at root.");
        } else {
            if (reading_len) {
                if (p < 0) throw std::runtime_error("Poorly formatted Newick! Length without node.");
                nodes[p].len = std::atof(tok.c_str());
            } else {
                size_t a = tok.find_first_not_of(' '), b = tok.find_last_not_of(' ');
                std::string name = (a == std::string::npos) ? "" : tok.substr(a, b - a + 1);
                if (!name.empty()) {
                    int taxon = -1;
                    for (int i = 0; i < T; ++i)
                        if (taxaNames[i] == name) taxon = i;
                    if (taxon < 0)
                        throw std::runtime_error("Token '" + name + "' is not in taxa names");
                    if (p < 0) throw std::runtime_error("Poorly formatted Newick! Tip outside a clade.");
                    nodes.push_back(N{true, taxon, p, 1.0, {}});
                    int id = (int)nodes.size() - 1;
                    nodes[p].kids.push_back(id);
                    p = id;
                }
            }
            reading_len = false;
        }
    }
    // ids: tips = taxon index, internal nodes T.. in creation order (Tree.cpp:152-157)
    std::vector<int> id(nodes.size(), -1);
    int next = T, ntips = 0;
    for (size_t i = 0; i < nodes.size(); ++i) {
        if (nodes[i].tip) {
            id[i] = nodes[i].taxon;
            ntips++;
        } else {
            id[i] = next++;
        }
    }
    if (ntips != T || next != 2 * T - 1)
        throw std::runtime_error("Newick tree is not binary over all taxa");
    TreeTopology t;
    t.children.assign((size_t)(T - 1) * 2, -1);
    t.branchLen.assign(2 * T - 1, 0.0);
    t.names = taxaNames;
    for (size_t i = 0; i < nodes.size(); ++i) {
        t.branchLen[id[i]] = nodes[i].len;
        if (!nodes[i].tip) {
            if (nodes[i].kids.size() != 2) throw std::runtime_error("Newick node is not binary");
            t.children[(size_t)(id[i] - T) * 2] = id[nodes[i].kids[0]];
            t.children[(size_t)(id[i] - T) * 2 + 1] = id[nodes[i].kids[1]];
        }
    }
    t.root = id[root];
    t.branchLen[t.root] = 0.0;
    return t;
}

std::string writeNewick(const TreeTopology &t)
{
    const int T = (int)t.names.size();
    std::function<void(int, std::string &)> rec = [&](int n, std::string &s) {
        if (n >= T) {
            s += "(";
            rec(t.children[(size_t)(n - T) * 2], s);
            s += ",";
            rec(t.children[(size_t)(n - T) * 2 + 1], s);
            s += ")";
            if (n != t.root) s += ":" + std::to_string(t.branchLen[n]);
        } else {
            s += t.names[n] + ":" + std::to_string(t.branchLen[n]);
        }
    };
    std::string s;
    rec(t.root, s);
    return s + ";";
}

// ---------------------------------------------------------------- engine
ParticleEngine::ParticleEngine(const Alignment &a, int numRates, int maxParticles, int maxClasses,
                               bool usingLG, int precision, int device, int workspaces)
    : aln(a), rates(numRates), kmax(maxClasses), lg(usingLG)
{
    smc_engine_config cfg{device, a.getNumTaxa(), a.getNumChar(), numRates, maxParticles,
                          maxClasses, precision, 0, 0, workspaces, 0};
    check(nullptr, smc_engine_create(&cfg, &eng), "smc_engine_create");
    check(eng, smc_set_alignment(eng, a.data().data()), "smc_set_alignment");
    slots.resize(maxParticles);
    lnls.assign(maxParticles, 0.0);
}

ParticleEngine::~ParticleEngine()
{
    if (eng) smc_engine_destroy(eng);
}

namespace {

// Sets the slot's current and committed parameters from the engine's copy.
void params_from_engine(smc_engine *eng, int slot, int T, ParticleEngine::Slot &sl)
{
    MoveParams q;
    q.ch.assign((size_t)(T - 1) * 2, 0);
    q.bl.assign(2 * T - 1, 0.0);
    q.base.assign(190, 0.0);
    check(eng, smc_get_particle_params(eng, slot, q.ch.data(), q.bl.data(), q.base.data(), &q.pinv, &q.shape),
          "smc_get_particle_params");
    sl.cur = q;
    sl.committed = q;
    sl.flags.assign(2 * T - 1, 0);
    sl.eig = 0;
    sl.unsent = sl.pending = false;
}

// The tree and record fields of a slot from its current parameters.
void sync_state(ParticleEngine::Slot &sl, int T)
{
    TreeTopology &t = sl.tree;
    t.children = sl.cur.ch;
    t.branchLen = sl.cur.bl;
    std::vector<char> is_child(2 * T - 1, 0);
    for (int32_t c : t.children) is_child[c] = 1;
    for (int n = T; n < 2 * T - 1; ++n)
        if (!is_child[n]) t.root = n;
    t.branchLen[t.root] = 0.0;
    std::copy(sl.cur.base.begin(), sl.cur.base.end(), sl.sp.baseMatrix.begin());
    sl.sp.pInvar = sl.cur.pinv;
    sl.sp.shape = sl.cur.shape;
    sl.sp.newick = writeNewick(t);
}

}  // namespace

void ParticleEngine::upload(int slot, const SerializedParticle &sp)
{
    uploadState(slot, parseNewick(sp.newick, aln.getTaxaNames()), sp);
}

void ParticleEngine::uploadState(int slot, TreeTopology t, const SerializedParticle &sp)
{
    const int T = aln.getNumTaxa(), S = aln.getNumChar();
    const int K = (int)sp.stationaries.size();
    if (K < 1 || K > kmax) throw std::runtime_error("number of classes out of range");
    if ((int)sp.assignments.size() != S) throw std::runtime_error("assignments size != sites");
    std::vector<double> logits((size_t)kmax * 20, 0.0);
    for (int k = 0; k < K; ++k)
        for (int i = 0; i < 20; ++i) logits[(size_t)k * 20 + i] = sp.stationaries[k][i];
    int32_t nc = K;
    double pinv = sp.pInvar, shape = sp.shape;
    (void)T;
    check(eng, smc_set_particles(eng, slot, 1, t.children.data(), t.branchLen.data(),
                                 sp.baseMatrix.data(), &nc, logits.data(), sp.assignments.data(),
                                 &pinv, &shape),
          "smc_set_particles");
    Slot &sl = slots.at(slot);
    sl.sp = sp;
    sl.tree = std::move(t);
    sl.valid = true;
    sl.haveLnl = false;
    params_from_engine(eng, slot, T, sl);
}

void ParticleEngine::refreshAll(int n, std::vector<double> &out)
{
    out.resize(n);
    for (int i = 0; i < n; ++i)
        if (slots.at(i).unsent || slots.at(i).pending)
            throw std::runtime_error("refreshAll: slot " + std::to_string(i) + " holds an unrefreshed proposal");
    check(eng, smc_refresh_likelihood(eng, 0, n, out.data(), nullptr), "smc_refresh_likelihood");
    for (int i = 0; i < n; ++i) {
        lnls[i] = out[i];
        Slot &sl = slots[i];
        sl.committed = sl.cur;  // a forced refresh commits
        sl.lnlCommitted = out[i];
        sl.haveLnl = true;
    }
}

void ParticleEngine::pull(int slot)
{
    Slot &sl = slots.at(slot);
    if (!sl.valid) throw std::runtime_error("slot not uploaded");
    const int T = aln.getNumTaxa(), S = aln.getNumChar();
    TreeTopology &t = sl.tree;
    t.children.assign((size_t)(T - 1) * 2, 0);
    t.branchLen.assign(2 * T - 1, 0.0);
    double pinv = 0.0, shape = 1.0;
    check(eng, smc_get_particle_params(eng, slot, t.children.data(), t.branchLen.data(),
                                       sl.sp.baseMatrix.data(), &pinv, &shape),
          "smc_get_particle_params");
    std::vector<int> parent(2 * T - 1, -1);
    for (int k = 0; k < T - 1; ++k) {
        parent[t.children[2 * k]] = T + k;
        parent[t.children[2 * k + 1]] = T + k;
    }
    for (int n = T; n < 2 * T - 1; ++n)
        if (parent[n] < 0) t.root = n;
    int32_t K = 0;
    std::vector<double> logits((size_t)kmax * 20);
    sl.sp.assignments.assign(S, 0);
    check(eng, smc_get_particle_state(eng, slot, &K, logits.data(), sl.sp.assignments.data()),
          "smc_get_particle_state");
    sl.sp.stationaries.assign(K, {});
    for (int k = 0; k < K; ++k)
        for (int i = 0; i < 20; ++i) sl.sp.stationaries[k][i] = logits[(size_t)k * 20 + i];
    sl.sp.categorySize.assign(K, 0);
    for (int a : sl.sp.assignments) sl.sp.categorySize[a]++;
    sl.sp.pInvar = pinv;
    sl.sp.shape = shape;
    sl.sp.newick = writeNewick(t);
    params_from_engine(eng, slot, T, sl);
    sl.haveLnl = true;  // the batched calls leave every moved slot refreshed and committed
    sl.lnlCommitted = lnls.at(slot);
}

void ParticleEngine::rejuvenate(int first, int count, double tempering, const smc_rejuvenation_options &opt,
                                std::vector<double> &out)
{
    out.resize(count);
    check(eng, smc_rejuvenate(eng, first, count, tempering, &opt, out.data(), nullptr), "smc_rejuvenate");
    for (int i = 0; i < count; ++i) {
        lnls[first + i] = out[i];
        pull(first + i);
    }
}

smc_sampler_result ParticleEngine::runSampler(const smc_sampler_options &opt, std::vector<double> &out,
                                              std::vector<double> &rawWeights)
{
    smc_sampler_result res{};
    out.resize(opt.num_particles);
    rawWeights.resize(opt.num_particles);
    check(eng, smc_run_sampler(eng, &opt, &res, out.data(), rawWeights.data()), "smc_run_sampler");
    for (int i = 0; i < opt.num_particles; ++i) {
        lnls[i] = out[i];
        pull(i);
    }
    return res;
}

// ---------------------------------------------------------------- Particle
void Particle::copyFromSerialized(SerializedParticle &sp)
{
    e.upload(slot, sp);
    e.slots.at(slot).rng = Rng(e.nextSeed(), (uint64_t)slot);
    refreshLikelihood(true);  // Particle.cpp:276
}

void Particle::writeToSerialized(SerializedParticle &sp)
{
    const auto &s = e.slots.at(slot);
    sp = s.sp;
    sp.newick = writeNewick(s.tree);
    // post-order branch lengths without the root (Particle.cpp:310-314)
    sp.branchLengths.clear();
    const int T = (int)s.tree.names.size();
    std::function<void(int)> post = [&](int n) {
        if (n >= T) {
            post(s.tree.children[(size_t)(n - T) * 2]);
            post(s.tree.children[(size_t)(n - T) * 2 + 1]);
        }
        if (n != s.tree.root) sp.branchLengths.push_back(s.tree.branchLen[n]);
    };
    post(s.tree.root);
    sp.categorySize.assign(s.sp.stationaries.size(), 0);
    for (int a : s.sp.assignments) sp.categorySize[a]++;
}

void Particle::sendParams()
{
    ParticleEngine::Slot &s = e.slots.at(slot);
    if (!s.unsent) return;
    check(e.handle(),
          smc_update_params(e.handle(), slot, 1, s.cur.ch.data(), s.cur.bl.data(), s.cur.base.data(), &s.cur.pinv,
                            &s.cur.shape),
          "smc_update_params");
    s.unsent = false;
}

void Particle::refreshLikelihood(bool forceUpdate)
{
    ParticleEngine::Slot &s = e.slots.at(slot);
    if (!s.valid) throw std::runtime_error("refreshLikelihood: slot not uploaded");
    sendParams();
    double l = 0.0;
    uint8_t nf = 0;
    if (forceUpdate || !s.haveLnl) {
        // every node, into the other workspace, and committed (Particle.cpp:566-571)
        check(e.handle(), smc_refresh_likelihood(e.handle(), slot, 1, &l, &nf), "smc_refresh_likelihood");
        s.committed = s.cur;
        s.lnlCommitted = l;
        s.haveLnl = true;
    } else if (s.pending) {
        // the dirty nodes of the proposals since the last refresh (Particle.cpp:577-612)
        check(e.handle(),
              smc_refresh_incremental(e.handle(), slot, 1, s.flags.data(), &s.eig, &l, &nf),
              "smc_refresh_incremental");
    } else {
        l = e.lnls.at(slot);  // nothing changed
    }
    std::fill(s.flags.begin(), s.flags.end(), 0);
    s.eig = 0;
    s.pending = false;
    e.lnls[slot] = l;
}

void Particle::accept()
{
    ParticleEngine::Slot &s = e.slots.at(slot);
    if (s.pending || s.unsent) refreshLikelihood(false);  // an accepted proposal is evaluated first
    const uint8_t yes = 1;
    check(e.handle(), smc_accept_reject(e.handle(), slot, 1, &yes), "smc_accept_reject");
    s.committed = s.cur;
    s.lnlCommitted = e.lnls.at(slot);
}

void Particle::reject()
{
    ParticleEngine::Slot &s = e.slots.at(slot);
    const uint8_t no = 0;
    // the engine returns to its committed flags, parameters, eigensystems and P
    check(e.handle(), smc_accept_reject(e.handle(), slot, 1, &no), "smc_accept_reject");
    s.cur = s.committed;
    std::fill(s.flags.begin(), s.flags.end(), 0);
    s.eig = 0;
    s.unsent = s.pending = false;
    e.lnls[slot] = s.lnlCommitted;
    sync_state(s, e.alignment().getNumTaxa());
}

smc_rejuvenation_options Particle::moveOptions() const
{
    smc_rejuvenation_options o{};
    o.scale_delta = scaleDelta;
    o.subtree_scale_delta = subtreeScaleDelta;
    o.shape_delta = shapeDelta;
    o.rate_delta = rateMatrixDelta;
    o.rate_full_scale = 0.5;  // Particle.hpp:92: Identity * 0.5
    o.invar_alpha = invarAlpha;
    o.rate_cholesky = rateMatrixCholesky.size() == 190 * 190 ? rateMatrixCholesky.data() : nullptr;
    return o;
}

void Particle::proposed(ParticleEngine::Slot &s)
{
    s.unsent = true;
    s.pending = true;
    sync_state(s, e.alignment().getNumTaxa());
}

double Particle::treeMove()
{
    ParticleEngine::Slot &s = e.slots.at(slot);
    const smc_rejuvenation_options o = moveOptions();
    const double h = tree_move(s.rng, s.cur, e.alignment().getNumTaxa(), s.flags, o);
    proposed(s);
    return h;
}

double Particle::shapeMove()
{
    ParticleEngine::Slot &s = e.slots.at(slot);
    if (e.numRates() < 2) return 0.0;  // no rate heterogeneity to move
    const double h = shape_move(s.rng, s.cur, s.flags, shapeDelta);
    proposed(s);
    return h;
}

double Particle::invarMove()
{
    ParticleEngine::Slot &s = e.slots.at(slot);
    const double h = invar_move(s.rng, s.cur, invarAlpha);
    proposed(s);
    return h;
}

double Particle::rateMatrixMove()
{
    ParticleEngine::Slot &s = e.slots.at(slot);
    const smc_rejuvenation_options o = moveOptions();
    const double h = rate_matrix_move(s.rng, s.cur, o);
    s.eig = 1;  // every class is re-decomposed and every P recomputed (Tree::updateAll)
    proposed(s);
    return h;
}

double Particle::lnPrior()
{
    sendParams();  // the prior of the current, possibly proposed, state
    double lp = 0.0;
    check(e.handle(), smc_ln_prior(e.handle(), slot, e.usingLG() ? 1 : 0, dppAlpha, &lp), "smc_ln_prior");
    return lp;
}

Particle::Particle(int seed, ParticleEngine &engine, int sl_index, bool initInvar, bool cat)
    : e(engine), slot(sl_index)
{
    const Alignment &aln = e.alignment();
    const int T = aln.getNumTaxa(), S = aln.getNumChar(), R = e.numRates();
    Rng g((uint64_t)(uint32_t)seed, (uint64_t)slot);
    // random tree (Tree.cpp:10-77): a root over tips 0 and 1, then tip i splits the
    // pendant edge of a uniformly chosen earlier tip; internal node T + i - 1 takes
    // children (i, r) in r's old child slot; every length ~ Exp(rate 10), the root's 0
    TreeTopology t;
    t.names = aln.getTaxaNames();
    t.children.assign((size_t)(T - 1) * 2, -1);
    t.branchLen.assign(2 * T - 1, 0.0);
    std::vector<int> parent(2 * T - 1, -1), pslot(2 * T - 1, 0);
    t.children[0] = 0;
    t.children[1] = 1;
    parent[0] = parent[1] = T;
    pslot[1] = 1;
    for (int i = 2; i < T; ++i) {
        const int r = std::min(i - 1, (int)(g.uniform() * i));
        const int nw = T + i - 1, a = parent[r], sl = pslot[r];
        t.children[(size_t)(a - T) * 2 + sl] = nw;
        parent[nw] = a;
        pslot[nw] = sl;
        t.children[(size_t)(nw - T) * 2] = i;
        t.children[(size_t)(nw - T) * 2 + 1] = r;
        parent[i] = nw;
        pslot[i] = 0;
        parent[r] = nw;
        pslot[r] = 1;
    }
    for (int n = 0; n < 2 * T - 1; ++n) t.branchLen[n] = -std::log1p(-g.uniform()) / 10.0;
    t.root = T;
    t.branchLen[T] = 0.0;
    SerializedParticle sp;
    sp.pInvar = initInvar ? g.uniform() : 0.0;                 // Particle.cpp:117-120
    sp.shape = R > 1 ? 2.0 * g.gamma(3.0) : 1.0;              // Particle.cpp:122-127
    if (e.usingLG()) {                                          // Particle.cpp:84-107
        smc_lg_base_matrix(sp.baseMatrix.data());
    } else {
        for (double &b : sp.baseMatrix) b = g.normal();
    }
    auto profile = [&]() {                                      // TransitionProbabilityClass.cpp:26-35
        std::array<double, 20> lg{};
        for (int j = 0; j < 19; ++j) lg[j] = g.normal();
        return lg;
    };
    sp.assignments.assign(S, 0);
    if (cat) {  // Chinese restaurant process (Particle.cpp:129-158)
        std::vector<int> size;
        for (int i = 0; i < S; ++i) {
            const double u = g.uniform();
            double tot = dppAlpha / (i + dppAlpha);
            if (tot > u) {
                if ((int)size.size() >= e.maxClasses())
                    throw std::runtime_error("Particle: initial classes exceed the engine's max_classes");
                sp.stationaries.push_back(profile());
                sp.assignments[i] = (int)size.size();
                size.push_back(1);
                continue;
            }
            int k = 0;
            for (; k < (int)size.size(); ++k) {
                tot += size[k] / (i + dppAlpha);
                if (tot > u) break;
            }
            k = std::min(k, (int)size.size() - 1);
            sp.assignments[i] = k;
            size[k]++;
        }
    } else {
        sp.stationaries.push_back(profile());
    }
    e.uploadState(slot, std::move(t), sp);
    e.slots.at(slot).rng = g;
    sync_state(e.slots.at(slot), T);
}

double Particle::gibbsPartitionMove(double tempering)
{
    const int32_t sl = slot;
    ParticleEngine::Slot &s = e.slots.at(slot);
    if (s.unsent || s.pending)
        throw std::runtime_error("gibbsPartitionMove: accept or reject the pending proposal first");
    check(e.handle(), smc_gibbs_partition_move(e.handle(), 1, &sl, tempering, e.nextSeed(), 0, dppAlpha, nullptr),
          "smc_gibbs_partition_move");
    const double l = e.lnls.at(slot);
    e.pull(slot);
    e.lnls[slot] = l;
    s.haveLnl = false;  // new classes: the next refresh recomputes every node
    return INFINITY;  // Particle.cpp:994
}

int Particle::getNumCategories() const { return (int)e.slots.at(slot).sp.stationaries.size(); }
double Particle::getPInvar() const { return e.slots.at(slot).sp.pInvar; }
double Particle::getShape() const { return e.slots.at(slot).sp.shape; }
std::vector<int> Particle::getAssignments() const { return e.slots.at(slot).sp.assignments; }
std::vector<std::array<double, 20>> Particle::getCategories() const
{
    return e.slots.at(slot).sp.stationaries;
}
std::string Particle::getNewick() const { return writeNewick(e.slots.at(slot).tree); }
std::vector<double> Particle::getRates() const
{
    std::vector<double> r;
    discretizeGamma(r, getShape(), e.numRates());
    if (e.numRates() == 1) r.assign(1, 1.0);
    return r;
}

// ---------------------------------------------------------------- free functions
void computeNextStep(ParticleEngine &engine, std::vector<double> &rawWeights,
                     std::vector<double> &rawLogLikelihoods, std::vector<double> &normalizedWeights,
                     double &ESS, const int &targetESS, double &currentTemp, const double &lastTemp)
{
    const int n = (int)rawWeights.size();
    normalizedWeights.resize(n);
    check(engine.handle(),
          smc_compute_next_step(engine.handle(), n, rawWeights.data(), rawLogLikelihoods.data(),
                                normalizedWeights.data(), &ESS, targetESS, &currentTemp, lastTemp),
          "smc_compute_next_step");
}

std::vector<int> systematicResample(ParticleEngine &engine, const std::vector<double> &normalized,
                                    double u)
{
    std::vector<int> out(normalized.size());
    check(engine.handle(),
          smc_systematic_resample(engine.handle(), (int)normalized.size(), normalized.data(), u,
                                  out.data()),
          "smc_systematic_resample");
    return out;
}

void discretizeGamma(std::vector<double> &outVec, double shape, int num)
{
    outVec.assign(num, 1.0);
    if (num > 1 && smc_discretize_gamma(shape, num, outVec.data()) != SMC_OK)
        throw std::runtime_error("discretizeGamma: invalid arguments");
}

}  // namespace smccat
