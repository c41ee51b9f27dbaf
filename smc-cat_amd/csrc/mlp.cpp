// The learned CAT-allocation proposal: its examples, its training on the GPU
// and the Metropolis-Hastings allocation move that uses it.
//
// There is no counterpart in the reference; the move replaces the Gibbs
// allocation move (Particle::gibbsPartitionMove, Particle.cpp:785-995) as the
// CAT move of rejuvenation (main.cpp:303-305).  It is Neal's (2000) Algorithm 8
// with one auxiliary class in which the exact draw from the site's conditional
// is replaced by one Metropolis-Hastings step targeting that same conditional:
// with the site removed, candidates are the remaining classes c (weight n_c)
// and one auxiliary class (weight alpha) whose profile is the site's own if it
// was a singleton, else a fresh draw from the logit prior
// (TransitionProbabilityClass.cpp:26-35).  The network proposes a candidate
// c' from q(. | site, classes), which does not depend on the site's current
// class, and the step accepts with probability
//   min(1, n_c' L_s(c')^T q(c) / (n_c L_s(c)^T q(c')))
// where L_s is the exact site likelihood (smc_site_class_lnl).  The move keeps
// the DP-CAT posterior invariant for any network; training only changes how
// often it moves.  Training examples are the Gibbs move's own draws: every
// draw of smc_gibbs_move records (site features, classes, chosen slot).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "engine_state.hpp"
#include "host_rng.hpp"
#include "mlp_layout.hpp"

using namespace smccat::mlp;

struct smc_mlp {
    smc_engine *e = nullptr;
    float *d_theta = nullptr, *d_m = nullptr, *d_v = nullptr, *d_grad = nullptr;
    int64_t step = 0;
    // examples (host; uploaded before training)
    std::vector<float> hdr, cls;
    std::vector<int32_t> K, label;
    std::vector<int64_t> coff;
    bool dev_current = false;
    int64_t cap_hdr = 0, cap_K = 0, cap_lab = 0, cap_coff = 0, cap_cls = 0;
    float *d_hdr = nullptr, *d_cls = nullptr;
    int *d_K = nullptr, *d_lab = nullptr, *d_idx = nullptr;
    int64_t *d_coff = nullptr;
    int64_t cap_idx = 0;
    // batch scratch
    int capB = 0;
    float *d_Z1 = nullptr, *d_H1 = nullptr, *d_Z2 = nullptr, *d_H2 = nullptr, *d_dH2 = nullptr;
    float *d_loss = nullptr, *d_part = nullptr, *d_logq = nullptr;
    int64_t cap_logq = 0;
    double *d_lacc = nullptr;
    // allocation-move batch (one example per particle and iteration)
    // per site step, one packed upload (examples, class tables, draw inputs) and one
    // packed readback (proposal log-weights, draws, site scores); see MoveStep
    uint8_t *d_mv_step = nullptr;
    int64_t cap_mv_step = 0;
};

namespace {

template <typename T>
hipError_t grow(T *&p, int64_t &cap, int64_t need)
{
    if (need <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    const int64_t n = std::max<int64_t>(need, cap * 3 / 2);
    hipError_t st = hipMalloc((void **)&p, (size_t)n * sizeof(T));
    cap = st == hipSuccess ? n : 0;
    return st;
}

int ensure_batch(smc_mlp *m, int B)
{
    smc_engine *e = m->e;
    if (B <= m->capB) return SMC_OK;
    float **bufs[] = {&m->d_Z1, &m->d_H1, &m->d_Z2, &m->d_H2, &m->d_dH2};
    for (float **b : bufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
        HIP_TRY(e, hipMalloc((void **)b, (size_t)B * kH * sizeof(float)));
    }
    if (m->d_loss) (void)hipFree(m->d_loss);
    if (m->d_part) (void)hipFree(m->d_part);
    m->d_loss = m->d_part = nullptr;
    HIP_TRY(e, hipMalloc((void **)&m->d_loss, (size_t)B * sizeof(float)));
    const int rows = (B + kRows - 1) / kRows;
    HIP_TRY(e, hipMalloc((void **)&m->d_part, (size_t)rows * kNParam * sizeof(float)));
    m->capB = B;
    return SMC_OK;
}

// Example header and class table of one Algorithm-8 context (as mlp_layout.hpp).
void make_example(const float *freq, const std::vector<std::array<double, 20>> &logits,
                  const std::vector<int> &sizes, double tau, double log_alpha, float *hdr,
                  std::vector<float> &cls)
{
    const int K = (int)logits.size();
    double mean[20] = {};
    double ntot = 0.0;
    for (int c = 0; c < K; ++c) {
        const auto &lg = logits[c];
        double mx = lg[0];
        for (int a = 1; a < 20; ++a) mx = std::max(mx, lg[a]);
        double z = 0.0;
        for (int a = 0; a < 20; ++a) z += std::exp(lg[a] - mx);
        const double lz = mx + std::log(z);
        float ent[kCls] = {};
        for (int a = 0; a < 20; ++a) {
            ent[a] = (float)(lg[a] - lz);
            mean[a] += sizes[c] * std::exp(lg[a] - lz);
        }
        ent[20] = (float)std::log((double)sizes[c]);
        cls.insert(cls.end(), ent, ent + kCls);
        ntot += sizes[c];
    }
    for (int a = 0; a < 20; ++a) {
        hdr[a] = freq[a];
        hdr[20 + a] = (float)(mean[a] / ntot);
    }
    hdr[40] = (float)tau;
    hdr[41] = (float)log_alpha;
    hdr[42] = hdr[43] = 0.f;
}

struct Recorder : AllocationSink {
    smc_mlp *m;
    explicit Recorder(smc_mlp *mm) : m(mm) {}
    void add(int site, const std::vector<std::array<double, 20>> &logits, const std::vector<int> &sizes,
             int label, double tempering, double dpp_alpha) override
    {
        const int K = (int)logits.size();
        if (K < 1 || K >= kMaxCand) return;  // outside what the kernels take
        m->coff.push_back((int64_t)(m->cls.size() / kCls));
        const size_t h = m->hdr.size();
        m->hdr.resize(h + kHdr);
        make_example(&m->e->h_sitefreq[(size_t)site * 20], logits, sizes, tempering, std::log(dpp_alpha),
                     &m->hdr[h], m->cls);
        m->K.push_back(K);
        m->label.push_back(label);
        m->dev_current = false;
    }
};

int upload_examples(smc_mlp *m)
{
    smc_engine *e = m->e;
    if (m->dev_current) return SMC_OK;
    const int64_t n = (int64_t)m->K.size(), nc = (int64_t)(m->cls.size() / kCls);
    HIP_TRY(e, grow(m->d_hdr, m->cap_hdr, std::max<int64_t>(n, 1) * kHdr));
    HIP_TRY(e, grow(m->d_K, m->cap_K, std::max<int64_t>(n, 1)));
    HIP_TRY(e, grow(m->d_lab, m->cap_lab, std::max<int64_t>(n, 1)));
    HIP_TRY(e, grow(m->d_coff, m->cap_coff, std::max<int64_t>(n, 1)));
    HIP_TRY(e, grow(m->d_cls, m->cap_cls, std::max<int64_t>(nc, 1) * kCls));
    if (n > 0) {
        HIP_TRY(e, hipMemcpyAsync(m->d_hdr, m->hdr.data(), (size_t)n * kHdr * 4, hipMemcpyHostToDevice, e->stream));
        HIP_TRY(e, hipMemcpyAsync(m->d_K, m->K.data(), (size_t)n * 4, hipMemcpyHostToDevice, e->stream));
        HIP_TRY(e, hipMemcpyAsync(m->d_lab, m->label.data(), (size_t)n * 4, hipMemcpyHostToDevice, e->stream));
        HIP_TRY(e, hipMemcpyAsync(m->d_coff, m->coff.data(), (size_t)n * 8, hipMemcpyHostToDevice, e->stream));
        HIP_TRY(e, hipMemcpyAsync(m->d_cls, m->cls.data(), (size_t)nc * kCls * 4, hipMemcpyHostToDevice,
                                  e->stream));
        HIP_TRY(e, hipStreamSynchronize(e->stream));
    }
    m->dev_current = true;
    return SMC_OK;
}

MlpArgs batch_args(smc_mlp *m, int B, const int *idx)
{
    MlpArgs a;
    a.B = B;
    a.idx = idx;
    a.hdr = m->d_hdr;
    a.K = m->d_K;
    a.label = m->d_lab;
    a.coff = m->d_coff;
    a.cls = m->d_cls;
    a.theta = m->d_theta;
    a.Z1 = m->d_Z1;
    a.H1 = m->d_H1;
    a.Z2 = m->d_Z2;
    a.H2 = m->d_H2;
    a.dH2 = m->d_dH2;
    a.loss = m->d_loss;
    a.part = m->d_part;
    a.inv_B = 1.f / (float)B;
    return a;
}

// forward + backward of B examples (rows idx or first..first+B), gradient into d_grad
int loss_grad_batch(smc_mlp *m, int B, const int *d_idx, double *d_loss_acc)
{
    int rc = ensure_batch(m, B);
    if (rc != SMC_OK) return rc;
    hipStream_t s = m->e->stream;
    MlpArgs a = batch_args(m, B, d_idx);
    launch_mlp_trunk(a, true, s);
    launch_mlp_head(a, true, s);
    launch_mlp_backward(a, s);
    launch_mlp_reduce((B + kRows - 1) / kRows, m->d_part, m->d_grad, m->d_loss, B, d_loss_acc, s);
    return SMC_OK;
}

int check_range(smc_mlp *m, int first, int n)
{
    if (first < 0 || n < 1 || (int64_t)first + n > (int64_t)m->K.size())
        return smc_fail(m->e, SMC_ERR_ARG, "example range out of bounds");
    return SMC_OK;
}

}  // namespace

extern "C" {

int smc_mlp_num_params(void) { return kNParam; }

int smc_mlp_create(smc_engine *e, uint64_t seed, smc_mlp **out)
{
    if (!e || !out) return SMC_ERR_ARG;
    smc_mlp *m = new smc_mlp();
    m->e = e;
    *out = nullptr;
    auto bad = [&](hipError_t st) {
        if (st == hipSuccess) return false;
        smc_fail(e, SMC_ERR_HIP, std::string("proposal network allocation failed: ") + hipGetErrorString(st));
        return true;
    };
    for (float **b : {&m->d_theta, &m->d_m, &m->d_v, &m->d_grad})
        if (bad(hipMalloc((void **)b, kNParam * sizeof(float)))) {
            smc_mlp_destroy(m);
            return SMC_ERR_HIP;
        }
    if (bad(hipMalloc((void **)&m->d_lacc, sizeof(double)))) {
        smc_mlp_destroy(m);
        return SMC_ERR_HIP;
    }
    // Glorot-uniform weights, zero biases; the new-slot bias starts at 0
    std::vector<float> th(kNParam, 0.f);
    smccat::Rng g(seed, 0x6d6c70ull);
    auto fill = [&](int off, int fan_out, int fan_in) {
        const double b = std::sqrt(6.0 / (fan_in + fan_out));
        for (int i = 0; i < fan_out * fan_in; ++i) th[off + i] = (float)((2.0 * g.uniform() - 1.0) * b);
    };
    fill(kW1, kH, kIn);
    fill(kW2, kH, kH);
    fill(kU, kH, 20);
    fill(kV, kH, 1);
    int rc = smc_mlp_set_params(m, th.data());
    if (rc != SMC_OK) {
        smc_mlp_destroy(m);
        return rc;
    }
    *out = m;
    return SMC_OK;
}

int smc_mlp_destroy(smc_mlp *m)
{
    if (!m) return SMC_OK;
    void *bufs[] = {m->d_theta, m->d_m, m->d_v, m->d_grad, m->d_hdr, m->d_cls, m->d_K, m->d_lab,
                    m->d_idx, m->d_coff, m->d_Z1, m->d_H1, m->d_Z2, m->d_H2, m->d_dH2, m->d_loss,
                    m->d_part, m->d_logq, m->d_lacc, m->d_mv_step};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete m;
    return SMC_OK;
}

int smc_mlp_set_params(smc_mlp *m, const float *theta)
{
    if (!m || !theta) return SMC_ERR_ARG;
    smc_engine *e = m->e;
    HIP_TRY(e, hipMemcpyAsync(m->d_theta, theta, kNParam * sizeof(float), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipMemsetAsync(m->d_m, 0, kNParam * sizeof(float), e->stream));
    HIP_TRY(e, hipMemsetAsync(m->d_v, 0, kNParam * sizeof(float), e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    m->step = 0;
    return SMC_OK;
}

int smc_mlp_get_params(smc_mlp *m, float *theta)
{
    if (!m || !theta) return SMC_ERR_ARG;
    smc_engine *e = m->e;
    HIP_TRY(e, hipMemcpyAsync(theta, m->d_theta, kNParam * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    return SMC_OK;
}

int smc_mlp_add_examples(smc_mlp *m, int n, const float *hdr, const int32_t *K, const int32_t *label,
                         const int64_t *cls_off, const float *cls, int64_t n_cls)
{
    if (!m) return SMC_ERR_ARG;
    if (n < 1 || !hdr || !K || !label || !cls_off || !cls || n_cls < 1)
        return smc_fail(m->e, SMC_ERR_ARG, "invalid examples");
    for (int i = 0; i < n; ++i)
        if (K[i] < 1 || K[i] >= kMaxCand || label[i] < 0 || label[i] > K[i] || cls_off[i] < 0 ||
            cls_off[i] + K[i] > n_cls)
            return smc_fail(m->e, SMC_ERR_ARG, "example " + std::to_string(i) + " out of range");
    const int64_t base = (int64_t)(m->cls.size() / kCls);
    m->hdr.insert(m->hdr.end(), hdr, hdr + (size_t)n * kHdr);
    m->K.insert(m->K.end(), K, K + n);
    m->label.insert(m->label.end(), label, label + n);
    for (int i = 0; i < n; ++i) m->coff.push_back(base + cls_off[i]);
    m->cls.insert(m->cls.end(), cls, cls + (size_t)n_cls * kCls);
    m->dev_current = false;
    return SMC_OK;
}

int64_t smc_mlp_num_examples(smc_mlp *m) { return m ? (int64_t)m->K.size() : -1; }

int smc_mlp_clear_examples(smc_mlp *m)
{
    if (!m) return SMC_ERR_ARG;
    m->hdr.clear();
    m->cls.clear();
    m->K.clear();
    m->label.clear();
    m->coff.clear();
    m->dev_current = false;
    return SMC_OK;
}

int smc_mlp_get_example(smc_mlp *m, int64_t i, float *hdr, int32_t *K, int32_t *label, float *cls)
{
    if (!m) return SMC_ERR_ARG;
    if (i < 0 || i >= (int64_t)m->K.size()) return smc_fail(m->e, SMC_ERR_ARG, "example index out of range");
    if (hdr) std::copy_n(&m->hdr[(size_t)i * kHdr], kHdr, hdr);
    if (K) *K = m->K[i];
    if (label) *label = m->label[i];
    if (cls) std::copy_n(&m->cls[(size_t)m->coff[i] * kCls], (size_t)m->K[i] * kCls, cls);
    return SMC_OK;
}

int smc_mlp_collect_gibbs(smc_mlp *m, int count, const int32_t *slots, double tempering, uint64_t seed,
                          int num_iter, double dpp_alpha, int64_t *stats_out)
{
    if (!m) return SMC_ERR_ARG;
    if (m->e->h_sitefreq.empty()) return smc_fail(m->e, SMC_ERR_STATE, "no alignment");
    Recorder rec(m);
    return smc_gibbs_move(m->e, count, slots, tempering, seed, num_iter, dpp_alpha, stats_out, &rec);
}

int smc_mlp_loss_grad(smc_mlp *m, int first, int n, double *loss_out, float *grad_out)
{
    if (!m) return SMC_ERR_ARG;
    int rc = check_range(m, first, n);
    if (rc == SMC_OK) rc = upload_examples(m);
    if (rc != SMC_OK) return rc;
    smc_engine *e = m->e;
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), first);
    HIP_TRY(e, grow(m->d_idx, m->cap_idx, n));
    HIP_TRY(e, hipMemcpyAsync(m->d_idx, idx.data(), (size_t)n * 4, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipMemsetAsync(m->d_lacc, 0, sizeof(double), e->stream));
    rc = loss_grad_batch(m, n, m->d_idx, m->d_lacc);
    if (rc != SMC_OK) return rc;
    double lacc = 0.0;
    HIP_TRY(e, hipMemcpyAsync(&lacc, m->d_lacc, sizeof(double), hipMemcpyDeviceToHost, e->stream));
    if (grad_out)
        HIP_TRY(e, hipMemcpyAsync(grad_out, m->d_grad, kNParam * sizeof(float), hipMemcpyDeviceToHost,
                                  e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    if (loss_out) *loss_out = lacc / n;
    return SMC_OK;
}

int smc_mlp_infer(smc_mlp *m, int first, int n, int out_stride, float *logq)
{
    if (!m || !logq) return SMC_ERR_ARG;
    int rc = check_range(m, first, n);
    if (rc == SMC_OK) rc = upload_examples(m);
    if (rc != SMC_OK) return rc;
    smc_engine *e = m->e;
    int kmax = 0;
    for (int i = first; i < first + n; ++i) kmax = std::max(kmax, (int)m->K[i]);
    if (out_stride < kmax + 1) return smc_fail(e, SMC_ERR_ARG, "out_stride below K + 1");
    rc = ensure_batch(m, n);
    if (rc != SMC_OK) return rc;
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), first);
    HIP_TRY(e, grow(m->d_idx, m->cap_idx, n));
    HIP_TRY(e, grow(m->d_logq, m->cap_logq, (int64_t)n * out_stride));
    HIP_TRY(e, hipMemcpyAsync(m->d_idx, idx.data(), (size_t)n * 4, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipMemsetAsync(m->d_logq, 0xff, (size_t)n * out_stride * 4, e->stream));  // NaN pad
    MlpArgs a = batch_args(m, n, m->d_idx);
    a.logq = m->d_logq;
    a.out_stride = out_stride;
    launch_mlp_trunk(a, false, e->stream);
    launch_mlp_head(a, false, e->stream);
    HIP_TRY(e, hipGetLastError());
    HIP_TRY(e, hipMemcpyAsync(logq, m->d_logq, (size_t)n * out_stride * 4, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    return SMC_OK;
}

int smc_mlp_train(smc_mlp *m, int epochs, int batch, double lr, uint64_t seed, double *epoch_loss)
{
    if (!m) return SMC_ERR_ARG;
    smc_engine *e = m->e;
    const int64_t N = (int64_t)m->K.size();
    if (epochs < 0 || batch < 1 || !(lr > 0.0) || N < 1 || N > INT32_MAX)
        return smc_fail(e, SMC_ERR_ARG, "invalid training arguments");
    int rc = upload_examples(m);
    if (rc != SMC_OK) return rc;
    const int B = (int)std::min<int64_t>(batch, N);
    rc = ensure_batch(m, B);
    if (rc != SMC_OK) return rc;
    HIP_TRY(e, grow(m->d_idx, m->cap_idx, N));
    std::vector<int> perm(N);
    smccat::Rng g(seed, 0x747261696eull);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ++ep) {
        std::iota(perm.begin(), perm.end(), 0);
        for (int64_t i = N - 1; i > 0; --i) {  // Fisher-Yates
            const int64_t j = std::min<int64_t>(i, (int64_t)(g.uniform() * (double)(i + 1)));
            std::swap(perm[i], perm[j]);
        }
        HIP_TRY(e, hipMemcpyAsync(m->d_idx, perm.data(), (size_t)N * 4, hipMemcpyHostToDevice, e->stream));
        HIP_TRY(e, hipMemsetAsync(m->d_lacc, 0, sizeof(double), e->stream));
        for (int64_t b0 = 0; b0 < N; b0 += B) {
            const int nb = (int)std::min<int64_t>(B, N - b0);
            rc = loss_grad_batch(m, nb, m->d_idx + b0, m->d_lacc);
            if (rc != SMC_OK) return rc;
            m->step++;
            const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)m->step)));
            const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)m->step)));
            launch_mlp_adam(m->d_theta, m->d_m, m->d_v, m->d_grad, (float)lr, b1, b2, eps, c1, c2, e->stream);
        }
        HIP_TRY(e, hipGetLastError());
        double lacc = 0.0;
        HIP_TRY(e, hipMemcpyAsync(&lacc, m->d_lacc, sizeof(double), hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(e, hipStreamSynchronize(e->stream));
        if (epoch_loss) epoch_loss[ep] = lacc / (double)N;
    }
    return SMC_OK;
}

}  // extern "C"

namespace {

struct MvState {
    std::vector<std::array<double, 20>> logits;
    std::vector<std::array<double, 20>> pi;  // softmax of logits
    std::vector<std::array<float, 20>> lpi;  // log pi, as the network reads it
    std::vector<int> orig;                   // engine class slot holding the class's eigensystem
    std::vector<int> free;                   // engine class slots no class uses
    std::vector<int> size;
    std::vector<int32_t> assign;
    double meansum[20];                      // sum over sites of pi of their class
};

void add_class(MvState &ps, const std::array<double, 20> &lg, int orig)
{
    ps.orig.push_back(orig);
    double mx = lg[0];
    for (int a = 1; a < 20; ++a) mx = std::max(mx, lg[a]);
    double z = 0.0;
    for (int a = 0; a < 20; ++a) z += std::exp(lg[a] - mx);
    const double lz = mx + std::log(z);
    std::array<double, 20> pi;
    std::array<float, 20> lpi;
    for (int a = 0; a < 20; ++a) {
        pi[a] = std::exp(lg[a] - lz);
        lpi[a] = (float)(lg[a] - lz);
    }
    ps.logits.push_back(lg);
    ps.pi.push_back(pi);
    ps.lpi.push_back(lpi);
    ps.size.push_back(0);
}

}  // namespace

// Byte layout of one site step's device region: inputs [hdr | K | coff | aux | codes |
// u | class table], then outputs [log q | draws | site scores], each part 256-byte
// aligned.
struct MoveStep {
    size_t o_hdr, o_K, o_coff, o_aux, o_codes, o_u, o_cls, in_bytes;
    size_t o_logq, o_pick, o_ll, out_bytes;
    MoveStep(int count, int kmax, int stride)
    {
        auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
        const size_t n = (size_t)count;
        o_hdr = 0;
        o_K = up(o_hdr + n * kHdr * 4);
        o_coff = up(o_K + n * 4);
        o_aux = up(o_coff + n * 8);
        o_codes = up(o_aux + n * 16);
        o_u = up(o_codes + n * (kmax + 1) * 4);
        o_cls = up(o_u + n * 8);
        in_bytes = up(o_cls + n * kmax * kCls * 4);
        o_logq = 0;
        o_pick = up(n * stride * 4);
        o_ll = up(o_pick + n * 4);
        out_bytes = up(o_ll + n * 2 * 8);
    }
};

extern "C" int smc_learned_partition_move(smc_mlp *m, int count, const int32_t *slots, double tempering,
                                          uint64_t seed, int num_iter, double dpp_alpha,
                                          double proposal_tempering, int64_t *stats_out)
{
    if (!m) return SMC_ERR_ARG;
    smc_engine *e = m->e;
    if (count < 1 || !slots || !(dpp_alpha > 0.0) || !(tempering >= 0.0))
        return smc_fail(e, SMC_ERR_ARG, "invalid allocation move arguments");
    const int S = e->S, Kmax = e->Kmax;
    if (S < 2) return smc_fail(e, SMC_ERR_ARG, "allocation move needs at least two sites");
    if (Kmax + 1 > kMaxCand) return smc_fail(e, SMC_ERR_ARG, "max_classes above the proposal's limit");
    if (e->h_sitefreq.empty()) return smc_fail(e, SMC_ERR_STATE, "no alignment");
    for (int i = 0; i < count; ++i)
        if (slots[i] < 0 || slots[i] >= e->Pmax || !e->h_valid[slots[i]])
            return smc_fail(e, SMC_ERR_ARG, "allocation move: slot not uploaded");
    const int iters = num_iter > 0 ? num_iter : (int)(S / 2.0);  // as the Gibbs move (Particle.cpp:791)
    const double tau_q = proposal_tempering < 0.0 ? tempering : proposal_tempering;
    const double log_alpha = std::log(dpp_alpha);

    std::vector<MvState> st(count);
    std::vector<smccat::Rng> rng;
    rng.reserve(count);
    for (int i = 0; i < count; ++i) {
        const int p = slots[i];
        MvState &ps = st[i];
        const int K = e->h_nclass[p];
        for (int k = 0; k < K; ++k) {
            std::array<double, 20> lg;
            for (int a = 0; a < 20; ++a) lg[a] = e->h_logits[((size_t)p * Kmax + k) * 20 + a];
            add_class(ps, lg, k);
        }
        ps.assign.assign(e->h_assign.begin() + (size_t)p * S, e->h_assign.begin() + (size_t)(p + 1) * S);
        for (int k = Kmax - 1; k >= K; --k) ps.free.push_back(k);
        std::fill(ps.meansum, ps.meansum + 20, 0.0);
        for (int a : ps.assign) {
            ps.size[a]++;
            for (int c = 0; c < 20; ++c) ps.meansum[c] += ps.pi[a][c];
        }
        rng.emplace_back(seed, (uint64_t)p);
    }
    // existing classes are scored with the engine's eigensystems: make them current
    int rc = smc_refresh_eigen(e, count, slots);
    if (rc != SMC_OK) return rc;
    // fresh profiles come from a pool of prior draws decomposed in batches (a
    // stream of their own per slot, consumed in draw order)
    const int pool_q = 16;
    std::vector<smccat::Rng> prof_rng;
    prof_rng.reserve(count);
    for (int i = 0; i < count; ++i) prof_rng.emplace_back(seed ^ 0x70726f66696c65ull, (uint64_t)slots[i]);
    auto draw = [&](int i, double *lg) {
        for (int a = 0; a < 19; ++a) lg[a] = prof_rng[i].normal();
        lg[19] = 0.0;
    };
    rc = smc_pool_refill(e, count, slots, pool_q, draw, 1);
    if (rc != SMC_OK) return rc;
    rc = ensure_batch(m, count);
    if (rc != SMC_OK) return rc;
    const int stride = Kmax + 1;
    // one device region per site step: the inputs (uploaded as one copy; the class
    // table, whose used length varies, last) and then the outputs (read back as one)
    const MoveStep ms(count, Kmax, stride);
    HIP_TRY(e, grow(m->d_mv_step, m->cap_mv_step, (int64_t)(ms.in_bytes + ms.out_bytes)));
    uint8_t *const dst = m->d_mv_step, *const dout = m->d_mv_step + ms.in_bytes;
    std::vector<uint8_t> step_in(ms.in_bytes), step_out(ms.out_bytes);
    rc = ensure_site_class_buffers(e, count, (int64_t)count * 2);
    if (rc != SMC_OK) return rc;

    std::vector<float> hdr((size_t)count * kHdr), cls((size_t)count * Kmax * kCls), logq((size_t)count * stride);
    std::vector<float> logn(S + 1, 0.f);  // log of class sizes
    for (int n = 1; n <= S; ++n) logn[n] = (float)std::log((double)n);
    std::vector<int32_t> Kv(count), site(count), cur(count);
    std::vector<int64_t> coff(count);
    std::vector<std::array<double, 20>> phi(count), phi_new(count);
    std::vector<int> phi_orig(count, -1);
    std::vector<uint8_t> was_single(count);
    std::vector<int32_t> prop(count), aux((size_t)count * 4), codes((size_t)count * (Kmax + 1));
    std::vector<int64_t> fresh_pool(count);
    std::vector<double> uu(count), ll((size_t)count * 2);
    int64_t updates = 0, proposed = 0, accepted = 0, created = 0, erased = 0;

    for (int it = 0; it < iters; ++it) {
        for (int i = 0; i < count; ++i)
            if (e->pool.next[i] >= pool_q) {  // a slot used its pool up: top every slot up
                rc = smc_pool_refill(e, count, slots, pool_q, draw, 0);
                if (rc != SMC_OK) return rc;
                break;
            }
        // remove a random site from its class; build the proposal's example
        for (int i = 0; i < count; ++i) {
            MvState &ps = st[i];
            smccat::Rng &g = rng[i];
            const int s = std::min(S - 1, (int)(g.uniform() * S));
            const int k0 = ps.assign[s];
            site[i] = s;
            for (int c = 0; c < 20; ++c) ps.meansum[c] -= ps.pi[k0][c];
            ps.size[k0]--;
            was_single[i] = ps.size[k0] == 0;
            if (was_single[i]) {  // the site's profile becomes the auxiliary class
                phi[i] = ps.logits[k0];
                phi_orig[i] = ps.orig[k0];
                ps.orig.erase(ps.orig.begin() + k0);
                ps.logits.erase(ps.logits.begin() + k0);
                ps.pi.erase(ps.pi.begin() + k0);
                ps.lpi.erase(ps.lpi.begin() + k0);
                ps.size.erase(ps.size.begin() + k0);
                for (int32_t &a : ps.assign)
                    if (a > k0) a--;
            }
            const int K = (int)ps.size.size();
            ps.assign[s] = -1;
            cur[i] = was_single[i] ? K : k0;
            Kv[i] = K;
            // the proposal's uniform (the device draws with it; the stream order is
            // site, proposal, acceptance, as in a host draw)
            uu[i] = g.uniform();
            int32_t *ax = &aux[(size_t)i * 4];
            ax[0] = slots[i];
            ax[1] = s;
            ax[2] = cur[i];
            ax[3] = was_single[i] ? phi_orig[i] : ps.orig[cur[i]];
            int32_t *cd = &codes[(size_t)i * (Kmax + 1)];
            for (int k = 0; k < K; ++k) cd[k] = ps.orig[k];
            // slot K: the site's own singleton profile (then it is the current class), or
            // the next fresh profile of the slot's pool
            fresh_pool[i] = (int64_t)i * pool_q + e->pool.next[i];
            cd[K] = was_single[i] ? phi_orig[i] : (int32_t)(-2 - fresh_pool[i]);
            coff[i] = i ? coff[i - 1] + Kv[i - 1] : 0;
            float *h = &hdr[(size_t)i * kHdr];
            for (int a = 0; a < 20; ++a) {
                h[a] = e->h_sitefreq[(size_t)s * 20 + a];
                h[20 + a] = (float)(ps.meansum[a] / (double)(S - 1));
            }
            h[40] = (float)tau_q;
            h[41] = (float)log_alpha;
            h[42] = h[43] = 0.f;
            for (int k = 0; k < K; ++k) {  // compact class table: entries coff[i] ..
                float *c = &cls[(size_t)(coff[i] + k) * kCls];
                std::copy(ps.lpi[k].begin(), ps.lpi[k].end(), c);
                c[20] = logn[ps.size[k]];
                c[21] = c[22] = c[23] = 0.f;
            }
        }
        // per-step copies go through the engine's pinned staging and readback buffers;
        // the network, the proposal draw and the exact site likelihoods of the drawn
        // and current classes run back to back, with one wait per site step
        const size_t n_ent = (size_t)(coff[count - 1] + Kv[count - 1]);
        std::memcpy(step_in.data() + ms.o_hdr, hdr.data(), hdr.size() * 4);
        std::memcpy(step_in.data() + ms.o_K, Kv.data(), (size_t)count * 4);
        std::memcpy(step_in.data() + ms.o_coff, coff.data(), (size_t)count * 8);
        std::memcpy(step_in.data() + ms.o_aux, aux.data(), aux.size() * 4);
        std::memcpy(step_in.data() + ms.o_codes, codes.data(), codes.size() * 4);
        std::memcpy(step_in.data() + ms.o_u, uu.data(), uu.size() * 8);
        std::memcpy(step_in.data() + ms.o_cls, cls.data(), n_ent * kCls * 4);
        rc = h2d(e, dst, step_in.data(), ms.o_cls + n_ent * kCls * 4);
        if (rc != SMC_OK) return rc;
        MlpArgs a = batch_args(m, count, nullptr);
        a.hdr = (const float *)(dst + ms.o_hdr);
        a.K = (const int *)(dst + ms.o_K);
        a.coff = (const int64_t *)(dst + ms.o_coff);
        a.cls = (const float *)(dst + ms.o_cls);
        a.logq = (float *)(dout + ms.o_logq);
        a.out_stride = stride;
        launch_mlp_trunk(a, false, e->stream);
        launch_mlp_head(a, false, e->stream);
        MoveDrawArgs da;
        da.count = count;
        da.stride = stride;
        da.kmax = Kmax;
        da.logq = a.logq;
        da.K = a.K;
        da.aux = (const int *)(dst + ms.o_aux);
        da.codes = (const int *)(dst + ms.o_codes);
        da.u = (const double *)(dst + ms.o_u);
        da.pick = (int *)(dout + ms.o_pick);
        da.ent = e->d_sc_ent;
        da.cand_cls = e->d_sc_cls;
        launch_mlp_draw(da, e->stream);
        HIP_TRY(e, hipGetLastError());
        rc = issue_site_class(e, count, 2, true, false, tempering,  // every candidate is coded
                              (double *)(dout + ms.o_ll));
        if (rc == SMC_OK) rc = d2h(e, step_out.data(), dout, ms.out_bytes);
        if (rc == SMC_OK) rc = sync_stream(e);
        if (rc != SMC_OK) return rc;
        std::memcpy(logq.data(), step_out.data() + ms.o_logq, logq.size() * 4);
        std::memcpy(prop.data(), step_out.data() + ms.o_pick, (size_t)count * 4);
        std::memcpy(ll.data(), step_out.data() + ms.o_ll, ll.size() * 8);
        for (int i = 0; i < count; ++i) {
            updates++;
            if (prop[i] == cur[i]) continue;
            proposed++;
            if (prop[i] == Kv[i]) {  // a fresh profile from the logit prior (the pool) is used up
                const int slot = e->pool.next[i]++;
                std::copy_n(&e->pool.logits[((size_t)i * pool_q + slot) * 20], 20, phi_new[i].begin());
            }
        }
        // accept or reject, then put the site back
        for (int i = 0; i < count; ++i) {
            MvState &ps = st[i];
            smccat::Rng &g = rng[i];
            const int K = Kv[i], s = site[i];
            int dest = cur[i];
            bool fresh = false;
            if (prop[i] != cur[i]) {
                const int c0 = cur[i], c1 = prop[i];
                const float *lq = &logq[(size_t)i * stride];
                auto logw = [&](int c) { return c == K ? log_alpha : std::log((double)ps.size[c]); };
                const double ratio = ll[2 * i + 1] - ll[2 * i] + logw(c1) - logw(c0) + (double)lq[c0] -
                                     (double)lq[c1];
                if (std::log(g.uniform()) <= ratio) {
                    accepted++;
                    dest = c1;
                    fresh = c1 == K;
                }
            }
            if (dest == K) {  // the auxiliary class becomes a real class
                int slot_k = phi_orig[i];
                if (fresh) {  // a pool profile: its eigensystem moves into a free class slot
                    slot_k = ps.free.back();
                    ps.free.pop_back();
                    rc = smc_pool_to_class(e, slots[i], (int)fresh_pool[i], slot_k);
                    if (rc != SMC_OK) return rc;
                    created++;
                }
                add_class(ps, fresh ? phi_new[i] : phi[i], slot_k);
            } else if (was_single[i]) {
                ps.free.push_back(phi_orig[i]);
                erased++;
            }
            ps.assign[s] = dest;
            ps.size[dest]++;
            for (int c = 0; c < 20; ++c) ps.meansum[c] += ps.pi[dest][c];
        }
    }
    std::vector<double> lg((size_t)Kmax * 20);
    for (int i = 0; i < count; ++i) {
        std::fill(lg.begin(), lg.end(), 0.0);
        const int K = (int)st[i].size.size();
        for (int k = 0; k < K; ++k) std::copy(st[i].logits[k].begin(), st[i].logits[k].end(), &lg[(size_t)k * 20]);
        rc = smc_upload_partition(e, slots[i], K, lg.data(), st[i].assign.data());
        if (rc != SMC_OK) return rc;
    }
    if (stats_out) {
        stats_out[0] = updates;
        stats_out[1] = proposed;
        stats_out[2] = accepted;
        stats_out[3] = created;
        stats_out[4] = erased;
    }
    return SMC_OK;
}

