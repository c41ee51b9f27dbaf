// Substitution-model kernels: batched eigen-decomposition of the reversible
// 20x20 rate matrix and the P(t) back-transform.
//
// Reference: TransitionProbabilityClass::recomputeEigens / recomputeTransitionProbs
// (TransitionProbabilityClass.cpp:50-95).  The reference runs a general complex
// eigensolver on Q and inverts V by LU.  Q = R diag(pi) / mu is reversible, so
// here the symmetric matrix S = Pi^{1/2} Q Pi^{-1/2} is diagonalised with a
// parallel cyclic Jacobi method (one wavefront per matrix, 10 disjoint rotations
// per round, round-robin ordering), giving Q = A diag(lam) B with
// A = Pi^{-1/2} U and B = U^T Pi^{1/2}: all real, FP64.
#include "internal.hpp"

namespace smccat {

namespace {

__device__ __forceinline__ int tri_index(int i, int j)
{
    // 190-vector index of the unordered pair (i, j), i != j (RateMatrices.hpp:97-111)
    return i > j ? i * (i - 1) / 2 + j : j * (j - 1) / 2 + i;
}

// Round-robin (circle method) pairing: in round r (0..18), pair k (0..9).
__device__ __forceinline__ void rr_pair(int r, int k, int &p, int &q)
{
    int a, b;
    if (k == 0) {
        a = r;
        b = 19;
    } else {
        a = (r + k) % 19;
        b = (r - k + 19) % 19;
    }
    p = a < b ? a : b;
    q = a < b ? b : a;
}

// Eigen-decomposition of one class by the calling 64-lane block: b = the 190
// log-exchangeabilities, lg = 20 logits, out = A | lam | B | pi (840 doubles).
__device__ void eigen_block(const double *__restrict__ b, const double *__restrict__ lg,
                            double *__restrict__ out)
{
    __shared__ double S[20][21];
    __shared__ double U[20][21];
    __shared__ double pi_s[20], sp_s[20], diag_s[20];
    __shared__ double rc[10], rs[10];
    __shared__ int rp[10], rq[10];
    __shared__ double red[64];

    const int lane = threadIdx.x;

    // pi = softmax(logits), TransitionProbabilityClass.cpp:101-109
    if (lane < 20) pi_s[lane] = exp(lg[lane]);
    __syncthreads();
    double tot = 0.0;
    for (int i = 0; i < 20; i++) tot += pi_s[i];
    __syncthreads();
    if (lane < 20) {
        pi_s[lane] = pi_s[lane] / tot;
        sp_s[lane] = sqrt(pi_s[lane]);
    }
    __syncthreads();

    // Q_ij = exp(b_c) pi_j (TransitionProbabilityClass.cpp:52-60)
    for (int idx = lane; idx < 400; idx += 64) {
        int i = idx / 20, j = idx % 20;
        S[i][j] = (i == j) ? 0.0 : exp(b[tri_index(i, j)]) * pi_s[j];
    }
    __syncthreads();
    if (lane < 20) {  // Q_ii = -sum_{j != i} Q_ij (:62-68)
        double off = 0.0;
        for (int j = 0; j < 20; j++)
            if (j != lane) off += S[lane][j];
        diag_s[lane] = -off;
    }
    __syncthreads();
    double mu = 0.0;  // mean rate (:70-75)
    for (int i = 0; i < 20; i++) mu += -diag_s[i] * pi_s[i];
    // symmetric form of Q / mu
    double sv[7];
#pragma unroll
    for (int m = 0; m < 7; m++) {
        int idx = lane + 64 * m;
        if (idx < 400) {
            int i = idx / 20, j = idx % 20;
            if (i == j) {
                sv[m] = diag_s[i] / mu;
            } else {
                double qij = S[i][j] / mu, qji = S[j][i] / mu;
                sv[m] = 0.5 * (sp_s[i] * qij / sp_s[j] + sp_s[j] * qji / sp_s[i]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 7; m++) {
        int idx = lane + 64 * m;
        if (idx < 400) {
            int i = idx / 20, j = idx % 20;
            S[i][j] = sv[m];
            U[i][j] = (i == j) ? 1.0 : 0.0;
        }
    }
    __syncthreads();

    // parallel cyclic Jacobi (Golub & Van Loan, Alg. 8.4.2 rotations; round-robin order)
    for (int sweep = 0; sweep < 40; sweep++) {
        double off = 0.0, all = 0.0;
        for (int idx = lane; idx < 400; idx += 64) {
            int i = idx / 20, j = idx % 20;
            double v = S[i][j] * S[i][j];
            all += v;
            if (i != j) off += v;
        }
        red[lane] = off;
        __syncthreads();
        double off_t = 0.0;
        for (int i = 0; i < 64; i++) off_t += red[i];
        __syncthreads();
        red[lane] = all;
        __syncthreads();
        double all_t = 0.0;
        for (int i = 0; i < 64; i++) all_t += red[i];
        __syncthreads();
        // converged: the off-diagonal mass is at the rounding floor (about 1e-32 of
        // the total after the sixth or seventh sweep; a tighter bound is never met
        // and would run every sweep)
        if (off_t <= 1e-30 * all_t) break;

        for (int r = 0; r < 19; r++) {
            if (lane < 10) {
                int pp, qq;
                rr_pair(r, lane, pp, qq);
                double apq = S[pp][qq];
                double c = 1.0, s = 0.0;
                if (apq != 0.0) {
                    double tau = (S[qq][qq] - S[pp][pp]) / (2.0 * apq);
                    double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                    c = 1.0 / sqrt(1.0 + t * t);
                    s = t * c;
                }
                rc[lane] = c;
                rs[lane] = s;
                rp[lane] = pp;
                rq[lane] = qq;
            }
            __syncthreads();
            for (int t = lane; t < 200; t += 64) {  // S <- S J, U <- U J
                int kk = t / 20, i = t % 20;
                int pp = rp[kk], qq = rq[kk];
                double c = rc[kk], s = rs[kk];
                double a0 = S[i][pp], a1 = S[i][qq];
                S[i][pp] = c * a0 - s * a1;
                S[i][qq] = s * a0 + c * a1;
                double u0 = U[i][pp], u1 = U[i][qq];
                U[i][pp] = c * u0 - s * u1;
                U[i][qq] = s * u0 + c * u1;
            }
            __syncthreads();
            for (int t = lane; t < 200; t += 64) {  // S <- J^T S
                int kk = t / 20, j = t % 20;
                int pp = rp[kk], qq = rq[kk];
                double c = rc[kk], s = rs[kk];
                double a0 = S[pp][j], a1 = S[qq][j];
                S[pp][j] = c * a0 - s * a1;
                S[qq][j] = s * a0 + c * a1;
            }
            __syncthreads();
        }
    }

    for (int idx = lane; idx < 400; idx += 64) {
        int i = idx / 20, kk = idx % 20;
        out[i * 20 + kk] = U[i][kk] / sp_s[i];          // A = Pi^{-1/2} U
        out[kEigB + kk * 20 + i] = U[i][kk] * sp_s[i];   // B = U^T Pi^{1/2}
    }
    if (lane < 20) {
        out[kEigLam + lane] = S[lane][lane];
        out[kEigPi + lane] = pi_s[lane];
    }
}

__global__ __launch_bounds__(64) void k_eigen_f64(const double *__restrict__ base,
                                                   const double *__restrict__ logits,
                                                   const int *__restrict__ num_classes,
                                                   double *__restrict__ eig, int max_classes,
                                                   int first, const int *__restrict__ plist)
{
    const int p = plist ? plist[blockIdx.x / max_classes] : first + blockIdx.x / max_classes;
    const int k = blockIdx.x % max_classes;
    if (k >= num_classes[p]) return;  // whole block exits together
    const size_t g = (size_t)p * max_classes + k;
    eigen_block(base + (size_t)p * 190, logits + g * 20, eig + g * kEigStride);
}

// Candidate profiles: block e * n_cand + c decomposes logits[e][c] with the base
// matrix of particle ent_particle[e], or, when cand_cls[e][c] >= 0, copies the
// eigensystem the engine already holds for that class of the particle.
__global__ __launch_bounds__(64) void k_eigen_cand(const double *__restrict__ base,
                                                    const int *__restrict__ ent_particle,
                                                    const int *__restrict__ ent_ncand,
                                                    const double *__restrict__ logits, int n_cand,
                                                    double *__restrict__ eig,
                                                    const int *__restrict__ cand_cls,
                                                    const double *__restrict__ eig_cache, int max_classes,
                                                    const double *__restrict__ pool, int copy_known)
{
    const int e = blockIdx.x / n_cand, c = blockIdx.x % n_cand;
    if (c >= ent_ncand[e]) return;
    const int k = cand_cls ? cand_cls[blockIdx.x] : -1;
    if (k >= 0 || k <= -2) {  // block-uniform: a copy of a decomposition made earlier
        if (!copy_known) return;  // read where it lies
        const double *src = k >= 0 ? eig_cache + ((size_t)ent_particle[e] * max_classes + k) * kEigStride
                                   : pool + (size_t)(-2 - k) * kEigStride;
        double *dst = eig + (size_t)blockIdx.x * kEigStride;
        for (int i = threadIdx.x; i < kEigStride; i += 64) dst[i] = src[i];
        return;
    }
    eigen_block(base + (size_t)ent_particle[e] * 190, logits + (size_t)blockIdx.x * 20,
                eig + (size_t)blockIdx.x * kEigStride);
}

// P(t) = A diag(e^{lam t r}) B of one class for a run of (node, rate) pairs,
// written row-major with a 21st column of row sums (P 1 = A diag(e) (B 1)).  P is
// reversible, pi_i P_ij = pi_j P_ji (A = Pi^-1/2 U, B = U^T Pi^1/2), so only the
// upper triangle (190 entries), the diagonal (20) and the row sums (20) are sums over
// k; the lower triangle is (pi_i / pi_j) P_ij at (j, i).  Each entry e (upper triangle
// in row order, then the diagonal, then the row sums) is the dot product of its
// outer-product column M[e][k] = A[i][k] Bx[k][j] (Bx: B with the row-sum column)
// with the pair's exponentials e_k = exp(lam_k t r).  Long runs (256 pairs; 128 measured
// 7% slower, the per-workgroup set-up being shared by fewer pairs): lane l of
// a pair of wavefronts keeps M of entries l and l + 64 (even wavefront) or 128 + l and
// 192 + l (odd one) in registers, so that two wavefronts cover the 230 entries of a
// (node, rate) pair and a step takes two pairs; an exponential read from LDS feeds two
// multiply-adds (LDS bandwidth bounds the one-entry form on long runs; four entries per
// lane, one wavefront per pair, ran 14% slower at two waves per SIMD).  Short runs (32
// pairs) give lane t entry t and two pairs a step: their per-workgroup set-up of M would
// cost twice as much in the two-entry form.  The entries of a step's pairs go through
// LDS (two buffers, one barrier a step) and leave in 16-byte consecutive-lane stores of
// the whole workgroup (measured faster than waves that run their own pairs with no
// barrier and store their own rows).  Node lists (a few pairs per workgroup) store each
// entry and its mirror straight from the lane: there the barriers cost more than the
// coalescing gains.  The exponentials of all pairs of the run are computed once, spread
// over the workgroup.
constexpr int kPmLanes = 256;
constexpr int kPmPairs = 32;       // (node, rate) pairs per workgroup
constexpr int kPmPairsLarge = 256;  // the same for launches with workgroups to spare
constexpr int64_t kPmLargeGrid = 32768;  // workgroups at 256 pairs from which launches use them
constexpr int kPmUpper = 190, kPmEntries = 230;
// forms: two entries per lane, staged; one entry per lane, staged; one entry, direct
constexpr int kPmTwo = 0, kPmOneStaged = 1, kPmOneDirect = 2;

// entry e of the list above -> (i, j); j = 20 for the row sums
__device__ __forceinline__ void pmat_entry(int e, int &i, int &j)
{
    if (e < kPmUpper) {
        i = 0;
        while (e >= 19 - i) {
            e -= 19 - i;
            i++;
        }
        j = i + 1 + e;
    } else if (e < kPmUpper + 20) {
        i = j = e - kPmUpper;
    } else if (e < kPmEntries) {
        i = e - kPmUpper - 20;
        j = 20;
    } else {
        i = j = 0;
    }
}

// node_of(i): node of the run's i-th node; pair_out(q): index n * R + r of pair q's
// matrix in the class's P block
template <typename TP, int PAIRS, int FORM, typename NodeOf, typename PairOut>
__device__ __forceinline__ void pmat_pairs(const double *__restrict__ eg, const double *__restrict__ bl,
                                           const double *__restrict__ rt, int R, int q0, int q1,
                                           NodeOf node_of, PairOut pair_out, TP *__restrict__ Pcls)
{
    static_assert(FORM == kPmTwo || FORM == kPmOneStaged || FORM == kPmOneDirect, "P form");
    constexpr int V = 16 / (int)sizeof(TP);  // entries per 16-byte store
    constexpr int NV = kPStride / V;
    static_assert(kPStride % V == 0, "P rows must split into 16-byte stores");
    __shared__ double bsum[20];
    __shared__ double ex[PAIRS][20];
    const int tid = threadIdx.x;
    if (tid < 20) {
        double a = 0.0;
        for (int j = 0; j < 20; j++) a += eg[kEigB + tid * 20 + j];
        bsum[tid] = a;
    }
    // the exponentials of every pair of the run, spread over the workgroup
    for (int x = tid; x < (q1 - q0) * 20; x += kPmLanes) {
        const int q = q0 + x / 20, k = x % 20;
        ex[x / 20][k] = exp(eg[kEigLam + k] * bl[node_of(q / R)] * rt[q % R]);
    }
    __syncthreads();
    // staged forms: the step's entries to LDS, then 16-byte runs of the step's matrices out
    constexpr int PS = 2;  // pairs per step
    __shared__ __align__(16) TP stage[FORM == kPmOneDirect ? 1 : 2][FORM == kPmOneDirect ? 1 : PS * kPStride];
    auto store_step = [&](const TP *st, int q, int nq) {
        __syncthreads();
        for (int x = tid; x < nq * NV; x += kPmLanes) {
            const int u = x / NV, v = x - u * NV;
            *(uint4 *)(Pcls + pair_out(q + u) * kPStride + v * V) = *(const uint4 *)(st + u * kPStride + v * V);
        }
    };
    int buf = 0;
    if constexpr (FORM == kPmTwo) {
        // wavefronts 2 p' and 2 p' + 1 take pair p' of the step: lane l of the even one
        // entries l and l + 64 (all in the upper triangle), of the odd one entries
        // 128 + l and 192 + l (the rest of the upper triangle, the diagonal, the row
        // sums, nothing from 230 on)
        const int lane = tid & 63, wave = tid >> 6, sub = wave & 1, pw = wave >> 1;
        int ei[2], ej[2];
        bool up[2], has[2];
        double m[2][20], ratio[2];
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const int e = 128 * sub + lane + 64 * c;
            pmat_entry(e, ei[c], ej[c]);
            up[c] = e < kPmUpper;
            has[c] = e < kPmEntries;
#pragma unroll
            for (int k = 0; k < 20; k++)
                m[c][k] = eg[ei[c] * 20 + k] * (ej[c] < 20 ? eg[kEigB + k * 20 + ej[c]] : bsum[k]);
            ratio[c] = eg[kEigPi + ei[c]] / eg[kEigPi + ej[c] % 20];
        }
        for (int q = q0; q < q1; q += PS, buf ^= 1) {
            const int nq = min(PS, q1 - q);
            TP *st = stage[buf];
            if (pw < nq) {
                const double *x = ex[q - q0 + pw];
                double acc[2] = {0.0, 0.0};
#pragma unroll
                for (int k = 0; k < 20; k++) {
                    const double e = x[k];
#pragma unroll
                    for (int c = 0; c < 2; c++) acc[c] = fma(m[c][k], e, acc[c]);
                }
                TP *sp = st + pw * kPStride;
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    if (has[c]) sp[ei[c] * 21 + ej[c]] = (TP)acc[c];
                    if (up[c]) sp[ej[c] * 21 + ei[c]] = (TP)(acc[c] * ratio[c]);
                }
            }
            store_step(st, q, nq);
        }
    } else if constexpr (FORM == kPmOneStaged) {
        // entry tid (none from 230 on), two pairs a step as independent chains
        int i, j;
        pmat_entry(tid, i, j);
        const bool up = tid < kPmUpper, has = tid < kPmEntries;
        double m[20];
#pragma unroll
        for (int k = 0; k < 20; k++) m[k] = eg[i * 20 + k] * (j < 20 ? eg[kEigB + k * 20 + j] : bsum[k]);
        const double ratio = eg[kEigPi + i] / eg[kEigPi + j % 20];
        for (int q = q0; q < q1; q += PS, buf ^= 1) {
            const int nq = min(PS, q1 - q);
            TP *st = stage[buf];
            double acc[PS];
#pragma unroll
            for (int u = 0; u < PS; u++) acc[u] = 0.0;
#pragma unroll
            for (int k = 0; k < 20; k++)
#pragma unroll
                for (int u = 0; u < PS; u++) acc[u] = fma(m[k], ex[min(q - q0 + u, q1 - q0 - 1)][k], acc[u]);
            if (has) {
#pragma unroll
                for (int u = 0; u < PS; u++) {
                    if (u < nq) {
                        st[u * kPStride + i * 21 + j] = (TP)acc[u];
                        if (up) st[u * kPStride + j * 21 + i] = (TP)(acc[u] * ratio);
                    }
                }
            }
            store_step(st, q, nq);
        }
    } else {
        // entry tid (none from 230 on), two pairs a step as independent chains, stored
        // straight from the lanes (node lists: staging through LDS measured slower on
        // these runs of a few pairs, whose barriers cost more than the coalescing gains)
        int i, j;
        pmat_entry(tid, i, j);
        const bool up = tid < kPmUpper;
        if (tid >= kPmEntries) return;  // no barrier follows
        double m[20];
#pragma unroll
        for (int k = 0; k < 20; k++) m[k] = eg[i * 20 + k] * (j < 20 ? eg[kEigB + k * 20 + j] : bsum[k]);
        const double ratio = eg[kEigPi + i] / eg[kEigPi + j % 20];
        auto put = [&](int q, double pa) {
            TP *out = Pcls + pair_out(q) * kPStride;
            out[i * 21 + j] = (TP)pa;
            if (up) out[j * 21 + i] = (TP)(pa * ratio);
        };
        int q = q0;
        for (; q + 2 <= q1; q += 2) {
            double pa = 0.0, pb = 0.0;
#pragma unroll
            for (int k = 0; k < 20; k++) {
                pa = fma(m[k], ex[q - q0][k], pa);
                pb = fma(m[k], ex[q - q0 + 1][k], pb);
            }
            put(q, pa);
            put(q + 1, pb);
        }
        if (q < q1) {
            double pa = 0.0;
#pragma unroll
            for (int k = 0; k < 20; k++) pa = fma(m[k], ex[q - q0][k], pa);
            put(q, pa);
        }
    }
}

// Every (node, rate) of every class of particles first.. (or plist[..]):
// workgroup (particle * max_classes + class, chunk of PAIRS pairs).  The per-class
// setup of a workgroup (row sums, per-entry products) is shared by its PAIRS pairs:
// launches with many workgroups take long runs, small ones short runs for parallelism.
template <typename TP, int PAIRS, int FORM>
__global__ __launch_bounds__(kPmLanes) void k_pmat(const double *__restrict__ eig,
                                                    const double *__restrict__ blen,
                                                    const double *__restrict__ rates,
                                                    const int *__restrict__ num_classes,
                                                    const int *__restrict__ pcls_off,
                                                    TP *__restrict__ Pout, int nn, int R,
                                                    int max_classes, int first,
                                                    const int *__restrict__ plist)
{
    const int p = plist ? plist[blockIdx.x / max_classes] : first + blockIdx.x / max_classes;
    const int k = blockIdx.x % max_classes;
    if (k >= num_classes[p]) return;  // whole workgroup
    const int q0 = blockIdx.y * PAIRS, q1 = min(q0 + PAIRS, nn * R);
    if (q0 >= q1) return;
    const size_t g = (size_t)p * max_classes + k;
    pmat_pairs<TP, PAIRS, FORM>(eig + g * kEigStride, blen + (size_t)p * nn, rates + (size_t)p * R, R, q0, q1,
                          [](int n) { return n; }, [](int q) { return (size_t)q; }, Pout + ((size_t)pcls_off[p] + k) * nn * R * kPStride);
}

// P of listed nodes only: workgroup (listed particle * max_classes + class, chunk
// of kPmPairs of its (node, rate) pairs).
template <typename TP>
__global__ __launch_bounds__(kPmLanes) void k_pmat_nodes(const double *__restrict__ eig,
                                                          const double *__restrict__ blen,
                                                          const double *__restrict__ rates,
                                                          const int *__restrict__ num_classes,
                                                          const int *__restrict__ pcls_off,
                                                          TP *__restrict__ Pout, int nn, int R, int max_classes,
                                                          const int *__restrict__ plist,
                                                          const int *__restrict__ noff,
                                                          const int *__restrict__ nodes)
{
    const int pi = blockIdx.x / max_classes, k = blockIdx.x % max_classes;
    const int p = plist[pi];
    if (k >= num_classes[p]) return;
    const int n0 = noff[pi], cnt = noff[pi + 1] - n0;
    const int q0 = blockIdx.y * kPmPairs, q1 = min(q0 + kPmPairs, cnt * R);
    if (q0 >= q1) return;
    const size_t g = (size_t)p * max_classes + k;
    const int *nl = nodes + n0;
    pmat_pairs<TP, kPmPairs, kPmOneDirect>(eig + g * kEigStride, blen + (size_t)p * nn, rates + (size_t)p * R, R, q0, q1,
                             [nl](int i) { return nl[i]; },
                             [nl, R](int q) { return (size_t)nl[q / R] * R + q % R; }, Pout + ((size_t)pcls_off[p] + k) * nn * R * kPStride);
}

}  // namespace

void launch_eigen(const EigenArgs &a, hipStream_t s)
{
    if (a.count <= 0) return;
    hipLaunchKernelGGL(k_eigen_f64, dim3(a.count * a.max_classes), dim3(64), 0, s, a.base,
                       a.logits, a.num_classes, a.eig, a.max_classes, a.first, a.plist);
}

void launch_eigen_cand(const double *base, const int *ent_particle, const int *ent_ncand,
                       const double *logits, int count, int n_cand, double *eig, const int *cand_cls,
                       const double *eig_cache, int max_classes, const double *pool, hipStream_t s,
                       bool copy_known)
{
    if (count <= 0 || n_cand <= 0) return;
    hipLaunchKernelGGL(k_eigen_cand, dim3(count * n_cand), dim3(64), 0, s, base, ent_particle,
                       ent_ncand, logits, n_cand, eig, cand_cls, eig_cache, max_classes, pool,
                       copy_known ? 1 : 0);
}

template <int PAIRS, int FORM>
void launch_pmat_pairs(const PmatArgs &a, hipStream_t s)
{
    dim3 grid(a.count * a.max_classes, (a.nn * a.R + PAIRS - 1) / PAIRS);
    if (a.fp32)
        hipLaunchKernelGGL((k_pmat<float, PAIRS, FORM>), grid, dim3(kPmLanes), 0, s, a.eig, a.blen, a.rates,
                           a.num_classes, a.pcls_off, (float *)a.P, a.nn, a.R, a.max_classes,
                           a.first, a.plist);
    else
        hipLaunchKernelGGL((k_pmat<double, PAIRS, FORM>), grid, dim3(kPmLanes), 0, s, a.eig, a.blen, a.rates,
                           a.num_classes, a.pcls_off, (double *)a.P, a.nn, a.R, a.max_classes,
                           a.first, a.plist);
}

// Launches with workgroups to spare take runs of 256 pairs in the two-entry form;
// the others runs of 32, staged when they cover many particles (full sweeps) and
// stored straight from the lanes when they cover a few (the whole-particle P of an
// MCMC round, which runs beside the round's pruning: measured faster there)
constexpr int kPmDirectParticles = 8;

void launch_pmat(const PmatArgs &a, hipStream_t s)
{
    if (a.count <= 0) return;
    const int64_t large = (int64_t)a.count * a.max_classes * ((a.nn * a.R + kPmPairsLarge - 1) / kPmPairsLarge);
    if (large >= kPmLargeGrid)
        launch_pmat_pairs<kPmPairsLarge, kPmTwo>(a, s);
    else if (a.count <= kPmDirectParticles)
        launch_pmat_pairs<kPmPairs, kPmOneDirect>(a, s);
    else
        launch_pmat_pairs<kPmPairs, kPmOneStaged>(a, s);
}

}  // namespace smccat

namespace smccat {

void launch_pmat_nodes(const PmatArgs &a, const int *plist, const int *noff, const int *nodes, int n_particles,
                       int max_nodes, hipStream_t s)
{
    if (n_particles <= 0 || max_nodes <= 0) return;
    const dim3 grid((unsigned)n_particles * a.max_classes, (unsigned)((max_nodes * a.R + kPmPairs - 1) / kPmPairs));
    if (a.fp32)
        hipLaunchKernelGGL(k_pmat_nodes<float>, grid, dim3(kPmLanes), 0, s, a.eig, a.blen, a.rates,
                           a.num_classes, a.pcls_off, (float *)a.P, a.nn, a.R, a.max_classes, plist, noff, nodes);
    else
        hipLaunchKernelGGL(k_pmat_nodes<double>, grid, dim3(kPmLanes), 0, s, a.eig, a.blen, a.rates,
                           a.num_classes, a.pcls_off, (double *)a.P, a.nn, a.R, a.max_classes, plist, noff, nodes);
}

}  // namespace smccat
