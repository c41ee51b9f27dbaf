// Split posteriors and the greedy consensus on the GPU.
//
// Reference: Tree::computeSplits / getSplitSet / getSplitBranchMap (Tree.cpp:485-535):
// every non-root node's set of tips below it, flipped when it does not hold taxon 0
// (so a split and its complement are one split); computeSplitPosteriors
// (Miscellaneous.hpp:115-150): the sum of the particles' weights over the particles
// whose split set holds the split; the greedy consensus and its branch lengths
// (main.cpp:338-383): splits in decreasing posterior, each kept when compatible with
// every split kept before (four-point condition), lengths the weighted mean over the
// particles holding the split.
//
// Layout: a split is W = ceil(T / 64) 64-bit words, bit t = taxon t.  Particle p
// contributes M = 2T - 3 entries (every non-root node but the root's second child,
// whose flipped split is the first child's; its length is added to that entry, as
// getSplitBranchMap adds lengths per split).  Entries are grouped by a 64-bit hash
// of their words with a stable radix sort, so each group's weights are summed in
// particle order: the sums do not depend on timing.  Two different splits with one
// hash are detected and reported.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>

#include "engine_state.hpp"

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    // SplitMix64 finaliser (Steele, Lea & Flood, "Fast splittable pseudorandom number
    // generators", OOPSLA 2014)
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return x;
}

// Tip sets of every node of particle blockIdx.x (nodebits[p][node][W]): lane w
// walks the post-order for word w.
__global__ __launch_bounds__(64) void k_node_sets(const int *__restrict__ children, const int *__restrict__ prog,
                                                  int T, int W, uint64_t *__restrict__ nodebits)
{
    const int p = blockIdx.x, w = threadIdx.x;
    if (w >= W) return;
    const int nn = 2 * T - 1;
    uint64_t *nb = nodebits + (size_t)p * nn * W;
    for (int t = 0; t < T; ++t) nb[(size_t)t * W + w] = (t >> 6) == w ? (1ull << (t & 63)) : 0ull;
    const int *ch = children + (size_t)p * (T - 1) * 2;
    const int *post = prog + (size_t)p * site_prog_stride(T);  // internal nodes, children first
    for (int k = 0; k < T - 1; ++k) {
        const int n = post[k];
        const int a = ch[2 * (n - T)], b = ch[2 * (n - T) + 1];
        nb[(size_t)n * W + w] = nb[(size_t)a * W + w] | nb[(size_t)b * W + w];
    }
}

// The M entries of each particle: flipped words, hash, length, particle.
__global__ __launch_bounds__(256) void k_split_entries(const uint64_t *__restrict__ nodebits,
                                                       const int *__restrict__ children, const int *__restrict__ root_of,
                                                       const double *__restrict__ blen, int n, int T, int W,
                                                       uint64_t *__restrict__ keys, int *__restrict__ vals,
                                                       uint64_t *__restrict__ ewords, double *__restrict__ ebl,
                                                       int *__restrict__ epart)
{
    const int nn = 2 * T - 1, M = 2 * T - 3;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * M) return;
    const int p = (int)(e / M), m = (int)(e % M);
    const int root = root_of[p];
    const int *ch = children + (size_t)p * (T - 1) * 2;
    const int first = ch[2 * (root - T)], second = ch[2 * (root - T) + 1];
    // entry m: the m-th node of 0..nn-1 other than the root and its second child
    int node = m;
    const int lo = root < second ? root : second, hi = root < second ? second : root;
    if (node >= lo) node++;
    if (node >= hi) node++;
    const uint64_t *nb = nodebits + ((size_t)p * nn + node) * W;
    const bool flip = (nb[0] & 1ull) == 0;
    const int tail = T - 64 * (W - 1);  // bits used in the last word
    const uint64_t last_mask = tail == 64 ? ~0ull : ((1ull << tail) - 1ull);
    uint64_t h = 0x9E3779B97F4A7C15ull ^ (uint64_t)W;
    for (int w = 0; w < W; ++w) {
        uint64_t x = flip ? ~nb[w] : nb[w];
        if (w == W - 1) x &= last_mask;
        ewords[(size_t)e * W + w] = x;
        h = mix64(h ^ x) + (uint64_t)w;
    }
    keys[e] = h;
    vals[e] = (int)e;
    double len = blen[(size_t)p * nn + node];
    if (node == first) len += blen[(size_t)p * nn + second];
    ebl[e] = len;
    epart[e] = p;
}

__device__ __forceinline__ bool same_words(const uint64_t *a, const uint64_t *b, int W)
{
    for (int w = 0; w < W; ++w)
        if (a[w] != b[w]) return false;
    return true;
}

// Group heads of the hash-sorted entries; a hash shared by different splits sets *collision.
__global__ __launch_bounds__(256) void k_split_heads(const uint64_t *__restrict__ skeys, const int *__restrict__ svals,
                                                     const uint64_t *__restrict__ ewords, int64_t E, int W,
                                                     int *__restrict__ head, int *__restrict__ collision)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E) return;
    const bool h = i == 0 || skeys[i] != skeys[i - 1];
    head[i] = h ? 1 : 0;
    if (!h && !same_words(ewords + (size_t)svals[i] * W, ewords + (size_t)svals[i - 1] * W, W))
        atomicOr(collision, 1);
}

// One thread per group head: its weight and weighted-length sums in entry order.
__global__ __launch_bounds__(256) void k_split_reduce(const uint64_t *__restrict__ skeys, const int *__restrict__ svals,
                                                      const int *__restrict__ head, const int *__restrict__ uid,
                                                      const uint64_t *__restrict__ ewords, const double *__restrict__ ebl,
                                                      const int *__restrict__ epart, const double *__restrict__ weight,
                                                      int64_t E, int W, uint64_t *__restrict__ uwords,
                                                      double *__restrict__ upost, double *__restrict__ ulen)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E || !head[i]) return;
    double sw = 0.0, swl = 0.0;
    for (int64_t j = i; j < E && (j == i || !head[j]); ++j) {
        const int en = svals[j];
        const double w = weight[epart[en]];
        sw += w;
        swl += w * ebl[en];
    }
    const int u = uid[i];
    for (int w = 0; w < W; ++w) uwords[(size_t)u * W + w] = ewords[(size_t)svals[i] * W + w];
    upost[u] = sw;
    ulen[u] = sw > 0.0 ? swl / sw : 0.0;
}

// Greedy consensus (main.cpp:344-365) over the splits in decreasing posterior:
// one workgroup; each candidate is checked against the kept splits in parallel.
constexpr int kConsThreads = 256;
__global__ __launch_bounds__(kConsThreads) void k_consensus(const uint64_t *__restrict__ uwords,
                                                            const int *__restrict__ order, int U, int T, int W,
                                                            int max_keep, int *__restrict__ kept, int *__restrict__ nkept)
{
    __shared__ int s_n, s_bad;
    const int tid = threadIdx.x;
    const int tail = T - 64 * (W - 1);
    const uint64_t last_mask = tail == 64 ? ~0ull : ((1ull << tail) - 1ull);
    if (tid == 0) s_n = 0;
    __syncthreads();
    for (int i = 0; i < U; ++i) {
        if (s_n >= max_keep) break;  // resolved: every further split is incompatible
        if (tid == 0) s_bad = 0;
        __syncthreads();
        const uint64_t *a = uwords + (size_t)order[i] * W;
        for (int k = tid; k < s_n; k += kConsThreads) {
            const uint64_t *b = uwords + (size_t)kept[k] * W;
            bool ab = false, anb = false, nab = false, nanb = false;  // non-empty intersections
            for (int w = 0; w < W; ++w) {
                const uint64_t m = w == W - 1 ? last_mask : ~0ull;
                ab |= (a[w] & b[w] & m) != 0;
                anb |= (a[w] & ~b[w] & m) != 0;
                nab |= (~a[w] & b[w] & m) != 0;
                nanb |= (~a[w] & ~b[w] & m) != 0;
            }
            if (ab && anb && nab && nanb) atomicOr(&s_bad, 1);
        }
        __syncthreads();
        if (tid == 0 && !s_bad) kept[s_n++] = order[i];
        __syncthreads();
    }
    if (tid == 0) *nkept = s_n;
}

// Bump allocator over the engine's split workspace (a null base only sums the sizes).
struct Arena {
    uint8_t *base;
    size_t off = 0;
    template <typename X>
    X *take(size_t count)
    {
        X *p = base ? reinterpret_cast<X *>(base + off) : nullptr;
        off += (count * sizeof(X) + 255) & ~(size_t)255;
        return p;
    }
};

struct SplitBufs {
    uint64_t *nodebits, *keys, *skeys, *ewords, *uwords;
    int *vals, *svals, *epart, *head, *uid, *flags;
    double *ebl, *weight, *upost, *ulen, *spost;
    int *order, *order_in, *kept;
    void *cub;
    size_t cub_bytes;
};

size_t layout(SplitBufs &b, uint8_t *base, int n, int T, int W, size_t cub_bytes)
{
    const int nn = 2 * T - 1;
    const size_t E = (size_t)n * (2 * T - 3);
    Arena a{base};
    b.nodebits = a.take<uint64_t>((size_t)n * nn * W);
    b.keys = a.take<uint64_t>(E);
    b.skeys = a.take<uint64_t>(E);
    b.ewords = a.take<uint64_t>(E * W);
    b.uwords = a.take<uint64_t>(E * W);
    b.vals = a.take<int>(E);
    b.svals = a.take<int>(E);
    b.epart = a.take<int>(E);
    b.head = a.take<int>(E);
    b.uid = a.take<int>(E);
    b.flags = a.take<int>(4);
    b.ebl = a.take<double>(E);
    b.weight = a.take<double>(n);
    b.upost = a.take<double>(E);
    b.ulen = a.take<double>(E);
    b.spost = a.take<double>(E);
    b.order = a.take<int>(E);
    b.order_in = a.take<int>(E);
    b.kept = a.take<int>(2 * T);
    b.cub = a.take<uint8_t>(cub_bytes);
    b.cub_bytes = cub_bytes;
    return a.off;
}

__global__ void k_iota(int *__restrict__ v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (int)i;
}

// The unique splits of slots [0, n) with posteriors and mean lengths, on the device;
// returns their number in *U (after a stream synchronisation).
int split_posteriors_device(smc_engine *e, int n, const double *weights, SplitBufs &b, int64_t *U_out)
{
    const int T = e->T, W = (T + 63) / 64;
    const int64_t E = (int64_t)n * (2 * T - 3);
    hipStream_t s = e->stream;
    // temporary storage of the library sorts and scan (the largest of them)
    size_t t1 = 0, t2 = 0, t3 = 0;
    HIP_TRY(e, hipcub::DeviceRadixSort::SortPairs(nullptr, t1, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                                  (int *)nullptr, (int *)nullptr, (int)E, 0, 64, s));
    HIP_TRY(e, hipcub::DeviceScan::ExclusiveSum(nullptr, t2, (int *)nullptr, (int *)nullptr, (int)E, s));
    HIP_TRY(e, hipcub::DeviceRadixSort::SortPairsDescending(nullptr, t3, (double *)nullptr, (double *)nullptr,
                                                            (int *)nullptr, (int *)nullptr, (int)E, 0, 64, s));
    const size_t cub_bytes = std::max(t1, std::max(t2, t3)) + 256;
    const size_t need = layout(b, nullptr, n, T, W, cub_bytes);
    if (need > e->split_ws_bytes) {
        HIP_TRY(e, hipStreamSynchronize(s));
        if (e->d_split_ws) (void)hipFree(e->d_split_ws);
        e->d_split_ws = nullptr;
        e->split_ws_bytes = 0;
        HIP_TRY(e, hipMalloc((void **)&e->d_split_ws, need));
        e->split_ws_bytes = need;
    }
    layout(b, e->d_split_ws, n, T, W, cub_bytes);
    int rc = h2d(e, b.weight, weights, (size_t)n * sizeof(double));
    if (rc != SMC_OK) return rc;
    HIP_TRY(e, hipMemsetAsync(b.flags, 0, 4 * sizeof(int), s));
    hipLaunchKernelGGL(k_node_sets, dim3(n), dim3(64), 0, s, e->d_children, e->d_prog, T, W, b.nodebits);
    const unsigned eg = (unsigned)((E + 255) / 256);
    hipLaunchKernelGGL(k_split_entries, dim3(eg), dim3(256), 0, s, b.nodebits, e->d_children, e->d_root, e->d_blen, n,
                       T, W, b.keys, b.vals, b.ewords, b.ebl, b.epart);
    size_t tb = b.cub_bytes;
    HIP_TRY(e, hipcub::DeviceRadixSort::SortPairs(b.cub, tb, b.keys, b.skeys, b.vals, b.svals, (int)E, 0, 64, s));
    hipLaunchKernelGGL(k_split_heads, dim3(eg), dim3(256), 0, s, b.skeys, b.svals, b.ewords, E, W, b.head, b.flags);
    tb = b.cub_bytes;
    HIP_TRY(e, hipcub::DeviceScan::ExclusiveSum(b.cub, tb, b.head, b.uid, (int)E, s));
    hipLaunchKernelGGL(k_split_reduce, dim3(eg), dim3(256), 0, s, b.skeys, b.svals, b.head, b.uid, b.ewords, b.ebl,
                       b.epart, b.weight, E, W, b.uwords, b.upost, b.ulen);
    HIP_TRY(e, hipGetLastError());
    int tailv[3];
    HIP_TRY(e, hipMemcpyAsync(&tailv[0], b.uid + (E - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(e, hipMemcpyAsync(&tailv[1], b.head + (E - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(e, hipMemcpyAsync(&tailv[2], b.flags, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(e, hipStreamSynchronize(s));
    if (tailv[2]) return smc_fail(e, SMC_ERR_STATE, "split posteriors: two different splits share a 64-bit hash");
    *U_out = (int64_t)tailv[0] + tailv[1];
    return SMC_OK;
}

int check_split_args(smc_engine *e, int n, const double *weights)
{
    if (n < 1 || n > e->Pmax || !weights) return smc_fail(e, SMC_ERR_ARG, "invalid split-posterior arguments");
    if ((int64_t)n * (2 * e->T - 3) > (int64_t)INT32_MAX)
        return smc_fail(e, SMC_ERR_ARG, "split posteriors: too many entries");
    for (int p = 0; p < n; ++p)
        if (!e->h_valid[p]) return smc_fail(e, SMC_ERR_STATE, "split posteriors: slot not uploaded");
    return SMC_OK;
}

}  // namespace

extern "C" int smc_split_posteriors(smc_engine *e, int n, const double *weights, int64_t max_splits,
                                    int64_t *n_splits, uint64_t *split_words, double *posterior,
                                    double *mean_length)
{
    if (!e) return SMC_ERR_ARG;
    if (!n_splits) return smc_fail(e, SMC_ERR_ARG, "n_splits is null");
    int rc = check_split_args(e, n, weights);
    if (rc != SMC_OK) return rc;
    SplitBufs b{};
    int64_t U = 0;
    rc = split_posteriors_device(e, n, weights, b, &U);
    if (rc != SMC_OK) return rc;
    *n_splits = U;
    if (U > max_splits) return SMC_OK;  // the caller sizes its arrays and calls again
    const int W = (e->T + 63) / 64;
    hipStream_t s = e->stream;
    if (split_words)
        HIP_TRY(e, hipMemcpyAsync(split_words, b.uwords, (size_t)U * W * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    if (posterior) HIP_TRY(e, hipMemcpyAsync(posterior, b.upost, (size_t)U * sizeof(double), hipMemcpyDeviceToHost, s));
    if (mean_length)
        HIP_TRY(e, hipMemcpyAsync(mean_length, b.ulen, (size_t)U * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(e, hipStreamSynchronize(s));
    return SMC_OK;
}

extern "C" int smc_consensus(smc_engine *e, int n, const double *weights, int32_t *n_selected,
                             uint64_t *split_words, double *posterior, double *mean_length)
{
    if (!e) return SMC_ERR_ARG;
    if (!n_selected) return smc_fail(e, SMC_ERR_ARG, "n_selected is null");
    int rc = check_split_args(e, n, weights);
    if (rc != SMC_OK) return rc;
    SplitBufs b{};
    int64_t U = 0;
    rc = split_posteriors_device(e, n, weights, b, &U);
    if (rc != SMC_OK) return rc;
    const int T = e->T, W = (T + 63) / 64;
    hipStream_t s = e->stream;
    // decreasing posterior; the stable sort keeps equal posteriors in hash order
    hipLaunchKernelGGL(k_iota, dim3((unsigned)((U + 255) / 256)), dim3(256), 0, s, b.order_in, U);
    size_t tb = b.cub_bytes;
    HIP_TRY(e, hipcub::DeviceRadixSort::SortPairsDescending(b.cub, tb, b.upost, b.spost, b.order_in, b.order, (int)U,
                                                            0, 64, s));
    const int max_keep = 2 * T - 3;  // a resolved binary tree's splits, trivial ones included
    hipLaunchKernelGGL(k_consensus, dim3(1), dim3(kConsThreads), 0, s, b.uwords, b.order, (int)U, T, W, max_keep,
                       b.kept, b.flags + 1);
    HIP_TRY(e, hipGetLastError());
    int nk = 0;
    std::vector<int> kept(max_keep);
    HIP_TRY(e, hipMemcpyAsync(&nk, b.flags + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(e, hipMemcpyAsync(kept.data(), b.kept, (size_t)max_keep * sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(e, hipStreamSynchronize(s));
    *n_selected = nk;
    for (int k = 0; k < nk; ++k) {
        const int u = kept[k];
        if (split_words)
            HIP_TRY(e, hipMemcpyAsync(split_words + (size_t)k * W, b.uwords + (size_t)u * W, W * sizeof(uint64_t),
                                      hipMemcpyDeviceToHost, s));
        if (posterior) HIP_TRY(e, hipMemcpyAsync(posterior + k, b.upost + u, sizeof(double), hipMemcpyDeviceToHost, s));
        if (mean_length) HIP_TRY(e, hipMemcpyAsync(mean_length + k, b.ulen + u, sizeof(double), hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(e, hipStreamSynchronize(s));
    return SMC_OK;
}
