// Training and inference kernels of the CAT-allocation proposal network
// (layout and definition in mlp_layout.hpp).  FP32 throughout.
//
// A batch of B examples is processed in rows of kRows = 16 examples:
//   k_mlp_trunk      one workgroup per row, thread j owns hidden unit j; W2 is
//                    staged transposed in LDS so the 128 x 128 layer reads it
//                    conflict-free; keeps z1, h1, z2 (training) and h2.
//   k_mlp_head       one wavefront per example (4 per wave, 16 per workgroup):
//                    lane l owns units l and l + 64 and scores classes l, l + 64,
//                    ... in parallel; log-softmax over K + 1 slots; in training
//                    also the loss, d loss / d h2 and the head's gradient.
//   k_mlp_trunk_bwd  the trunk's backward pass for one row.
// Gradients are written per row ("partials", [rows][kNParam]) and summed by
// k_mlp_reduce in a fixed order, so a gradient does not depend on timing.
// k_mlp_adam applies Adam (Kingma & Ba 2015, Algorithm 1).
#include "internal.hpp"
#include "mlp_layout.hpp"

namespace smccat {

namespace {

using namespace mlp;

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * erfcf(-z * 0.70710678118654752f); }
__device__ __forceinline__ float gelu_grad(float z)
{
    return 0.5f * erfcf(-z * 0.70710678118654752f) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ int example_of(const int *idx, int r) { return idx ? idx[r] : r; }

__global__ __launch_bounds__(128) void k_mlp_trunk(int B, const int *__restrict__ idx,
                                                   const float *__restrict__ hdr,
                                                   const float *__restrict__ theta, float *__restrict__ Z1,
                                                   float *__restrict__ H1, float *__restrict__ Z2,
                                                   float *__restrict__ H2)
{
    __shared__ float xs[kRows][kIn];
    __shared__ float h1s[kRows][kH];
    __shared__ float w2t[kH][kH + 1];  // w2t[i][j] = W2[j][i]
    const int j = threadIdx.x;
    const int r0 = blockIdx.x * kRows;
    const int nr = min(kRows, B - r0);
    for (int f = j; f < kRows * kIn; f += kH) {
        const int r = f / kIn, i = f % kIn;
        xs[r][i] = r < nr ? hdr[(size_t)example_of(idx, r0 + r) * kHdr + i] : 0.f;
    }
    for (int f = j; f < kH * kH; f += kH) w2t[f % kH][f / kH] = theta[kW2 + f];
    __syncthreads();
    float acc[kRows];
    const float b1 = theta[kB1 + j];
#pragma unroll
    for (int r = 0; r < kRows; r++) acc[r] = b1;
    for (int i = 0; i < kIn; i++) {
        const float w = theta[kW1 + j * kIn + i];
#pragma unroll
        for (int r = 0; r < kRows; r++) acc[r] = fmaf(w, xs[r][i], acc[r]);
    }
#pragma unroll
    for (int r = 0; r < kRows; r++) {
        const float h = gelu(acc[r]);
        h1s[r][j] = h;
        if (r < nr && Z1) {
            Z1[(size_t)(r0 + r) * kH + j] = acc[r];
            H1[(size_t)(r0 + r) * kH + j] = h;
        }
    }
    __syncthreads();
    const float b2 = theta[kB2 + j];
#pragma unroll
    for (int r = 0; r < kRows; r++) acc[r] = b2;
    for (int i = 0; i < kH; i++) {
        const float w = w2t[i][j];
#pragma unroll
        for (int r = 0; r < kRows; r++) acc[r] = fmaf(w, h1s[r][i], acc[r]);
    }
#pragma unroll
    for (int r = 0; r < kRows; r++)
        if (r < nr) {
            if (Z2) Z2[(size_t)(r0 + r) * kH + j] = acc[r];
            H2[(size_t)(r0 + r) * kH + j] = gelu(acc[r]);
        }
}

// the value held for class c by lane (c & 63) in slot (c >> 6), broadcast to the wave
__device__ __forceinline__ float slot_get(const float (&v)[kMaxCand / 64], int c)
{
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kMaxCand / 64; k++)
        if (k == (c >> 6)) s = v[k];
    return __shfl(s, c & 63, 64);
}

// The head works in the class-parallel form l_c = w0 + w . log pi_c with
// w[a] = sum_j h2[j] U[j][a] and w0 = h2 . u (20 + 1 wave reductions per
// example); lane l then scores classes l, l + 64, ... independently.  The
// backward pass uses the same form: with rr = sum_c dl_c log pi_c and
// sd = sum_c dl_c, dh2 = U rr + u sd + v dl_new and dU = h2 rr^T.
template <bool TRAIN>
__global__ __launch_bounds__(256) void k_mlp_head(int B, const int *__restrict__ idx,
                                                  const float *__restrict__ hdr, const int *__restrict__ Kc,
                                                  const int *__restrict__ label,
                                                  const long long *__restrict__ coff,
                                                  const float *__restrict__ cls,
                                                  const float *__restrict__ theta,
                                                  const float *__restrict__ H2, float *__restrict__ logq,
                                                  int out_stride, float *__restrict__ dH2,
                                                  float *__restrict__ loss, float *__restrict__ part,
                                                  float inv_B)
{
    constexpr int NH = kNParam - kHeadFirst;
    constexpr int NS = kMaxCand / 64;
    __shared__ float red[TRAIN ? 4 : 1][TRAIN ? NH : 1];
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int j0 = lane, j1 = lane + 64;
    float U0[20], U1[20];
#pragma unroll
    for (int a = 0; a < 20; a++) {
        U0[a] = theta[kU + j0 * 20 + a];
        U1[a] = theta[kU + j1 * 20 + a];
    }
    const float ub0 = theta[kUB + j0], ub1 = theta[kUB + j1];
    const float v0 = theta[kV + j0], v1 = theta[kV + j1];
    const float bn = theta[kBN];
    float dU0[20], dU1[20], du0 = 0.f, du1 = 0.f, dv0 = 0.f, dv1 = 0.f, dbn = 0.f;
#pragma unroll
    for (int a = 0; a < 20; a++) dU0[a] = dU1[a] = 0.f;

    for (int q = 0; q < 4; q++) {
        const int r = blockIdx.x * kRows + wave * 4 + q;
        if (r >= B) break;  // wave-uniform; no barrier inside this loop
        const int ex = example_of(idx, r);
        const float tau = hdr[(size_t)ex * kHdr + 40], la = hdr[(size_t)ex * kHdr + 41];
        const int K = Kc[ex];
        const float *ce = cls + (size_t)coff[ex] * kCls;
        const float h0 = H2[(size_t)r * kH + j0], h1 = H2[(size_t)r * kH + j1];
        float w[20];
#pragma unroll
        for (int a = 0; a < 20; a++) w[a] = wave_sum(fmaf(h0, U0[a], h1 * U1[a]));
        const float w0 = wave_sum(fmaf(h0, ub0, h1 * ub1));
        const float ln = wave_sum(fmaf(h0, v0, h1 * v1)) + bn;
        float lg[NS];
#pragma unroll
        for (int k = 0; k < NS; k++) {
            const int c = 64 * k + lane;
            lg[k] = -INFINITY;
            if (c < K) {
                const float *e = ce + (size_t)c * kCls;
                float l = w0;
#pragma unroll
                for (int a = 0; a < 20; a++) l = fmaf(w[a], e[a], l);
                lg[k] = fmaf(tau, l, e[20]);
            } else if (c == K) {
                lg[k] = fmaf(tau, ln, la);
            }
        }
        float m = -INFINITY;
#pragma unroll
        for (int k = 0; k < NS; k++) m = fmaxf(m, lg[k]);
        m = wave_max(m);
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < NS; k++) s += (lg[k] == -INFINITY) ? 0.f : expf(lg[k] - m);
        const float lse = m + logf(wave_sum(s));
        if constexpr (!TRAIN) {
#pragma unroll
            for (int k = 0; k < NS; k++) {
                const int c = 64 * k + lane;
                if (c <= K) logq[(size_t)r * out_stride + c] = lg[k] - lse;
            }
        } else {
            const int lab = label[ex];
            const float llab = slot_get(lg, lab);
            if (lane == 0) loss[r] = lse - llab;
            float rr[20];
#pragma unroll
            for (int a = 0; a < 20; a++) rr[a] = 0.f;
            float sd = 0.f, dln = 0.f;
#pragma unroll
            for (int k = 0; k < NS; k++) {
                const int c = 64 * k + lane;
                if (c > K) continue;
                const float dl = tau * (expf(lg[k] - lse) - (c == lab ? 1.f : 0.f)) * inv_B;
                if (c == K) {
                    dln = dl;
                } else {
                    const float *e = ce + (size_t)c * kCls;
#pragma unroll
                    for (int a = 0; a < 20; a++) rr[a] = fmaf(dl, e[a], rr[a]);
                    sd += dl;
                }
            }
#pragma unroll
            for (int a = 0; a < 20; a++) rr[a] = wave_sum(rr[a]);
            sd = wave_sum(sd);
            dln = wave_sum(dln);  // held by one lane
            float dh0 = fmaf(ub0, sd, v0 * dln), dh1 = fmaf(ub1, sd, v1 * dln);
#pragma unroll
            for (int a = 0; a < 20; a++) {
                dh0 = fmaf(U0[a], rr[a], dh0);
                dh1 = fmaf(U1[a], rr[a], dh1);
            }
            dH2[(size_t)r * kH + j0] = dh0;
            dH2[(size_t)r * kH + j1] = dh1;
#pragma unroll
            for (int a = 0; a < 20; a++) {
                dU0[a] = fmaf(h0, rr[a], dU0[a]);
                dU1[a] = fmaf(h1, rr[a], dU1[a]);
            }
            du0 = fmaf(h0, sd, du0);
            du1 = fmaf(h1, sd, du1);
            dv0 = fmaf(dln, h0, dv0);
            dv1 = fmaf(dln, h1, dv1);
            dbn += dln;
        }
    }
    if constexpr (TRAIN) {
        float *rw = red[wave];
#pragma unroll
        for (int a = 0; a < 20; a++) {
            rw[j0 * 20 + a] = dU0[a];
            rw[j1 * 20 + a] = dU1[a];
        }
        rw[(kUB - kU) + j0] = du0;
        rw[(kUB - kU) + j1] = du1;
        rw[(kV - kU) + j0] = dv0;
        rw[(kV - kU) + j1] = dv1;
        if (lane == 0) rw[kBN - kU] = dbn;
        __syncthreads();
        for (int f = threadIdx.x; f < NH; f += 256)
            part[(size_t)blockIdx.x * kNParam + kHeadFirst + f] =
                (red[0][f] + red[1][f]) + (red[2][f] + red[3][f]);
    }
}

__global__ __launch_bounds__(128) void k_mlp_trunk_bwd(int B, const int *__restrict__ idx,
                                                       const float *__restrict__ hdr,
                                                       const float *__restrict__ theta,
                                                       const float *__restrict__ Z1,
                                                       const float *__restrict__ H1,
                                                       const float *__restrict__ Z2,
                                                       const float *__restrict__ dH2,
                                                       float *__restrict__ part)
{
    __shared__ float xs[kRows][kIn];
    __shared__ float h1s[kRows][kH];
    __shared__ float dz2s[kRows][kH];
    __shared__ float dz1s[kRows][kH];
    const int t = threadIdx.x;
    const int r0 = blockIdx.x * kRows;
    const int nr = min(kRows, B - r0);
    float *pr = part + (size_t)blockIdx.x * kNParam;
    for (int f = t; f < kRows * kIn; f += kH) {
        const int r = f / kIn, i = f % kIn;
        xs[r][i] = r < nr ? hdr[(size_t)example_of(idx, r0 + r) * kHdr + i] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < kRows; r++) {
        const bool ok = r < nr;
        const size_t o = (size_t)(r0 + r) * kH + t;
        h1s[r][t] = ok ? H1[o] : 0.f;
        dz2s[r][t] = ok ? dH2[o] * gelu_grad(Z2[o]) : 0.f;
    }
    __syncthreads();
    // dW2[k][t] = sum_r dz2[r][k] h1[r][t]; coalesced over t
    for (int k = 0; k < kH; k++) {
        float acc = 0.f;
#pragma unroll
        for (int r = 0; r < kRows; r++) acc = fmaf(dz2s[r][k], h1s[r][t], acc);
        pr[kW2 + k * kH + t] = acc;
    }
    {
        float acc = 0.f;
#pragma unroll
        for (int r = 0; r < kRows; r++) acc += dz2s[r][t];
        pr[kB2 + t] = acc;
    }
    // dh1[r][t] = sum_j W2[j][t] dz2[r][j]
    float acc[kRows];
#pragma unroll
    for (int r = 0; r < kRows; r++) acc[r] = 0.f;
    for (int jj = 0; jj < kH; jj++) {
        const float w = theta[kW2 + jj * kH + t];
#pragma unroll
        for (int r = 0; r < kRows; r++) acc[r] = fmaf(w, dz2s[r][jj], acc[r]);
    }
#pragma unroll
    for (int r = 0; r < kRows; r++)
        dz1s[r][t] = r < nr ? acc[r] * gelu_grad(Z1[(size_t)(r0 + r) * kH + t]) : 0.f;
    __syncthreads();
    for (int k = 0; k < kIn; k++) {
        const int f = t + kH * k;  // f = j * kIn + i
        const int j = f / kIn, i = f % kIn;
        float a = 0.f;
#pragma unroll
        for (int r = 0; r < kRows; r++) a = fmaf(dz1s[r][j], xs[r][i], a);
        pr[kW1 + f] = a;
    }
    {
        float a = 0.f;
#pragma unroll
        for (int r = 0; r < kRows; r++) a += dz1s[r][t];
        pr[kB1 + t] = a;
    }
}

__global__ __launch_bounds__(256) void k_mlp_reduce(int rows, const float *__restrict__ part,
                                                    float *__restrict__ grad, const float *__restrict__ loss,
                                                    int B, double *__restrict__ loss_acc)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < kNParam) {
        float s = 0.f;
        for (int r = 0; r < rows; r++) s += part[(size_t)r * kNParam + p];
        grad[p] = s;
    }
    if (blockIdx.x == 0 && loss_acc) {
        __shared__ double ws[4];
        double s = 0.0;
        for (int b = threadIdx.x; b < B; b += 256) s += (double)loss[b];
        for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
        if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) *loss_acc += (ws[0] + ws[1]) + (ws[2] + ws[3]);
    }
}

__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ theta, float *__restrict__ m,
                                                  float *__restrict__ v, const float *__restrict__ g,
                                                  float lr, float b1, float b2, float eps, float c1,
                                                  float c2)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kNParam) return;
    const float gp = g[p];
    const float mp = b1 * m[p] + (1.f - b1) * gp;
    const float vp = b2 * v[p] + (1.f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    theta[p] -= lr * (mp * c1) / (sqrtf(vp * c2) + eps);
}

// The learned move's proposal draw, one lane per particle: class c of the nslot
// slots (the K classes, plus the new-class slot while a class slot is free) with
// probability exp(logq[c]) / sum, by the walk of cumulative weights against
// u * sum, in double; then the site-class entry of the MH step -- the current and
// the drawn class's codes, or no candidates when the draw is the current class.
__global__ __launch_bounds__(256) void k_mlp_draw(MoveDrawArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    const float *lq = a.logq + (size_t)i * a.stride;
    const int K = a.K[i], nslot = K < a.kmax ? K + 1 : K;
    double tot = 0.0;
    for (int c = 0; c < nslot; ++c) tot += exp((double)lq[c]);
    const double x = a.u[i] * tot;
    int pick = nslot - 1;
    double acc = 0.0;
    for (int c = 0; c < nslot; ++c) {
        acc += exp((double)lq[c]);
        if (acc > x) {
            pick = c;
            break;
        }
    }
    const int *ax = a.aux + 4 * (size_t)i;
    a.pick[i] = pick;
    a.ent[i] = ax[0];
    a.ent[a.count + i] = ax[1];
    a.ent[2 * a.count + i] = pick != ax[2] ? 2 : 0;
    a.cand_cls[2 * i] = ax[3];
    a.cand_cls[2 * i + 1] = a.codes[(size_t)i * (a.kmax + 1) + pick];
}

}  // namespace

void launch_mlp_draw(const MoveDrawArgs &a, hipStream_t s)
{
    if (a.count <= 0) return;
    hipLaunchKernelGGL(k_mlp_draw, dim3((a.count + 255) / 256), dim3(256), 0, s, a);
}

void launch_mlp_trunk(const MlpArgs &a, bool train, hipStream_t s)
{
    if (a.B <= 0) return;
    const unsigned rows = (unsigned)((a.B + kRows - 1) / kRows);
    hipLaunchKernelGGL(k_mlp_trunk, dim3(rows), dim3(128), 0, s, a.B, a.idx, a.hdr, a.theta,
                       train ? a.Z1 : nullptr, train ? a.H1 : nullptr, train ? a.Z2 : nullptr, a.H2);
}

void launch_mlp_head(const MlpArgs &a, bool train, hipStream_t s)
{
    if (a.B <= 0) return;
    const unsigned rows = (unsigned)((a.B + kRows - 1) / kRows);
    const long long *coff = reinterpret_cast<const long long *>(a.coff);
    if (train)
        hipLaunchKernelGGL((k_mlp_head<true>), dim3(rows), dim3(256), 0, s, a.B, a.idx, a.hdr, a.K,
                           a.label, coff, a.cls, a.theta, a.H2, nullptr, 0, a.dH2, a.loss, a.part,
                           a.inv_B);
    else
        hipLaunchKernelGGL((k_mlp_head<false>), dim3(rows), dim3(256), 0, s, a.B, a.idx, a.hdr, a.K,
                           nullptr, coff, a.cls, a.theta, a.H2, a.logq, a.out_stride, nullptr, nullptr,
                           nullptr, 0.f);
}

void launch_mlp_backward(const MlpArgs &a, hipStream_t s)
{
    if (a.B <= 0) return;
    const unsigned rows = (unsigned)((a.B + kRows - 1) / kRows);
    hipLaunchKernelGGL(k_mlp_trunk_bwd, dim3(rows), dim3(128), 0, s, a.B, a.idx, a.hdr, a.theta, a.Z1,
                       a.H1, a.Z2, a.dH2, a.part);
}

void launch_mlp_reduce(int rows, const float *part, float *grad, const float *loss, int B,
                       double *loss_acc, hipStream_t s)
{
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kNParam + 255) / 256), dim3(256), 0, s, rows, part, grad, loss,
                       B, loss_acc);
}

void launch_mlp_adam(float *theta, float *m, float *v, const float *g, float lr, float b1, float b2,
                     float eps, float c1, float c2, hipStream_t s)
{
    hipLaunchKernelGGL(k_mlp_adam, dim3((kNParam + 255) / 256), dim3(256), 0, s, theta, m, v, g, lr, b1, b2,
                       eps, c1, c2);
}

}  // namespace smccat
