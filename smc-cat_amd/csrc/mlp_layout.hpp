// Layout of the CAT-allocation proposal network and of its examples (shared by
// the HIP kernels and the host code; not installed).
//
// The network scores Algorithm 8's candidates for one site (Neal 2000; the
// move of Particle::gibbsPartitionMove, Particle.cpp:785-995): the site is
// removed from its class, C are the remaining classes with sizes n_c, and one
// auxiliary "new" class stands for all fresh profiles.
//   x   = [site amino-acid frequencies (20), size-weighted mean profile (20)]
//   h1  = gelu(W1 x + b1), h2 = gelu(W2 h1 + b2)              (40 -> 128 -> 128)
//   l_c = h2 . (U log pi_c + u)   for c in C,   l_new = h2 . v + b_new
//   logits = tau * l_c + log n_c  and  tau * l_new + log alpha;  q = softmax
// l plays the part of the site log-likelihood, so the proposal is tempered
// like its target.  Parameters are one float32 vector theta.
#pragma once

namespace smccat {
namespace mlp {

constexpr int kH = 128;       // hidden units
constexpr int kIn = 40;       // inputs
constexpr int kW1 = 0;        // [kH][kIn]
constexpr int kB1 = 5120;     // [kH]
constexpr int kW2 = 5248;     // [kH][kH]
constexpr int kB2 = 21632;    // [kH]
constexpr int kU = 21760;     // [kH][20]
constexpr int kUB = 24320;    // [kH]
constexpr int kV = 24448;     // [kH]
constexpr int kBN = 24576;    // scalar
constexpr int kNParam = 24577;
constexpr int kHeadFirst = kU;  // parameters written by the head kernel: [kU, kNParam)
constexpr int kHdr = 44;        // example header: x[40], tau, log alpha, 2 pad
constexpr int kCls = 24;        // class entry: log pi[20], log n, 3 pad
constexpr int kMaxCand = 256;   // classes per example the kernels accept
constexpr int kRows = 16;       // examples per workgroup (one row of partial gradients)

}  // namespace mlp
}  // namespace smccat
