"""mlp_oracle.py -- TEST INFRASTRUCTURE ONLY.

CPU float64 restatement of the learned CAT-allocation proposal network, used as
the checker for the HIP training and inference kernels.  Nothing in the product
library imports this file.

The reference has no learned proposal; the network proposes the site-to-class
reassignment of Particle::gibbsPartitionMove (Particle.cpp:785-995, Neal 2000
Algorithm 8), so its context is Algorithm 8's: site s removed from its class,
classes C with sizes n_c (without s), one auxiliary "new" class.

Network (parameters theta, float32 on the GPU, layout below):
    x   = [f_s (20 site amino-acid frequencies), m (20 size-weighted mean profile)]
    h1  = gelu(W1 x + b1)            40 -> 128
    h2  = gelu(W2 h1 + b2)           128 -> 128
    l_c = h2 . (U lp_c + u)          existing class c, lp_c = log pi_c (20)
    l_n = h2 . v + b_new             the "new" slot
    logits: tau * l_c + log n_c  and  tau * l_n + log alpha
    q = softmax(logits)  over K + 1 slots
l approximates a site log-likelihood, so the proposal tempers with the target.
Training loss: mean over examples of -log q(label), label in [0, K] (K = new).
gelu(z) = z * Phi(z) = z erfc(-z / sqrt 2) / 2 (exact).

Pins (tests/test_mlp_oracle.py): central finite differences of the loss in
float64 pin every gradient block; q sums to one; zero weights give the CRP
weights exactly; gelu against scipy's normal cdf; Adam against its published
update (Kingma & Ba 2015, Algorithm 1) on a closed-form quadratic.
"""
import numpy as np
from scipy.special import erfc

H = 128
NIN = 40
W1, B1 = 0, 5120
W2, B2 = 5248, 21632
U, UB = 21760, 24320
V, BN = 24448, 24576
NPARAM = 24577
CLS_STRIDE = 24   # class-table entry: 20 log pi, log n, 3 pad
HDR = 44          # example header: 40 features, tau, log alpha, 2 pad


def unpack(theta):
    t = np.asarray(theta, dtype=np.float64)
    return dict(W1=t[W1:B1].reshape(H, NIN), b1=t[B1:W2], W2=t[W2:B2].reshape(H, H), b2=t[B2:U],
                U=t[U:UB].reshape(H, 20), u=t[UB:V], v=t[V:BN], bn=t[BN])


def gelu(z):
    return 0.5 * z * erfc(-z / np.sqrt(2.0))


def gelu_grad(z):
    return 0.5 * erfc(-z / np.sqrt(2.0)) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def features(freq, logits, sizes):
    """Example header and class table of one Algorithm-8 context.
    freq[20]: the site's amino-acid frequencies; logits[K][20] class logits;
    sizes[K]: class sizes without the site."""
    logits = np.asarray(logits, dtype=np.float64)
    lp = logits - np.log(np.exp(logits - logits.max(1, keepdims=True)).sum(1, keepdims=True)) \
        - logits.max(1, keepdims=True)
    pi = np.exp(lp)
    n = np.asarray(sizes, dtype=np.float64)
    m = (n[:, None] * pi).sum(0) / n.sum()
    cls = np.zeros((len(n), CLS_STRIDE))
    cls[:, :20] = lp
    cls[:, 20] = np.log(n)
    return np.concatenate([freq, m]), cls


def forward(theta, x, tau, log_alpha, cls_list):
    """Logits over K+1 slots for each example; returns (list of logits, cache)."""
    p = unpack(theta)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ p["W1"].T + p["b1"]
    h1 = gelu(z1)
    z2 = h1 @ p["W2"].T + p["b2"]
    h2 = gelu(z2)
    out = []
    for e, cls in enumerate(cls_list):
        cls = np.asarray(cls, dtype=np.float64)
        g = cls[:, :20] @ p["U"].T + p["u"]             # [K, H]
        l = g @ h2[e]
        ln = h2[e] @ p["v"] + p["bn"]
        out.append(np.concatenate([tau[e] * l + cls[:, 20], [tau[e] * ln + log_alpha[e]]]))
    return out, (x, z1, h1, z2, h2)


def log_softmax(v):
    m = v.max()
    return v - m - np.log(np.exp(v - m).sum())


def loss_grad(theta, x, tau, log_alpha, cls_list, labels):
    """Mean cross-entropy and its gradient with respect to theta (float64)."""
    p = unpack(theta)
    logits, (x, z1, h1, z2, h2) = forward(theta, x, tau, log_alpha, cls_list)
    B = len(cls_list)
    g = np.zeros(NPARAM)
    dU = g[U:UB].reshape(H, 20)
    dh2 = np.zeros_like(h2)
    loss = 0.0
    for e in range(B):
        ls = log_softmax(logits[e])
        loss -= ls[labels[e]]
        d = np.exp(ls)
        d[labels[e]] -= 1.0
        d /= B
        dl = tau[e] * d                                 # d loss / d l
        cls = np.asarray(cls_list[e], dtype=np.float64)
        K = len(cls)
        gc = cls[:, :20] @ p["U"].T + p["u"]
        dh2[e] = dl[:K] @ gc + dl[K] * p["v"]
        dU += np.outer(h2[e], dl[:K] @ cls[:, :20])
        g[UB:V] += dl[:K].sum() * h2[e]
        g[V:BN] += dl[K] * h2[e]
        g[BN] += dl[K]
    dz2 = dh2 * gelu_grad(z2)
    g[W2:B2] = (dz2.T @ h1).ravel()
    g[B2:U] = dz2.sum(0)
    dz1 = (dz2 @ p["W2"]) * gelu_grad(z1)
    g[W1:B1] = (dz1.T @ x).ravel()
    g[B1:W2] = dz1.sum(0)
    return loss / B, g


def adam(theta, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1); step counts from 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
