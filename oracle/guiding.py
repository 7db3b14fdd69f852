"""Path-guiding oracle - TEST INFRASTRUCTURE ONLY.

Float64 numpy restatement of the guiding model that the GPU kernels
implement (craystal_amd/csrc/guiding.hip).  It has no counterpart in the
reference; it is pinned by finite differences (backward), quadrature (vMF
normalisation), histograms (vMF sampling) and a closed-form Adam step
(tests/test_guiding_oracle.py).

Model
  features  p = (x - bmin) / (bmax - bmin); for axis a, level l < 5:
            sin(2^l pi p_a), cos(2^l pi p_a)  -> 30; then 2 zeros (32);
            shading normal n (3); outgoing direction wo (3); zero pad to 64.
  MLP       64 -> w -> w -> w -> w -> 48 (ReLU after the first four; hidden
            width w = 64, or 128), the last layer's outputs 40..47 unused.
  head      for lobe j < 8: a = o[5j], m = o[5j+1:5j+4], k = o[5j+4];
            weights softmax(a), mean m/|m|, kappa softplus(k) + 0.01.
  vMF       v(w; mu, k) = k / (2 pi (1 - e^{-2k})) * exp(k (mu.w - 1))
  loss      per sample  -(y / q_s) * log q(w)  (KL divergence up to a
            constant), y = target radiance, q_s = the pdf w was sampled
            with; averaged over the batch.
  Adam      beta1 0.9, beta2 0.999, eps 1e-8.
"""
from __future__ import annotations

import numpy as np

N_IN = 64
WIDTH = 64
N_OUT = 48
N_LOBES = 8
N_FREQ = 5
LAYERS = [(N_IN, WIDTH), (WIDTH, WIDTH), (WIDTH, WIDTH), (WIDTH, WIDTH), (WIDTH, N_OUT)]


def layers(width=WIDTH):
    """(fan-in, fan-out) of the 5 layers of a network of hidden width `width`."""
    return [(N_IN, width), (width, width), (width, width), (width, width), (width, N_OUT)]
KAPPA_MIN = 0.01


def bf16(x):
    """Round to bfloat16 (nearest even) and back, as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def features(pos, nrm, wo, bmin, bmax):
    """(N, 64) float64 input rows."""
    pos = np.asarray(pos, np.float64)
    ext = np.asarray(bmax, np.float64) - np.asarray(bmin, np.float64)
    p = (pos - np.asarray(bmin, np.float64)) / np.where(ext > 0, ext, 1.0)
    n = len(pos)
    X = np.zeros((n, N_IN))
    for a in range(3):
        for l in range(N_FREQ):
            X[:, a * 10 + 2 * l] = np.sin((2.0 ** l) * np.pi * p[:, a])
            X[:, a * 10 + 2 * l + 1] = np.cos((2.0 ** l) * np.pi * p[:, a])
    X[:, 32:35] = nrm
    X[:, 35:38] = wo
    return X


def init_params(seed=1234, width=WIDTH):
    """He-uniform weights (fan-in), zero biases; float64 master copy."""
    rng = np.random.default_rng(seed)
    P = []
    for fi, fo in layers(width):
        lim = np.sqrt(6.0 / fi)
        P.append([rng.uniform(-lim, lim, (fi, fo)), np.zeros(fo)])
    return P


def forward(P, X, quant=False):
    """Returns (outputs (N, 48), activations list incl. input)."""
    acts = [bf16(X) if quant else X]
    h = acts[0]
    for i, (W, b) in enumerate(P):
        Wq = bf16(W) if quant else W
        z = h @ Wq + b
        if i < len(P) - 1:
            h = np.maximum(z, 0.0)
            if quant:
                h = bf16(h)
            acts.append(h)
        else:
            out = z
    return out, acts


def softplus(x):
    return np.where(x > 20, x, np.log1p(np.exp(np.minimum(x, 20))))


def head(out):
    """(weights (N,8), means (N,8,3), kappas (N,8))."""
    o = out[:, :5 * N_LOBES].reshape(-1, N_LOBES, 5)
    a = o[:, :, 0]
    a = a - a.max(1, keepdims=True)
    w = np.exp(a)
    w /= w.sum(1, keepdims=True)
    m = o[:, :, 1:4]
    mu = m / np.maximum(np.linalg.norm(m, axis=2, keepdims=True), 1e-6)
    kap = softplus(o[:, :, 4]) + KAPPA_MIN
    return w, mu, kap


def vmf_pdf(mu, kap, d):
    """Stable vMF density; mu (...,3), kap (...), d (...,3) broadcastable."""
    c = kap / (2 * np.pi * (-np.expm1(-2 * kap)))
    return c * np.exp(kap * (np.sum(mu * d, -1) - 1))


def mixture_pdf(out, d):
    w, mu, kap = head(out)
    return np.sum(w * vmf_pdf(mu, kap, d[:, None, :]), 1)


def sample_vmf(mu, kap, u1, u2):
    """Jakob (2012), numerically stable."""
    w = 1 + np.log(u1 + (1 - u1) * np.exp(-2 * kap)) / kap
    s = np.sqrt(np.maximum(0.0, 1 - w * w))
    phi = 2 * np.pi * u2
    a = np.where(np.abs(mu[..., 0:1]) > 0.9, [0.0, 1.0, 0.0], [1.0, 0.0, 0.0])
    t = np.cross(a, mu)
    t /= np.linalg.norm(t, axis=-1, keepdims=True)
    b = np.cross(mu, t)
    return w[..., None] * mu + s[..., None] * (np.cos(phi)[..., None] * t + np.sin(phi)[..., None] * b)


def loss_and_grad_out(out, d, y, q_s):
    """Mean loss and d loss / d out (N, 48)."""
    n = len(out)
    o = out[:, :5 * N_LOBES].reshape(-1, N_LOBES, 5)
    w, mu, kap = head(out)
    v = vmf_pdf(mu, kap, d[:, None, :])
    q = np.sum(w * v, 1)
    wt = y / q_s
    loss = float(np.mean(-wt * np.log(q)))
    g = -wt / n                                   # d loss / d log q
    resp = w * v / q[:, None]                     # pi_j v_j / q
    G = np.zeros_like(out)
    Go = G[:, :5 * N_LOBES].reshape(-1, N_LOBES, 5)
    Go[:, :, 0] = g[:, None] * (resp - w)
    k = o[:, :, 4]
    e = np.exp(-2 * kap)
    dlogv_dk = 1 / kap - 2 * e / (1 - e) + (np.sum(mu * d[:, None, :], 2) - 1)
    sig = 1 / (1 + np.exp(-k))
    Go[:, :, 4] = g[:, None] * resp * dlogv_dk * sig
    m = o[:, :, 1:4]
    nm = np.maximum(np.linalg.norm(m, axis=2, keepdims=True), 1e-6)
    dmu = (g[:, None] * resp * kap)[:, :, None] * d[:, None, :]      # d/d mu
    proj = dmu - np.sum(dmu * mu, 2, keepdims=True) * mu
    Go[:, :, 1:4] = proj / nm
    G[:, :5 * N_LOBES] = Go.reshape(n, -1)
    return loss, G


def backward(P, acts, G):
    """Parameter gradients [(dW, db), ...] given d loss / d outputs."""
    grads = [None] * len(P)
    delta = G
    for i in range(len(P) - 1, -1, -1):
        W, _ = P[i]
        h = acts[i]
        grads[i] = (h.T @ delta, delta.sum(0))
        if i > 0:
            delta = (delta @ W.T) * (acts[i] > 0)
    return grads


def loss_fn(P, X, d, y, q_s):
    out, _ = forward(P, X)
    return loss_and_grad_out(out, d, y, q_s)[0]


class Adam:
    def __init__(self, P, lr=1e-2, b1=0.9, b2=0.999, eps=1e-8):
        self.lr, self.b1, self.b2, self.eps, self.t = lr, b1, b2, eps, 0
        self.m = [[np.zeros_like(W), np.zeros_like(b)] for W, b in P]
        self.v = [[np.zeros_like(W), np.zeros_like(b)] for W, b in P]

    def step(self, P, grads):
        self.t += 1
        c1 = 1 - self.b1 ** self.t
        c2 = 1 - self.b2 ** self.t
        for i, (gW, gb) in enumerate(grads):
            for j, g in enumerate((gW, gb)):
                self.m[i][j] = self.b1 * self.m[i][j] + (1 - self.b1) * g
                self.v[i][j] = self.b2 * self.v[i][j] + (1 - self.b2) * g * g
                P[i][j] = P[i][j] - self.lr * (self.m[i][j] / c1) / (np.sqrt(self.v[i][j] / c2) + self.eps)
        return P
