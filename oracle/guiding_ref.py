"""TEST INFRASTRUCTURE ONLY: float64 numpy restatement of the path-guiding model.

The guiding proposal has no counterpart in the reference (it is an addition
of this build).  Only tests/ and bench.py's checks use this module.

Model (matches rainbow_amd/csrc/guiding.hip):
  inputs    x in [0,1]^9: position normalised by the scene bounds, (normal+1)/2, (wo+1)/2
  encoding  for each scalar v and frequency k = 0..11: sin(2^k pi v), cos(2^k pi v)
            -> 216 features, zero-padded to 224 columns
  MLP       224 -> 64 -> 64 -> 64 -> 64 -> 64 (ReLU) -> 40 (linear), weights W[l] (out x in), b[l]
  head      8 lobes j: raw o[5j..5j+4] = (m_x, m_y, m_z, k, a)
            mu_j = m / |m|, kappa_j = softplus(k) + KAPPA_MIN (capped at KAPPA_MAX),
            w = softmax(a)
  pdf       q(w) = sum_j w_j * vMF(w; mu_j, kappa_j),
            vMF = kappa / (2 pi (1 - exp(-2 kappa))) * exp(kappa (mu.w - 1))
  loss      per sample i with weight c_i = L_i / p_i (radiance over sampling pdf):
            loss = -sum_i c_i log q(w_i) / sum_i c_i
  Adam      beta1 0.9, beta2 0.99, eps 1e-8 (bias-corrected)
"""
from __future__ import annotations

import math

import numpy as np

N_FREQ = 12
N_IN = 9
N_FEAT = N_IN * N_FREQ * 2   # 216
N_FEAT_PAD = 224
HIDDEN = 64
N_HIDDEN_LAYERS = 5
N_LOBES = 8
N_OUT = 5 * N_LOBES          # 40
KAPPA_MIN = 0.01
KAPPA_MAX = 500.0
LAYER_DIMS = [(N_FEAT_PAD, HIDDEN)] + [(HIDDEN, HIDDEN)] * (N_HIDDEN_LAYERS - 1) + [(HIDDEN, N_OUT)]


# padded parameter layout shared with the HIP kernels (include/rbw.h RBW_GUIDE_PARAMS)
PAD_IN = [224, 64, 64, 64, 64, 64]
PAD_OUT = [64, 64, 64, 64, 64, 48]


def to_padded(Ws, bs):
    out = []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        Wp = np.zeros((PAD_OUT[l], PAD_IN[l]))
        Wp[:W.shape[0], :W.shape[1]] = W
        bp = np.zeros(PAD_OUT[l])
        bp[:len(b)] = b
        out += [Wp.reshape(-1), bp]
    return np.concatenate(out).astype(np.float32)


def from_padded(flat):
    Ws, bs = [], []
    o = 0
    for l, (fi, fo) in enumerate(LAYER_DIMS):
        Wp = flat[o:o + PAD_OUT[l] * PAD_IN[l]].reshape(PAD_OUT[l], PAD_IN[l])
        o += PAD_OUT[l] * PAD_IN[l]
        bp = flat[o:o + PAD_OUT[l]]
        o += PAD_OUT[l]
        Ws.append(np.array(Wp[:fo, :fi], np.float64))
        bs.append(np.array(bp[:fo], np.float64))
    return Ws, bs


def n_params(padded=False):
    tot = 0
    for i, (fi, fo) in enumerate(LAYER_DIMS):
        fi_real = N_FEAT if (i == 0 and not padded) else fi
        tot += fi_real * fo + fo
    return tot


def encode(x):
    """x (n,9) in [0,1] -> (n,224) features; column f = 24*v + 2*k + (0 sin | 1 cos)."""
    x = np.asarray(x, np.float64)
    n = x.shape[0]
    out = np.zeros((n, N_FEAT_PAD))
    for v in range(N_IN):
        for k in range(N_FREQ):
            a = (2.0 ** k) * math.pi * x[:, v]
            out[:, 24 * v + 2 * k] = np.sin(a)
            out[:, 24 * v + 2 * k + 1] = np.cos(a)
    return out


def init_params(seed=0):
    """He-uniform init of W (out x in) and zero biases; padded input columns are zero."""
    rng = np.random.default_rng(seed)
    Ws, bs = [], []
    for i, (fi, fo) in enumerate(LAYER_DIMS):
        fan_in = N_FEAT if i == 0 else fi
        lim = math.sqrt(6.0 / fan_in) * (0.5 if i == len(LAYER_DIMS) - 1 else 1.0)
        W = rng.uniform(-lim, lim, (fo, fi))
        if i == 0:
            W[:, N_FEAT:] = 0.0
        Ws.append(W)
        bs.append(np.zeros(fo))
    return Ws, bs


def forward(Ws, bs, X):
    """returns raw outputs (n,40) and the list of layer inputs/pre-activations for backward"""
    acts = [X]
    pre = []
    A = X
    for i, (W, b) in enumerate(zip(Ws, bs)):
        Z = A @ W.T + b
        pre.append(Z)
        A = np.maximum(Z, 0.0) if i < len(Ws) - 1 else Z
        acts.append(A)
    return A, acts, pre


def softplus(x):
    return np.where(x > 20, x, np.log1p(np.exp(np.minimum(x, 20))))


def sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def head(o):
    """raw outputs (n,40) -> mu (n,8,3), kappa (n,8), weights (n,8)"""
    o = o.reshape(-1, N_LOBES, 5)
    m = o[:, :, 0:3]
    nm = np.linalg.norm(m, axis=-1, keepdims=True)
    mu = m / np.maximum(nm, 1e-12)
    kappa = np.minimum(softplus(o[:, :, 3]) + KAPPA_MIN, KAPPA_MAX)
    a = o[:, :, 4]
    a = a - a.max(axis=1, keepdims=True)
    w = np.exp(a)
    w /= w.sum(axis=1, keepdims=True)
    return mu, kappa, w


def vmf_pdf(d, mu, kappa):
    """d (n,3), mu (n,...,3), kappa (n,...) -> pdf"""
    cos = np.einsum("n...c,nc->n...", mu, d) if mu.ndim == 3 else (mu * d).sum(-1)
    norm = kappa / (2 * math.pi * (-np.expm1(-2 * kappa)))
    return norm * np.exp(kappa * (cos - 1.0))


def mixture_pdf(o, d):
    mu, kappa, w = head(o)
    return (w * vmf_pdf(d, mu, kappa)).sum(axis=1)


def loss_and_grad_out(o, d, c):
    """loss and d loss / d o for raw outputs o (n,40), directions d (n,3), weights c (n,)"""
    n = o.shape[0]
    o3 = o.reshape(n, N_LOBES, 5)
    mu, kappa, w = head(o)
    v = vmf_pdf(d, mu, kappa)                       # (n,8)
    q = (w * v).sum(axis=1)                         # (n,)
    csum = c.sum()
    loss = -(c * np.log(q)).sum() / csum
    r = w * v / q[:, None]                          # responsibilities
    g = -(c / csum)[:, None]                        # d loss / d log q  per sample (n,1)
    dlogit = g * (r - w)
    cos = np.einsum("njc,nc->nj", mu, d)
    dlogv_dk = 1.0 / kappa - 2.0 * np.exp(-2 * kappa) / (-np.expm1(-2 * kappa)) + (cos - 1.0)
    sp = softplus(o3[:, :, 3]) + KAPPA_MIN
    dk_draw = np.where(sp < KAPPA_MAX, sigmoid(o3[:, :, 3]), 0.0)
    dkraw = g * r * dlogv_dk * dk_draw
    # d log v / d mu = kappa d ; mu = m/|m| -> d/dm = (I - mu mu^T)/|m| applied
    m = o3[:, :, 0:3]
    nm = np.maximum(np.linalg.norm(m, axis=-1, keepdims=True), 1e-12)
    gmu = (g * r * kappa)[:, :, None] * d[:, None, :]
    dm = (gmu - (gmu * mu).sum(-1, keepdims=True) * mu) / nm
    grad = np.zeros_like(o3)
    grad[:, :, 0:3] = dm
    grad[:, :, 3] = dkraw
    grad[:, :, 4] = dlogit
    return loss, grad.reshape(n, N_OUT)


def backward(Ws, bs, acts, pre, dout):
    """gradients of W and b given d loss / d raw output"""
    dWs = [None] * len(Ws)
    dbs = [None] * len(Ws)
    dZ = dout
    for i in range(len(Ws) - 1, -1, -1):
        A = acts[i]
        dWs[i] = dZ.T @ A
        dbs[i] = dZ.sum(axis=0)
        if i > 0:
            dA = dZ @ Ws[i]
            dZ = dA * (pre[i - 1] > 0)
    return dWs, dbs


# ---- the kernels' rounding, restated (FP16 operands, wide accumulation)

def f16(a):
    """round to IEEE half and back"""
    return np.asarray(a, np.float64).astype(np.float16).astype(np.float64)


def encode_f16(x):
    """the kernels' encoding: t = 2^k x / 2 in float32 (exact), the fraction of t in float32,
    sin and cos of 2 pi times it rounded to half"""
    x32 = np.asarray(x, np.float32)
    out = np.zeros((x32.shape[0], N_FEAT_PAD))
    for v in range(N_IN):
        for k in range(N_FREQ):
            t = x32[:, v] * np.float32(2.0 ** k * 0.5)
            f = (t - np.floor(t)).astype(np.float64)
            out[:, 24 * v + 2 * k] = f16(np.sin(2 * math.pi * f))
            out[:, 24 * v + 2 * k + 1] = f16(np.cos(2 * math.pi * f))
    return out


def forward_f16(Ws, bs, x):
    """raw outputs with half weights and half post-ReLU activations (biases and sums wide);
    returns (outputs, layer inputs)"""
    A = encode_f16(x)
    acts = [A]
    for i, (W, b) in enumerate(zip(Ws, bs)):
        Wh = f16(W)
        if i == 0:
            Wh = np.pad(Wh, ((0, 0), (0, N_FEAT_PAD - Wh.shape[1]))) if Wh.shape[1] < N_FEAT_PAD else Wh
        Z = A @ Wh.T + b
        if i < len(Ws) - 1:
            A = f16(np.maximum(Z, 0.0))
            acts.append(A)
        else:
            A = Z
    return A, acts


def grads_f16(Ws, bs, x, d, c):
    """loss and mean gradients as the training kernel forms them: the head's per-record
    gradient at weight c / mean(c) rounded to half (clamped to +-30000), data gradients
    rounded to half after the ReLU mask, weight gradients summed wide over the records"""
    n = x.shape[0]
    o, acts = forward_f16(Ws, bs, x)
    loss, dout = loss_and_grad_out(o, d, c)
    clip = lambda a: np.clip(a, -30000.0, 30000.0)
    dZ = f16(clip(dout * n))
    dWs, dbs = [None] * len(Ws), [None] * len(Ws)
    for i in range(len(Ws) - 1, -1, -1):
        A = acts[i]
        dW = dZ.T @ A / n
        dWs[i] = dW[:, :N_FEAT] if i == 0 else dW
        dbs[i] = dZ.sum(axis=0) / n
        if i > 0:
            dA = dZ @ f16(Ws[i])
            dZ = f16(clip(dA)) * (A > 0)
    return loss, dWs, dbs


def loss_fn(Ws, bs, X, d, c):
    o, _, _ = forward(Ws, bs, X)
    return loss_and_grad_out(o, d, c)[0]


class Adam:
    def __init__(self, params, lr=1e-2, b1=0.9, b2=0.99, eps=1e-8):
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps
        self.m = [np.zeros_like(p) for p in params]
        self.v = [np.zeros_like(p) for p in params]
        self.t = 0

    def step(self, params, grads):
        self.t += 1
        out = []
        for i, (p, g) in enumerate(zip(params, grads)):
            self.m[i] = self.b1 * self.m[i] + (1 - self.b1) * g
            self.v[i] = self.b2 * self.v[i] + (1 - self.b2) * g * g
            mh = self.m[i] / (1 - self.b1 ** self.t)
            vh = self.v[i] / (1 - self.b2 ** self.t)
            out.append(p - self.lr * mh / (np.sqrt(vh) + self.eps))
        return out


def vmf_sample(mu, kappa, u1, u2):
    """Jakob's vMF sampler around mu (3,), scalar kappa; returns a unit vector."""
    W = 1.0 + math.log(u1 + (1.0 - u1) * math.exp(-2.0 * kappa)) / kappa
    s = math.sqrt(max(0.0, 1.0 - W * W))
    phi = 2.0 * math.pi * u2
    local = np.array([s * math.cos(phi), s * math.sin(phi), W])
    z = np.asarray(mu, float)
    a = np.array([1.0, 0.0, 0.0]) if abs(z[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    x = np.cross(a, z)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    return local[0] * x + local[1] * y + local[2] * z
