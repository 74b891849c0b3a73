"""Test infrastructure only: FP64 numpy oracle of the path-guiding network.

Restates lambda_amd/csrc/guide.h and guide_mlp.h (which have no counterpart in
the reference): frequency encoding, the 5-layer MLP, the von Mises-Fisher
mixture built from its 40 outputs, the loss -c log q(w) with its gradient,
backpropagation and Adam.  With bf16=True it rounds to bfloat16 (round to
nearest even) exactly where the GPU stores bfloat16: the encoded input, the
weights, the hidden activations and the backpropagated dZ; everything else is
float64.  Pinned by finite differences of the float64 loss
(tests/test_guide_oracle.py) and by the closed forms of the vMF density.
"""
from __future__ import annotations

import numpy as np

SCALARS, BANDS = 6, 8
N_IN = SCALARS * BANDS * 2
HIDDEN = 64
LAYERS = 5
LOBES = 8
N_OUT = 5 * LOBES
KAPPA_MIN, KAPPA_MAX = 1e-3, 2000.0


def layer_in(l: int) -> int:
    return N_IN if l == 0 else HIDDEN


def w_offset(l: int) -> int:
    return 0 if l == 0 else N_IN * 64 + 64 + (l - 1) * (64 * 64 + 64)


def b_offset(l: int) -> int:
    return w_offset(l) + layer_in(l) * 64


N_PARAMS = N_IN * 64 + 64 + 4 * (64 * 64 + 64)


def unpack(params: np.ndarray):
    """Per layer (W [K, 64], b [64]) views of a flat parameter vector."""
    out = []
    for l in range(LAYERS):
        k = layer_in(l)
        w = params[w_offset(l):w_offset(l) + k * 64].reshape(k, 64)
        b = params[b_offset(l):b_offset(l) + 64]
        out.append((w, b))
    return out


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) through float32, returned as float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def encode(features: np.ndarray) -> np.ndarray:
    """[n, 6] features in [0, 1] -> [n, 96]: input i = scalar i // 16, band
    (i % 16) // 2, sin for even i, cos for odd i."""
    f = np.asarray(features, np.float64)
    i = np.arange(N_IN)
    s, b = i // 16, (i % 16) // 2
    x = f[:, s] * (2.0 ** b) * np.pi
    return np.where(i % 2 == 1, np.cos(x), np.sin(x))


def forward(params: np.ndarray, features: np.ndarray, bf: bool = False):
    """Returns (outputs [n, 64], activations [input, H1..H4])."""
    rnd = bf16 if bf else (lambda a: a)
    h = rnd(encode(features))
    acts = [h]
    layers = unpack(np.asarray(params, np.float64))
    for l, (w, b) in enumerate(layers):
        z = h @ rnd(w) + b
        if l < LAYERS - 1:
            h = rnd(np.maximum(z, 0.0))
            acts.append(h)
        else:
            return z, acts
    raise AssertionError


def kappa_of(raw):
    """Concentration from its raw output: exp, clamped (guide.h guide_lobes)."""
    return np.exp(np.minimum(raw, 80.0))


def lobes(o: np.ndarray):
    """(mu [n, 8, 3], kappa [n, 8], weight [n, 8]) from raw outputs [n, >=40]."""
    q = o[:, :N_OUT].reshape(-1, LOBES, 5)
    m = q[:, :, :3]
    l = np.maximum(np.linalg.norm(m, axis=2), 1e-6)
    mu = m / l[:, :, None]
    kappa = np.clip(kappa_of(q[:, :, 3]), KAPPA_MIN, KAPPA_MAX)
    logits = q[:, :, 4]
    wgt = np.exp(logits - logits.max(axis=1, keepdims=True))
    wgt /= wgt.sum(axis=1, keepdims=True)
    return mu, kappa, wgt


def vmf_norm(kappa):
    return kappa / (2 * np.pi * -np.expm1(-2 * kappa))


def pdf(o: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Mixture density of directions w [n, 3] under outputs o [n, >=40]."""
    mu, kappa, wgt = lobes(o)
    comp = vmf_norm(kappa) * np.exp(kappa * (np.einsum("nkd,nd->nk", mu, w) - 1))
    return (wgt * comp).sum(axis=1)


def loss_grad(o: np.ndarray, w: np.ndarray, c: np.ndarray):
    """(loss [n], d loss / d o [n, 40]) of -c log q(w)."""
    o = np.asarray(o, np.float64)
    mu, kappa, wgt = lobes(o)
    cosang = np.einsum("nkd,nd->nk", mu, w)
    comp = vmf_norm(kappa) * np.exp(kappa * (cosang - 1))
    q = np.maximum((wgt * comp).sum(axis=1), 1e-30)
    resp = wgt * comp / q[:, None]
    c = np.asarray(c, np.float64)[:, None]
    d = np.zeros((len(o), LOBES, 5))
    d[:, :, 4] = -c * (resp - wgt)
    e = np.exp(-2 * kappa)
    dlog_dk = 1 / kappa - 2 * e / (-np.expm1(-2 * kappa)) + (cosang - 1)
    raw = o[:, :N_OUT].reshape(-1, LOBES, 5)
    ek = kappa_of(raw[:, :, 3])
    dk_draw = np.where((ek > KAPPA_MIN) & (ek < KAPPA_MAX), ek, 0.0)
    d[:, :, 3] = -c * resp * dlog_dk * dk_draw
    m = raw[:, :, :3]
    l = np.maximum(np.linalg.norm(m, axis=2), 1e-6)
    gm = kappa[:, :, None] * w[:, None, :]
    gm = (gm - mu * np.einsum("nkd,nkd->nk", mu, gm)[:, :, None]) / l[:, :, None]
    d[:, :, :3] = -(c * resp)[:, :, None] * gm
    return -c[:, 0] * np.log(q), d.reshape(-1, N_OUT)


def loss(params: np.ndarray, samples: np.ndarray) -> float:
    """Summed loss of samples [n, 10] = {features[6], direction[3], c}."""
    s = np.asarray(samples, np.float64)
    o, _ = forward(params, s[:, :6])
    l, _ = loss_grad(o, s[:, 6:9], s[:, 9])
    return float(l.sum())


def gradient(params: np.ndarray, samples: np.ndarray, bf: bool = False):
    """(summed gradient [N_PARAMS], summed loss) over samples [n, 10]."""
    rnd = bf16 if bf else (lambda a: a)
    s = np.asarray(samples, np.float64)
    o, acts = forward(params, s[:, :6], bf)
    lval, d = loss_grad(o, s[:, 6:9], s[:, 9])
    dz = np.zeros((len(s), 64))
    dz[:, :N_OUT] = d
    dz = rnd(dz)
    layers = unpack(np.asarray(params, np.float64))
    g = np.zeros(N_PARAMS)
    for l in range(LAYERS - 1, -1, -1):
        w, _ = layers[l]
        hprev = acts[l]
        g[w_offset(l):w_offset(l) + w.size] = (hprev.T @ dz).ravel()
        g[b_offset(l):b_offset(l) + 64] = dz.sum(axis=0)
        if l > 0:
            dh = dz @ rnd(w).T
            dz = rnd(np.where(hprev > 0, dh, 0.0))
    return g, float(lval.sum())


def adam(params, m, v, grad, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step; returns (params, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
