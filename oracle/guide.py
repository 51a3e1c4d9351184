"""Test infrastructure: FP64 numpy restatement of the learned path guide.

Only tests, the smoke check and the bench's CPU-baseline leg use this module,
as the checker of the HIP guide kernels.  The guide has no counterpart in the
reference renderer (LavaFrame samples the BSDF only, pathtrace.glsl:268-291);
it is specified here and pinned by checks that do not retype it:

* ``loss_grad``'s analytic gradients against central finite differences of
  ``loss`` (tests/test_guide.py::test_analytic_gradient_matches_finite_differences);
* ``vmf_pdf`` integrates to 1 over the sphere, and ``vmf_sample`` follows
  ``vmf_pdf`` (histogram against quadrature);
* ``adam_step`` against the closed form of the first Adam step
  (|delta| = lr for every parameter with a non-zero gradient).

Model (the same definitions as lavaframe_amd/csrc/dev_guide.h):
  encoding, 64 features of (x in [0,1]^3 scene-normalised position, n, wo):
    [0, 36)  sin / cos (2^f pi x_a), f = 0..5, a = 0..2   index (a*6+f)*2 + {0 sin, 1 cos}
    [36, 48) sin / cos (2^f pi n_a), f = 0..1             36 + (a*2+f)*2 + s
    [48, 60) sin / cos (2^f pi wo_a), f = 0..1            48 + (a*2+f)*2 + s
    [60, 63) wo;  63: 1
  MLP 64 -> 64 -> 64 -> 64 -> 40, ReLU after the three hidden layers,
  y = x W + b with W stored [in][out].
  Output lobe k (8 lobes): raw direction r = o[5k:5k+3], s = o[5k+3], a = o[5k+4];
  mu = r / |r|, kappa = exp(clip(s, LOGK_MIN, LOGK_MAX)), pi = softmax(a).
  vMF lobe pdf  v(d) = kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.d - 1)).
  Loss over a batch of N samples: L = -(1/N) sum_i w_i log p(d_i | x_i).
"""
from __future__ import annotations

import math

import numpy as np

N_IN = 64
N_HID = 64
N_LOBES = 8
N_OUT = 5 * N_LOBES
LOGK_MIN = -3.0
LOGK_MAX = 7.0
LAYERS = [(N_IN, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_OUT)]


def n_params():
    return sum(i * o + o for i, o in LAYERS)


def layer_slice(l):
    """(offset, size) of layer l's W and b in the flat parameter vector."""
    off = sum(i * o + o for i, o in LAYERS[:l])
    i, o = LAYERS[l]
    return off, i * o + o


def unpack(params):
    """Flat parameter vector -> [(W [in][out], b [out])] in layer order
    (each layer: W row-major, then b)."""
    out, k = [], 0
    for i, o in LAYERS:
        W = params[k:k + i * o].reshape(i, o)
        k += i * o
        b = params[k:k + o]
        k += o
        out.append((W, b))
    return out


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def encode(x, n, wo):
    """(N,3) normalised position, normal, outgoing direction -> (N,64) features."""
    x, n, wo = (np.asarray(a, np.float64) for a in (x, n, wo))
    N = x.shape[0]
    e = np.empty((N, 64))
    for a in range(3):
        for f in range(6):
            ang = (2.0 ** f) * math.pi * x[:, a]
            e[:, (a * 6 + f) * 2] = np.sin(ang)
            e[:, (a * 6 + f) * 2 + 1] = np.cos(ang)
        for f in range(2):
            ang = (2.0 ** f) * math.pi * n[:, a]
            e[:, 36 + (a * 2 + f) * 2] = np.sin(ang)
            e[:, 36 + (a * 2 + f) * 2 + 1] = np.cos(ang)
            ang = (2.0 ** f) * math.pi * wo[:, a]
            e[:, 48 + (a * 2 + f) * 2] = np.sin(ang)
            e[:, 48 + (a * 2 + f) * 2 + 1] = np.cos(ang)
    e[:, 60:63] = wo
    e[:, 63] = 1.0
    return e


def normalise_pos(p, lo, hi):
    lo = np.asarray(lo, np.float64)
    ext = np.maximum(np.asarray(hi, np.float64) - lo, 1e-6)
    return np.clip((np.asarray(p, np.float64) - lo) / ext, 0.0, 1.0)


def forward(params, enc, emulate_bf16=False):
    """Returns the raw outputs (N,40) and the hidden activations [h1, h2, h3].
    With emulate_bf16, every matrix-product operand (inputs, activations,
    weights) is rounded to bfloat16 first, as the MFMA kernels do; sums stay
    in double."""
    layers = unpack(np.asarray(params, np.float64))
    h = np.asarray(enc, np.float64)
    acts = []
    for li, (W, b) in enumerate(layers):
        if emulate_bf16:
            y = bf16(h) @ bf16(W) + b
        else:
            y = h @ W + b
        if li < 3:
            y = np.maximum(y, 0.0)
            acts.append(y)
        h = y
    return h, acts


def lobes(out):
    """Raw outputs (N,40) -> mu (N,8,3), kappa (N,8), pi (N,8)."""
    o = np.asarray(out, np.float64).reshape(-1, N_LOBES, 5)
    r = o[:, :, 0:3]
    rn = np.linalg.norm(r, axis=2, keepdims=True)
    mu = r / np.maximum(rn, 1e-12)
    kappa = np.exp(np.clip(o[:, :, 3], LOGK_MIN, LOGK_MAX))
    a = o[:, :, 4]
    a = a - a.max(axis=1, keepdims=True)
    pi = np.exp(a)
    pi /= pi.sum(axis=1, keepdims=True)
    return mu, kappa, pi


def vmf_log_norm(kappa):
    """log of kappa / (2 pi (1 - exp(-2 kappa)))."""
    return np.log(kappa) - math.log(2 * math.pi) - np.log(-np.expm1(-2.0 * kappa))


def mixture_pdf(mu, kappa, pi, d):
    d = np.asarray(d, np.float64)
    cos = np.einsum("nkc,nc->nk", mu, d)
    return (pi * np.exp(vmf_log_norm(kappa) + kappa * (cos - 1.0))).sum(axis=1)


def vmf_pdf(mu, kappa, d):
    """One lobe: mu (3,), kappa scalar, d (M,3)."""
    cos = np.asarray(d, np.float64) @ np.asarray(mu, np.float64)
    return np.exp(vmf_log_norm(kappa) + kappa * (cos - 1.0))


def vmf_sample(mu, kappa, u1, u2):
    """Inverse-CDF sampling of one vMF lobe (Jakob 2012, "Numerically stable
    sampling of the von Mises Fisher distribution on S2"): cos theta =
    1 + log(u1 + (1 - u1) exp(-2 kappa)) / kappa about mu, phi = 2 pi u2."""
    mu = np.asarray(mu, np.float64)
    w = 1.0 + np.log(u1 + (1.0 - u1) * np.exp(-2.0 * kappa)) / kappa
    w = np.clip(w, -1.0, 1.0)
    s = np.sqrt(np.maximum(0.0, 1.0 - w * w))
    t, b = frame(mu)
    ph = 2 * math.pi * u2
    return w[:, None] * mu + s[:, None] * (np.cos(ph)[:, None] * t + np.sin(ph)[:, None] * b)


def frame(n):
    """Tangent frame used by the device sampler (Duff et al. 2017, "Building an
    orthonormal basis, revisited")."""
    sgn = 1.0 if n[2] >= 0 else -1.0
    a = -1.0 / (sgn + n[2])
    b = n[0] * n[1] * a
    t = np.array([1.0 + sgn * n[0] * n[0] * a, sgn * b, -sgn * n[0]])
    bb = np.array([b, sgn + n[1] * n[1] * a, -n[1]])
    return t, bb


def loss(params, enc, dirs, w):
    out, _ = forward(params, enc)
    mu, kappa, pi = lobes(out)
    p = mixture_pdf(mu, kappa, pi, dirs)
    return float(-(np.asarray(w, np.float64) * np.log(p)).sum() / len(w))


def output_grad(out, dirs, w):
    """d loss / d raw outputs (N,40), analytically (responsibilities gamma_k =
    pi_k v_k / p):
      d/da_k = (w/N) (pi_k - gamma_k)
      d/ds_k = -(w/N) gamma_k kappa_k (1/kappa_k - 2 e^{-2 kappa_k} / (1 - e^{-2 kappa_k}) + mu_k.d - 1)
               (zero where s_k is clipped)
      d/dr_k = -(w/N) gamma_k kappa_k (d - (mu_k.d) mu_k) / |r_k|"""
    out = np.asarray(out, np.float64)
    N = out.shape[0]
    o = out.reshape(N, N_LOBES, 5)
    r = o[:, :, 0:3]
    rn = np.maximum(np.linalg.norm(r, axis=2), 1e-12)
    mu, kappa, pi = lobes(out)
    d = np.asarray(dirs, np.float64)
    cos = np.einsum("nkc,nc->nk", mu, d)
    logv = vmf_log_norm(kappa) + kappa * (cos - 1.0)
    lw = np.log(pi) + logv
    m = lw.max(axis=1, keepdims=True)
    gamma = np.exp(lw - m)
    gamma /= gamma.sum(axis=1, keepdims=True)
    sc = (np.asarray(w, np.float64) / N)[:, None]
    g = np.zeros_like(o)
    g[:, :, 4] = sc * (pi - gamma)
    e2 = np.exp(-2.0 * kappa)
    dlogv = 1.0 / kappa - 2.0 * e2 / (-np.expm1(-2.0 * kappa)) + (cos - 1.0)
    inside = (o[:, :, 3] > LOGK_MIN) & (o[:, :, 3] < LOGK_MAX)
    g[:, :, 3] = np.where(inside, -sc * gamma * kappa * dlogv, 0.0)
    tang = d[:, None, :] - cos[:, :, None] * mu
    g[:, :, 0:3] = -(sc * gamma * kappa / rn)[:, :, None] * tang
    return g.reshape(N, N_OUT)


def loss_grad(params, enc, dirs, w):
    """Gradient of ``loss`` w.r.t. the flat parameter vector (backpropagation
    through the ReLU MLP from ``output_grad``)."""
    layers = unpack(np.asarray(params, np.float64))
    out, acts = forward(params, enc)
    g = output_grad(out, dirs, w)
    ins = [np.asarray(enc, np.float64)] + acts
    grads = [None] * 4
    for li in range(3, -1, -1):
        W, _ = layers[li]
        x = ins[li]
        grads[li] = (x.T @ g, g.sum(axis=0))
        if li > 0:
            g = (g @ W.T) * (ins[li] > 0)
    return np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])


def adam_step(params, grad, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1); t counts from 1.
    Returns new (params, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
