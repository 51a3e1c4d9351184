"""Test infrastructure: numpy restatement of the reference's post-process pass
(shaders/postprocess.glsl:1-171) and of GetOutputBuffer's RGB8 read-back
(TiledRenderer.cpp:382-397).  Only tests use it, as the checker of the HIP
post-process kernel.

The pass reads the accumulation texture (RGB32F, GL_LINEAR, GL_MIRRORED_REPEAT,
TiledRenderer.cpp:163-171) at the pixel's texture coordinate
((x + 0.5) / W, (y + 0.5) / H), scales by invSampleCounter, optionally with
chromatic aberration (three channel reads at shifted coordinates), applies
tonemap tonemapIndex (0-6) and an optional vignette.  The RGB8 read-back
converts each clamped channel c to round(255 c).

Pinned by tests/test_post.py against closed forms (tonemap values at chosen
inputs, bilinear reads at texel centres, the mirrored-repeat index map on
small textures, the vignette at the centre) rather than by retyping the
formulas; GL's exact filtering arithmetic is the driver's and is followed as
the GL specification states it (parity unpinned beyond that).
"""
from __future__ import annotations

import numpy as np


def mirror_index(i, n):
    """GL_MIRRORED_REPEAT texel index (OpenGL 4.6 spec, 8.14.2): mirror(i) =
    i for i >= 0, -(1 + i) otherwise, on the period 2n."""
    i = np.asarray(i, np.int64)
    m = np.mod(i, 2 * n)
    return np.where(m < n, m, 2 * n - 1 - m)


def sample_linear(tex, u, v):
    """GL_LINEAR read of an (H, W, C) texture at normalised (u, v) (v from the
    bottom row), mirrored-repeat wrap."""
    H, W = tex.shape[:2]
    x = np.asarray(u, np.float64) * W - 0.5
    y = np.asarray(v, np.float64) * H - 0.5
    i0 = np.floor(x).astype(np.int64)
    j0 = np.floor(y).astype(np.int64)
    a = x - i0
    b = y - j0
    i0m, i1m = mirror_index(i0, W), mirror_index(i0 + 1, W)
    j0m, j1m = mirror_index(j0, H), mirror_index(j0 + 1, H)
    t00 = tex[j0m, i0m]
    t10 = tex[j0m, i1m]
    t01 = tex[j1m, i0m]
    t11 = tex[j1m, i1m]
    a = a[..., None]
    b = b[..., None]
    return (1 - a) * (1 - b) * t00 + a * (1 - b) * t10 + (1 - a) * b * t01 + a * b * t11


def tonemap_linear(c, limit):
    lum = 0.3 * c[..., 0:1] + 0.6 * c[..., 1:2] + 0.1 * c[..., 2:3]
    return c * 1.0 / (1.0 + lum / limit)


def tonemap_aces(c):
    a, b, y, d, e = 2.51, 0.03, 2.43, 0.59, 0.14
    return np.clip((c * (a * c + b)) / (c * (y * c + d) + e), 0.0, 1.0)


def tonemap_reinhard(c):
    return np.clip(c / (c + 1), 0.0, 1.0)


def tonemap_hejl(c):
    c = np.maximum(0.0, c - 0.004)
    return (c * (6.2 * c + 0.5)) / (c * (6.2 * c + 1.7) + 0.06)


def tonemap_kanjero(c):
    a, b, y, d, e = 1.2295, 0.3135, 1.1935, 0.4655, 0.073
    c = np.power((c * (c * (a * c + b) + y * d) / (c * (y * c + d) + e)), 1.7)
    c = np.power(c, 1.0 / 0.8)
    c = c * 0.8
    return np.clip(c, 0.0, 1.0)


def tonemap_uncharted2(c):
    A, B, C, D, E, F = 0.15, 0.50, 0.10, 0.20, 0.02, 0.30
    return ((c * (A * c + C * B) + D * E) / (c * (A * c + B) + D * F)) - E / F


def gamma(c):
    return np.power(c, 1.0 / 2.2)


def tonemap(c, index):
    if index == 1:
        return gamma(tonemap_linear(c, 2.0))
    if index == 2:
        return gamma(tonemap_aces(c))
    if index == 3:
        return gamma(tonemap_reinhard(c))
    if index == 4:
        return gamma(tonemap_kanjero(c))
    if index == 5:
        return tonemap_hejl(c)
    if index == 6:
        return gamma(tonemap_uncharted2(c)) * 1.75
    return c


def postprocess(accum, inv, pp):
    """accum: (H, W, 3) sums, row 0 at the bottom; pp: dict of the options
    (tonemap_index, use_ca, use_ca_distortion, ca_distance, ca_p1..p3,
    use_vignette, vignette_intensity, vignette_power)."""
    H, W = accum.shape[:2]
    acc = np.asarray(accum, np.float64)
    u = ((np.arange(W) + 0.5) / W)[None, :].repeat(H, 0)
    v = ((np.arange(H) + 0.5) / H)[:, None].repeat(W, 1)
    if pp.get("use_ca"):
        off = pp["ca_distance"]
        out = np.zeros((H, W, 3))
        if pp.get("use_ca_distortion"):
            dist = np.power(np.hypot(u - pp["ca_p3"], v - pp["ca_p3"]), pp["ca_p1"]) * pp["ca_p2"]
            s = off * dist
            out[..., 0] = sample_linear(acc, u + s, v + s)[..., 0] * inv
            out[..., 1] = sample_linear(acc, u, v)[..., 1] * inv
            out[..., 2] = sample_linear(acc, u - s, v - s)[..., 2] * inv
        else:
            s = off * 0.025 * pp["ca_p2"]
            out[..., 0] = sample_linear(acc, u + s, v + s)[..., 0] * inv
            out[..., 1] = sample_linear(acc, u, v)[..., 1] * inv
            out[..., 2] = sample_linear(acc, u - s, v - s)[..., 2] * inv
    else:
        out = acc * inv
    out = tonemap(out, int(pp.get("tonemap_index", 0)))
    if pp.get("use_vignette"):
        dist = 1.0 - np.power(np.hypot(u - 0.5, v - 0.5), pp["vignette_power"]) * pp["vignette_intensity"]
        out = out * dist[..., None]
    return out


def to_rgb8(c):
    """GL's float -> normalised unsigned byte conversion on read-back: round to
    nearest (OpenGL 4.6 spec, 2.3.5.2), ties upward (the spec leaves them open)."""
    return np.floor(np.clip(np.asarray(c, np.float64), 0.0, 1.0) * 255.0 + 0.5).astype(np.uint8)
