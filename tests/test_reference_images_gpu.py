"""Rendered frames against the FP64 oracle's 65,536-spp references written by
tests/golden/make_fp64_references.py (independent random numbers: frames from
1,000,001 on): the Cornell box at 256 x 256 (every pixel) and the instanced
scene at 1920 x 1080 (1024 pixels, one per cell of a 32 x 32 grid).

The bar is statistical: per pixel and channel the difference over its
standard error (the reference's per-sample deviation, scaled to both sample
counts) stays within 4.5 for 99.5% of the values (99% on the 3072 values of
the instanced scene, whose sky and lights give heavier tails), block means and
the frame mean show no bias, and the RMSE equals what the two noises predict
within 10% (on the instanced scene: at most 25% above it)."""
from pathlib import Path

import numpy as np
import pytest

from lavaframe_amd import lfh, scenes

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def _reference(name):
    """(mean, std, spp, pixels or None) of a stored reference."""
    f = GOLD / name
    if not f.exists():
        pytest.skip(f"{name} not generated (tests/golden/make_fp64_references.py)")
    z = np.load(f)
    px = z["pixels"] if "pixels" in z.files else None
    return (z["mean"].reshape(-1, 3).astype(np.float64), z["std"].reshape(-1, 3).astype(np.float64), int(z["spp"]),
            px)


def _frame(sc, spp):
    ctx = lfh.Context.from_scene(sc)
    ctx.update(sc.camera(), reset=True)
    ctx.render_passes(spp)
    s, c = ctx.accum()
    o = sc.options()
    return ((s.astype(np.float64) - c) / spp)[..., :3].reshape(o.height * o.width, 3)


def _check(img, mean, std, spp, ref_spp, frac=0.995, ratio=(0.9, 1.1)):
    se = std * np.sqrt(1.0 / spp + 1.0 / ref_spp)
    z = np.abs(img - mean) / np.maximum(se, 1e-7)
    assert np.mean(z < 4.5) > frac, np.mean(z < 4.5)
    rmse = np.sqrt(np.mean((img - mean) ** 2))
    expected = np.sqrt(np.mean(se ** 2))
    assert ratio[0] < rmse / expected < ratio[1], (rmse, expected)
    se_all = np.sqrt(np.sum(se ** 2)) / img.size
    assert abs(img.mean() - mean.mean()) < 4 * se_all, (img.mean(), mean.mean(), se_all)
    return rmse, expected


def test_cornell_frame_matches_fp64_64k_reference():
    print("[ref] cornell 256^2 x 4096 spp", flush=True)
    mean, std, ref_spp, _ = _reference("cornell_c0_fp64_64k.npz")
    spp = 4096
    img = _frame(scenes.cornell(256, 256, max_depth=4), spp)
    _check(img, mean, std, spp, ref_spp)
    # 8 x 8-pixel block means: no block biased
    blk = lambda a: a.reshape(32, 8, 32, 8, 3).mean((1, 3))  # noqa: E731
    se_b = np.sqrt(blk(std.reshape(256, 256, 3) ** 2) / 64 * (1.0 / spp + 1.0 / ref_spp))
    zb = np.abs(blk(img.reshape(256, 256, 3)) - blk(mean.reshape(256, 256, 3))) / np.maximum(se_b, 1e-7)
    assert np.mean(zb < 4) > 0.99 and zb.max() < 6, (np.mean(zb < 4), zb.max())


def test_instanced_frame_matches_fp64_64k_reference_on_1024_pixels():
    print("[ref] instanced 1080p x 256 spp", flush=True)
    mean, std, ref_spp, px = _reference("instanced_c2_fp64_64k.npz")
    spp = 256
    img = _frame(scenes.instanced(), spp)[px]
    # rare bright samples (sun lobe, lights seen after glossy bounces) carry
    # most of the per-sample variance, so a typical frame falls below the
    # predicted RMSE (0.74 of it measured at 256 spp); only an excess would
    # mean bias
    _check(img, mean, std, spp, ref_spp, frac=0.99, ratio=(0.5, 1.25))
