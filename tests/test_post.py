"""Pins the post-process oracle (oracle/post.py) with closed forms: the
mirrored-repeat index sequence, bilinear reads at texel centres and midpoints,
tonemap values at chosen inputs, the vignette at the centre and a corner, and
the RGB8 conversion."""
import math

import numpy as np

from oracle import post as P


def test_mirrored_repeat_sequence():
    # a 4-texel texture repeats 0 1 2 3 3 2 1 0 in both directions
    i = np.arange(-8, 16)
    expect = np.array([0, 1, 2, 3, 3, 2, 1, 0] * 3)
    np.testing.assert_array_equal(P.mirror_index(i, 4), expect)


def test_bilinear_reads():
    rng = np.random.default_rng(0)
    tex = rng.random((5, 7, 3))
    H, W = tex.shape[:2]
    ys, xs = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    # texel centres return the texels
    np.testing.assert_allclose(P.sample_linear(tex, (xs + 0.5) / W, (ys + 0.5) / H), tex, atol=1e-12)
    # the midpoint between horizontal neighbours is their average
    np.testing.assert_allclose(P.sample_linear(tex, (xs[:, :-1] + 1.0) / W, (ys[:, :-1] + 0.5) / H),
                               0.5 * (tex[:, :-1] + tex[:, 1:]), atol=1e-12)
    # at the border the mirror repeats the edge texel
    np.testing.assert_allclose(P.sample_linear(tex, np.zeros(H), (np.arange(H) + 0.5) / H), tex[:, 0], atol=1e-12)


def test_tonemap_closed_forms():
    one = np.ones((1, 3))
    zero = np.zeros((1, 3))
    g = 1 / 2.2
    np.testing.assert_allclose(P.tonemap(one * 0.7, 0), one * 0.7)
    np.testing.assert_allclose(P.tonemap(one, 1), one * (2.0 / 3.0) ** g)        # c / (1 + lum / 2), lum = 1
    np.testing.assert_allclose(P.tonemap(zero, 2), zero)
    np.testing.assert_allclose(P.tonemap(one * 1e6, 2), one)                      # ACES saturates at 1
    np.testing.assert_allclose(P.tonemap(one, 3), one * 0.5 ** g)                 # Reinhard c / (c + 1)
    np.testing.assert_allclose(P.tonemap(zero, 4), zero)
    np.testing.assert_allclose(P.tonemap(one * 0.004, 5), zero, atol=1e-15)       # Hejl-Dawson offset
    np.testing.assert_allclose(P.tonemap(zero, 6), zero, atol=1e-15)              # Uncharted 2 fixes 0
    c = np.array([[0.2, 1.5, 4.0]])
    for k in range(1, 7):                                                         # non-decreasing per channel
        t = P.tonemap(c, k)
        assert np.all(np.diff(t[0]) >= 0) and t[0, 0] < t[0, 2], k


def test_vignette_centre_and_corner():
    acc = np.ones((5, 5, 3))
    pp = dict(tonemap_index=0, use_vignette=1, vignette_intensity=0.8, vignette_power=1.5)
    out = P.postprocess(acc, 1.0, pp)
    assert np.allclose(out[2, 2], 1.0)
    d = math.hypot(0.5 - 0.1, 0.5 - 0.1)
    assert np.allclose(out[0, 0], 1.0 - d ** 1.5 * 0.8)


def test_rgb8_conversion():
    c = np.array([-1.0, 0.0, 0.5 / 255, 0.4 / 255, 0.5, 1.0, 7.0])
    np.testing.assert_array_equal(P.to_rgb8(c), [0, 0, 1, 0, 128, 255, 255])


def test_scene_file_sets_post_process_options(tmp_path):
    """The Renderer block's post-process keys (Loader.cpp:262-282) reach
    lfh_scene_get_postprocess; absent keys keep the defaults (Renderer.h:34-43)."""
    from lavaframe_amd import lfh
    (tmp_path / "q.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 3//1\n")
    cam = "Camera\n{\n\tposition 0 0 -3\n\tlookAt 0 0 0\n\tfov 40\n}\n\nmesh\n{\n\tfile q.obj\n}\n"
    (tmp_path / "s.lfs").write_text(
        "Renderer\n{\n\tresolution 8 8\n\ttonemapIndex 4\n\tuseChromaticAberration 1\n\tCADistance 0.2\n"
        "\tuseCAdistortion 1\n\tCAAngularity 3\n\tCADirectionality -1\n\tCACenter 0.4\n\tuseVignette 1\n"
        "\tvignetteIntensity 0.7\n\tvignettePower 2.5\n}\n\n" + cam)
    got = lfh.Scene().load(tmp_path / "s.lfs").postprocess().as_dict()
    want = dict(tonemap_index=4, use_ca=1, use_ca_distortion=1, ca_distance=0.2, ca_p1=3.0, ca_p2=-1.0, ca_p3=0.4,
                use_vignette=1, vignette_intensity=0.7, vignette_power=2.5)
    for k, v in want.items():
        assert abs(got[k] - v) < 1e-6, (k, got[k], v)
    (tmp_path / "d.lfs").write_text("Renderer\n{\n\tresolution 8 8\n}\n\n" + cam)
    got = lfh.Scene().load(tmp_path / "d.lfs").postprocess().as_dict()
    want = dict(tonemap_index=2, use_ca=0, use_ca_distortion=0, ca_distance=0.05, ca_p1=5.0, ca_p2=-0.5, ca_p3=0.5,
                use_vignette=0, vignette_intensity=0.0, vignette_power=1.0)
    for k, v in want.items():
        assert abs(got[k] - v) < 1e-6, (k, got[k], v)
