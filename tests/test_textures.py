"""Texture maps on the CPU: the image decoders (PNG, binary PPM/PGM) against
PIL's decoding of the same files, the loader's texture keys
(Loader.cpp:127-147), the one-size rule of the array texture, and the
oracle's texture() read pinned by closed forms (texel centres, midpoints,
GL_REPEAT wrapping, layer clamping)."""
import io

import numpy as np
import pytest
from PIL import Image

from lavaframe_amd import lfh, scenes
from oracle import native


def _decode(tmp_path, name, data):
    f = tmp_path / name
    f.write_bytes(data)
    s = lfh.Scene()
    i = s.add_texture_file(f)
    s.add_material(lfh.MATERIAL_DEFAULT)
    s.add_instance(s.add_mesh(scenes.soup_triangles(1)), scenes.affine(), 0)
    s.flatten()
    return s.arrays()["textures"][i]


def _png(img, **kw):
    b = io.BytesIO()
    img.save(b, format="PNG", **kw)
    return b.getvalue()


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "LA", "P"])
@pytest.mark.parametrize("level", [0, 1, 9])
def test_png_decodes_like_pil(tmp_path, mode, level):
    rng = np.random.default_rng(len(mode) * 10 + level)
    base = rng.integers(0, 256, (23, 37, 4), dtype=np.uint8)
    base[5:12, 3:30] = (40, 200, 90, 255)          # runs compress differently from noise
    img = Image.fromarray(base, "RGBA").convert(mode) if mode != "P" else \
        Image.fromarray(base[..., :3], "RGB").quantize(colors=50)
    got = _decode(tmp_path, f"t_{mode}_{level}.png", _png(img, compress_level=level))
    want = np.asarray(img.convert("RGBA"))
    assert got.shape == want.shape
    assert np.array_equal(got, want)


def test_png_16_bit_and_1_bit(tmp_path):
    """16-bit gray keeps the high byte; 1-bit gray scales 1 to 255 (stb_image's
    conversions)."""
    g16 = (np.arange(16 * 9, dtype=np.uint32).reshape(9, 16) * 457).astype(np.uint16)
    got = _decode(tmp_path, "g16.png", _png(Image.fromarray(g16)))
    assert np.array_equal(got[..., 0], (g16 >> 8).astype(np.uint8)) and (got[..., 3] == 255).all()
    bits = (np.random.default_rng(1).random((7, 13)) > 0.5)
    got = _decode(tmp_path, "b1.png", _png(Image.fromarray(bits)))
    assert np.array_equal(got[..., 1], bits.astype(np.uint8) * 255)


def test_binary_ppm_and_pgm(tmp_path):
    rgb = np.random.default_rng(2).integers(0, 256, (5, 6, 3), dtype=np.uint8)
    got = _decode(tmp_path, "a.ppm", b"P6\n# comment\n6 5\n255\n" + rgb.tobytes())
    assert np.array_equal(got[..., :3], rgb) and (got[..., 3] == 255).all()
    g = np.arange(12, dtype=np.uint8).reshape(3, 4) * 20
    got = _decode(tmp_path, "a.pgm", b"P5 4 3 255\n" + g.tobytes())
    assert np.array_equal(got[..., 2], g)


def test_unreadable_textures_are_reported(tmp_path):
    (tmp_path / "x.jpg").write_bytes(b"\xff\xd8\xff\xe0 not decoded here")
    s = lfh.Scene()
    with pytest.raises(lfh.LfhError):
        s.add_texture_file(tmp_path / "x.jpg")
    with pytest.raises(lfh.LfhError):
        s.add_texture_file(tmp_path / "missing.png")


def test_loader_reads_texture_keys(tmp_path):
    """albedoTexture / metallicRoughnessTexture / normalTexture /
    emissionTexture name files relative to the scene file; the same file twice
    is one texture; a file that cannot be read leaves the id at -1."""
    a, m, n, e = scenes.texture_maps(8)
    for name, img in (("a.png", a), ("m.png", m), ("n.png", n), ("e.png", e)):
        Image.fromarray(img, "RGBA").save(tmp_path / name)
    (tmp_path / "q.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nvn 0 0 1\n"
                                    "f 1/1/1 2/2/1 3/3/1\n")
    (tmp_path / "s.lfs").write_text(
        "material one\n{\n\tcolor 1 1 1\n\talbedoTexture a.png\n\tmetallicRoughnessTexture m.png\n"
        "\tnormalTexture n.png\n\temissionTexture e.png\n}\n\n"
        "material two\n{\n\tcolor 1 1 1\n\talbedoTexture a.png\n\tnormalTexture missing.png\n}\n\n"
        "Camera\n{\n\tposition 0 0 -3\n\tlookAt 0 0 0\n\tfov 40\n}\n\n"
        "mesh\n{\n\tfile q.obj\n\tmaterial one\n}\n")
    s = lfh.Scene().load(tmp_path / "s.lfs")
    arr = s.arrays()
    assert arr["textures"].shape == (4, 8, 8, 4)
    assert np.array_equal(arr["textures"][2], n)
    mats = arr["materials"]
    assert list(mats[1, 24:28]) == [0, 1, 2, 3]
    assert list(mats[2, 24:28]) == [0, -1, -1, -1]


def test_textures_of_different_sizes_are_rejected():
    s = lfh.Scene()
    s.add_texture(np.zeros((4, 4, 4), np.uint8), "a")
    s.add_texture(np.zeros((8, 4, 4), np.uint8), "b")
    s.add_material(lfh.MATERIAL_DEFAULT)
    s.add_instance(s.add_mesh(scenes.soup_triangles(1)), scenes.affine(), 0)
    with pytest.raises(lfh.LfhError) as e:
        s.build()
    assert "one size" in str(e.value)


def _orc_with(tex):
    s = lfh.Scene()
    for i, t in enumerate(tex):
        s.add_texture(t, f"t{i}")
    s.add_material(lfh.MATERIAL_DEFAULT)
    s.add_instance(s.add_mesh(scenes.soup_triangles(1)), scenes.affine(), 0)
    s.build()
    return native.OracleScene(s.arrays())


def test_oracle_texture_read_closed_forms():
    rng = np.random.default_rng(5)
    W, H = 6, 4
    t0 = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    t1 = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    orc = _orc_with([t0, t1])
    # texel centres return the texel (row j at t = (j + 0.5) / H)
    jj, ii = np.mgrid[0:H, 0:W]
    st = np.stack([(ii + 0.5) / W, (jj + 0.5) / H], -1).reshape(-1, 2)
    np.testing.assert_allclose(orc.texture_sample(st, 0), t0.reshape(-1, 4) / 255.0, atol=1e-12)
    np.testing.assert_allclose(orc.texture_sample(st, 1), t1.reshape(-1, 4) / 255.0, atol=1e-12)
    # halfway between two horizontal neighbours: their mean
    st2 = np.array([[(2 + 1.0) / W, (1 + 0.5) / H]])
    np.testing.assert_allclose(orc.texture_sample(st2, 0)[0], (t0[1, 2] / 255.0 + t0[1, 3] / 255.0) / 2, atol=1e-12)
    # GL_REPEAT: whole periods, negative coordinates included; the left edge
    # blends the last and first columns
    sh = st + rng.integers(-3, 4, st.shape)
    np.testing.assert_allclose(orc.texture_sample(sh, 0), orc.texture_sample(st, 0), atol=1e-9)
    edge = orc.texture_sample(np.array([[0.0, 0.5 / H]]), 0)[0]
    np.testing.assert_allclose(edge, (t0[0, 0] / 255.0 + t0[0, W - 1] / 255.0) / 2, atol=1e-12)
    # layers outside the array clamp to the first / last layer
    np.testing.assert_allclose(orc.texture_sample(st, 7), orc.texture_sample(st, 1), atol=0)
    np.testing.assert_allclose(orc.texture_sample(st, -2), orc.texture_sample(st, 0), atol=0)
