"""The headless renderer (lavaframe_amd/_lib/lfh_render) on a .lfs scene
written here: it loads the file with the reference grammar (Loader.cpp:45-386),
renders maxSamples passes through the Renderer surface (Update, then Render,
Main.cpp:747-753) and exports accum / maxSamples as PFM; the image equals the
library's render of the same file through the C ABI bit for bit."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from lavaframe_amd import lfh, scenes

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "lavaframe_amd" / "_lib" / "lfh_render"

OBJ_QUAD = """v -2 0 -2
v 2 0 -2
v 2 0 2
v -2 0 2
vn 0 1 0
f 1//1 4//1 3//1 2//1
"""

LFS = """Renderer
{
	resolution 64 48
	maxDepth 3
	tonemapIndex %d
	useVignette %d
	vignetteIntensity 0.6
	vignettePower 1.5
}

Camera
{
	position 0 1.2 -3.5
	lookAt 0 0.4 0
	fov 45
}

material floor
{
	color 0.7 0.7 0.7
}

material red
{
	color 0.7 0.1 0.1
	roughness 0.3
}

light
{
	position -0.5 2.5 -0.5
	v1 0.5 2.5 -0.5
	v2 -0.5 2.5 0.5
	emission 8 8 8
	type Quad
}

mesh
{
	file floor.obj
	material floor
}

mesh
{
	file floor.obj
	material red
	position 0 0.6 0
	scale 0.2 1 0.2
}
"""


def _render_like_main(scene, ms):
    sc = lfh.Scene().load(scene).build()
    ctx = lfh.Context.from_scene(sc)
    cam = sc.camera()
    while True:
        ctx.update(cam)
        if ctx.sample_count() == ms + 1:
            break
        ctx.render()
    return sc, ctx


def test_cli_renders_scene_file_like_the_library(tmp_path):
    print("[cli] lfh_render", flush=True)
    (tmp_path / "floor.obj").write_text(OBJ_QUAD)
    scene = tmp_path / "scene.lfs"
    scene.write_text(LFS % (0, 0))
    ms = 8
    r = subprocess.run([str(CLI), "-s", str(scene), "-ms", str(ms), "-ou", str(tmp_path), "-et", "pfm"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    data = (tmp_path / "scene.pfm").read_bytes()
    header = b"PF\n64 48\n-1.0\n"
    assert data.startswith(header)
    img = np.frombuffer(data[len(header):], np.float32).reshape(48, 64, 3)
    assert np.isfinite(img).all() and img.mean() > 0
    # the same file through the C ABI and the same Update / Render sequence:
    # tonemap 0 exports accum / maxSamples
    sc, ctx = _render_like_main(scene, ms)
    s, c = ctx.accum()
    ref = (s - c) * np.float32(1.0 / ms)
    np.testing.assert_array_equal(img, ref)


def test_cli_ppm_export_is_the_post_processed_rgb8(tmp_path):
    """-et ppm writes GetOutputBuffer's image: ACES tonemap and vignette from
    the Renderer block, RGB8, rows top to bottom."""
    print("[cli] lfh_render ppm", flush=True)
    (tmp_path / "floor.obj").write_text(OBJ_QUAD)
    scene = tmp_path / "scene.lfs"
    scene.write_text(LFS % (2, 1))
    ms = 6
    r = subprocess.run([str(CLI), "-s", str(scene), "-ms", str(ms), "-ou", str(tmp_path), "-et", "ppm"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    data = (tmp_path / "scene.ppm").read_bytes()
    header = b"P6\n64 48\n255\n"
    assert data.startswith(header)
    img = np.frombuffer(data[len(header):], np.uint8).reshape(48, 64, 3)[::-1]
    sc, ctx = _render_like_main(scene, ms)
    pp = sc.postprocess()
    assert (pp.tonemap_index, pp.use_vignette) == (2, 1)
    np.testing.assert_array_equal(img, ctx.output_ldr(pp, 1.0 / ms))
    assert img.mean() > 5


def test_cli_device_bvh_renders_the_same_image(tmp_path):
    """-dbvh builds the BVHs on the GPU: the exported image is bitwise the
    one rendered over the scene's own BVH."""
    print("[cli] lfh_render -dbvh", flush=True)
    (tmp_path / "floor.obj").write_text(OBJ_QUAD)
    scene = tmp_path / "scene.lfs"
    scene.write_text(LFS % (0, 0))
    imgs = []
    for extra in ([], ["-dbvh"]):
        out = tmp_path / ("d" if extra else "h")
        out.mkdir()
        r = subprocess.run([str(CLI), "-s", str(scene), "-ms", "6", "-ou", str(out), *extra],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        imgs.append((out / "scene.pfm").read_bytes())
    assert imgs[0] == imgs[1]


def test_cli_textured_scene_like_the_library(tmp_path):
    """A floor with a PNG albedo map (OBJ texture coordinates, loader key
    albedoTexture): the CLI's image equals the library's over the same file."""
    from PIL import Image
    print("[cli] textured scene", flush=True)
    Image.fromarray(scenes.texture_maps(16)[0], "RGBA").save(tmp_path / "chk.png")
    (tmp_path / "floor.obj").write_text("v -2 0 -2\nv 2 0 -2\nv 2 0 2\nv -2 0 2\nvt 0 0\nvt 3 0\nvt 3 3\nvt 0 3\n"
                                        "vn 0 1 0\nf 1/1/1 4/4/1 3/3/1 2/2/1\n")
    scene = tmp_path / "scene.lfs"
    text = (LFS % (0, 0)).replace("material floor\n{\n\tcolor 0.7 0.7 0.7\n}",
                                  "material floor\n{\n\tcolor 0.9 0.9 0.9\n\talbedoTexture chk.png\n}")
    assert "albedoTexture" in text
    scene.write_text(text)
    ms = 6
    r = subprocess.run([str(CLI), "-s", str(scene), "-ms", str(ms), "-ou", str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    data = (tmp_path / "scene.pfm").read_bytes()
    header = b"PF\n64 48\n-1.0\n"
    img = np.frombuffer(data[len(header):], np.float32).reshape(48, 64, 3)
    sc, ctx = _render_like_main(scene, ms)
    assert sc.arrays()["textures"].shape[0] == 1
    s, c = ctx.accum()
    np.testing.assert_array_equal(img, (s - c) * np.float32(1.0 / ms))
