"""The drop-in adapter of include/lavaframe/HipTiledRenderer.h compiles
against the reference's own headers (LavaFrame::Renderer, Renderer.h:71-117;
LavaFrame::Scene, Scene.h:55-101): oracle/Makefile builds
oracle/_ref/adapter_check.o, which instantiates lfh::HipRendererT over
LavaFrame::Scene and derives HipTiledRenderer from LavaFrame::Renderer with
every virtual overridden (`override` on each).  A CPU test; it needs the
reference sources, so it is skipped where they are absent."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference/LavaFrame")
OBJ = ROOT / "oracle" / "_ref" / "adapter_check.o"


@pytest.mark.skipif(not REF.exists(), reason="reference sources not present")
def test_adapter_compiles_against_the_reference_headers():
    r = subprocess.run(["make", "-f", "Makefile", "ref"], cwd=ROOT / "oracle", capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert OBJ.exists()
    syms = subprocess.run(["nm", "-C", str(OBJ)], capture_output=True, text=True, check=True).stdout
    # the class's vtable and every override were emitted from the reference's declarations
    for s in ("vtable for LavaFrame::HipTiledRenderer", "lfh_adapter_make",
              "LavaFrame::HipTiledRenderer::Init()", "LavaFrame::HipTiledRenderer::Update(float)",
              "LavaFrame::HipTiledRenderer::GetOutputBufferHDR(float**, int&, int&)",
              "lfh::HipRendererT<LavaFrame::Scene>::Init()"):
        assert s in syms, s
    # the Renderer base it derives from is the reference's (its constructor is referenced)
    assert "LavaFrame::Renderer::Renderer(LavaFrame::Scene*" in syms
