"""Input checks of the device BVH build, on the CPU: lfh_create rejects a
scene desc without BVH nodes whose mesh / instance tables are inconsistent
before it touches the GPU, and lfh_scene_flatten fills those tables."""
import numpy as np
import pytest

from lavaframe_amd import lfh, scenes


def _flat_arrays():
    sc = scenes.cornell(16, 16, build=False)
    arr = sc.arrays()
    assert arr["nodes"].size == 0
    return sc, arr


def test_flatten_fills_mesh_and_instance_tables():
    sc, arr = _flat_arrays()
    off = arr["mesh_tri_offsets"]
    assert off[0] == 0 and off[-1] == len(arr["vert_indices"]) and (np.diff(off) > 0).all()
    assert list(arr["instance_meshes"]) == [0, 1, 2, 3, 3]
    assert list(arr["instance_materials"]) == [1, 2, 3, 1, 1]
    # triangles in mesh order: vertex indices 3k, 3k+1, 3k+2
    vi = arr["vert_indices"]
    assert np.array_equal(vi[:, 0], np.arange(len(vi)) * 3)
    built = scenes.cornell(16, 16).arrays()
    assert np.array_equal(built["mesh_tri_offsets"], off)


@pytest.mark.parametrize("edit,msg", [
    (lambda a: a.__setitem__("mesh_tri_offsets", a["mesh_tri_offsets"][:-1]), "offsets"),
    (lambda a: a["mesh_tri_offsets"].__setitem__(2, a["mesh_tri_offsets"][1]), "empty mesh"),
    (lambda a: a["instance_meshes"].__setitem__(4, 9), "instance mesh out of range"),
    (lambda a: a["vertices"].__setitem__((5, 1), np.nan), "non-finite vertex"),
])
def test_inconsistent_tables_are_rejected(edit, msg):
    sc, arr = _flat_arrays()
    arr = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in arr.items()}
    edit(arr)
    desc, keep = lfh.desc_from_arrays(arr)
    with pytest.raises(lfh.LfhError) as e:
        lfh.Context(desc, sc.options())
    assert msg in str(e.value)
