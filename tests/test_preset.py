"""Scene ingest: OBJ / MTL models and .NPRTENV presets (the reference's own
experiment scenes, copied under tests/golden/presets by make_preset_fixtures.py).

The material order (a default material, then the .mtl's entries in file
order) is pinned by the presets the reference application wrote: their
`materials` tables list the application's materials in its order, and the
colours the .mtl defines appear at exactly the indices this order gives."""
import math
import os

import numpy as np
import pytest

from cudapt_amd import preset

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "presets")


def _read(name):
    with open(os.path.join(GOLD, name)) as f:
        return f.read()


def _load(name):
    return preset.load_preset(os.path.join(GOLD, name), model_dirs=[GOLD])


def test_tokenizer_drops_comments():
    t = preset._tokens("// header line 1 2\ncampos 1 2 3 // trailing\n/* block\n rttech 9 */ rttech 5\n")
    assert t == ["campos", "1", "2", "3", "rttech", "5"]
    p = preset.parse_preset("campos 1 2 3\ncampitchyaw 0.5 -0.25\nrttech 5\nunknown 7 8\n")
    assert p["campos"] == (1.0, 2.0, 3.0) and p["campitchyaw"] == (0.5, -0.25) and p["rttech"] == 5


def test_fly_camera_closed_forms():
    c = preset.fly_camera((0, 0, 0), 0.0, 0.0)
    assert np.allclose(c["dir"], (0, 0, 1)) and c["up"] == (0.0, 1.0, 0.0) and c["fov"] == pytest.approx(math.pi / 4)
    assert np.allclose(preset.fly_camera((0, 0, 0), math.pi / 2, 0.0)["dir"], (1, 0, 0), atol=1e-7)
    assert np.allclose(preset.fly_camera((0, 0, 0), 0.0, math.pi / 2)["dir"], (0, 1, 0), atol=1e-7)
    assert np.allclose(preset.fly_camera((0, 0, 0), 0.0, -math.pi / 2)["dir"], (0, -1, 0), atol=1e-7)


def test_cornell_preset_header():
    p = _load("cornellwtranball.NPRTENV")
    assert p["rttech"] == 3 and p["reference_mode"] == "cudaRTPTStream" and p["mode"] == 4
    assert p["camera"]["pos"] == pytest.approx((-1.82495, -2.40754, -2.74389))
    pitch, yaw = -0.13, -0.215
    want = np.array([math.cos(yaw) * math.sin(pitch), math.sin(yaw), math.cos(yaw) * math.cos(pitch)])
    assert np.allclose(p["camera"]["dir"], want / np.linalg.norm(want), atol=1e-6)


@pytest.mark.parametrize("name", ["cornellwtranball.NPRTENV", "prismA.NPRTENV"])
def test_material_table_order_matches_preset(name):
    """The table built from the OBJ/MTL has as many entries as the preset's
    `materials` record and the default material (diffuse 0.6) at index 0."""
    pre = preset.parse_preset(_read(name))
    model = os.path.basename(pre["models"][0].replace("\\", "/"))
    m = preset.load_model(os.path.join(GOLD, model))
    assert len(m["mats"]) == len(pre["materials"])
    assert np.allclose(m["mats"][0]["diffuse"], 0.6)
    assert np.allclose(pre["materials"][0][3:6], 0.6)


def test_cornell_mtl_colours_at_preset_indices():
    pre = preset.parse_preset(_read("cornellwtranball.NPRTENV"))
    m = preset.load_model(os.path.join(GOLD, "cornell_box_with_sphere.obj"))
    for k, rec in enumerate(pre["materials"]):
        assert np.allclose(m["mats"][k]["diffuse"], rec[3:6]), k          # diffuse: values 3..5
        if k > 0:
            assert np.allclose(m["mats"][k]["emissive"], rec[7:10]), k    # emissive (Ka + Ke): values 7..9
    # the .mtl's names in file order after the default material
    names = [n for n, _ in preset.read_mtl(_read("cornell_box_with_sphere.mtl"))]
    assert names == ["green", "light", "red", "test_sphere_initialShadingGroup", "white"]


def _faces_per_material(obj_text, mtl_names):
    """Triangles per material index, counted from the face lines directly."""
    idx = {n: i + 1 for i, n in enumerate(mtl_names)}
    counts = np.zeros(len(mtl_names) + 1, np.int64)
    cur = 0
    for line in obj_text.splitlines():
        if line.startswith("usemtl"):
            cur = idx[line.split(None, 1)[1].strip()]
        elif line.startswith("f "):
            counts[cur] += len(line.split()) - 3
    return counts


@pytest.mark.parametrize("model", ["cornell_box_with_sphere", "test_prism"])
def test_face_materials_and_positions(model):
    text, mtl = _read(model + ".obj"), _read(model + ".mtl")
    m = preset.read_obj_model(text, mtl)
    names = [n for n, _ in preset.read_mtl(mtl)]
    assert np.array_equal(np.bincount(m["mat"], minlength=len(names) + 1), _faces_per_material(text, names))
    # corners: the same positions as the plain mesh fixture of the same file
    gold = np.load(os.path.join(HERE, "golden", "meshes.npz"))
    if model + "_pos" in gold:
        p, i = gold[model + "_pos"], gold[model + "_idx"]
        assert np.array_equal(m["pos"][m["idx"]], p[i])
    assert len(m["pos"]) == 3 * len(m["idx"]) == len(m["nrm"]) == len(m["uv"])


def test_normals_uvs_and_fan():
    text = "\n".join(["v 0 0 0", "v 1 0 0", "v 1 1 0", "v 0 1 0", "vt 0.25 0.75", "vn 0 0 -1",
                      "f 1/1/1 2/1/1 3/1/1 4/1/1", "f 1 3 2"])
    m = preset.read_obj_model(text)
    assert len(m["idx"]) == 3 and np.all(m["mat"] == 0)
    # fan (0, k, k + 1) of the quad
    assert np.array_equal(m["pos"][:6], np.float32([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 0, 0], [1, 1, 0], [0, 1, 0]]))
    assert np.allclose(m["nrm"][:6], (0, 0, -1))                 # the file's vn
    assert np.allclose(m["uv"][:6], (0.25, 0.25))                # v flipped
    assert np.allclose(m["nrm"][6:], (0, 0, -1))                 # flat normal of 1-3-2 (clockwise seen from +z)
    assert np.allclose(m["mats"][0]["diffuse"], 0.6)


def test_preset_materials_applied_only_when_table_is_large_enough():
    rec = [0.0] * preset.PRESET_MATERIAL_VALUES
    rec[3:6] = [0.1, 0.2, 0.3]
    rec[13] = 0.9                                                   # roughness
    mats = np.asarray([preset.rt_material()], preset.MATERIAL_DTYPE)
    same = preset.apply_preset_materials(mats, [rec, rec])
    assert np.array_equal(same, mats)
    new = preset.apply_preset_materials(mats, [rec])
    assert np.allclose(new[0]["diffuse"], (0.1, 0.2, 0.3)) and new[0]["roughness"] == pytest.approx(0.9)


def test_prism_preset_skips_spectral_block():
    p = _load("prismA.NPRTENV")
    mats = p["scene"]["mats"]
    assert len(mats) == 6 and mats[4]["emissive"][0] == 500.0 and mats[5]["transparency"] == 1.0
    assert p["rttech"] == 3 and p["converged"] is None
    assert len(p["scene"]["idx"]) == int(_faces_per_material(_read("test_prism.obj"),
                                                             [n for n, _ in preset.read_mtl(_read("test_prism.mtl"))]).sum())


def test_missing_model_is_an_error(tmp_path):
    f = tmp_path / "x.NPRTENV"
    f.write_text("models 1 ..\\asset\\model\\nowhere.obj\n")
    with pytest.raises(FileNotFoundError):
        preset.load_preset(str(f))
