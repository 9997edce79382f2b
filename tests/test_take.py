"""sprout's scene formats (sprout_amd.take): the mesh, scene and take files
of tests/golden/take_fixture, written in sprout's formats, and the
reference's own sample mesh when the reference is here."""

import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "take_fixture")
SAMPLE = "/root/reference/doc/samples/mis"


def test_mesh_parts():
    from sprout_amd import take

    pos, idx, part = take.load_mesh(os.path.join(FIX, "plane.json"))
    assert pos.shape == (5, 3) and idx.shape == (3, 3)
    assert part.tolist() == [0, 0, 1]


def test_rotation_convention():
    from sprout_amd import take

    r = take.rotation_matrix([30, 40, 50])
    assert np.allclose(r @ r.T, np.eye(3)) and np.isclose(np.linalg.det(r), 1)
    # Rz Rx Ry with sprout's element layout: a pure x rotation by 90 degrees
    # takes the object's +z row to +y... as set_rotation_x writes it
    rx = take.rotation_matrix([90, 0, 0])
    assert np.allclose(rx, [[1, 0, 0], [0, 0, -1], [0, 1, 0]])
    m = take.object_to_world({"position": [1, 2, 3], "scale": [2, 3, 4], "rotation": [0, 0, 0]})
    assert np.allclose(np.array([1, 1, 1, 1.0]) @ m, [3, 5, 7, 1])


def test_take_fixture():
    from sprout_amd import scenes, take

    t = take.Take(os.path.join(FIX, "small.take"))
    pos, idx, tmat, mats = t.scene
    assert len(idx) == 5 and len(mats) == 3
    assert mats["kind"].tolist() == [scenes.SUBSTITUTE, scenes.SUBSTITUTE, scenes.LIGHT]
    assert tmat.tolist() == [0, 0, 1, 2, 2]
    # the floor is 4 x 4
    a, b, c = (pos[idx[:2, k]] for k in range(3))
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1).sum()
    assert np.isclose(area, 16)
    # the lamp faces down at y = 2.5
    la, lb, lc = (pos[idx[3, k]] for k in range(3))
    n = np.cross(lb - la, lc - la)
    assert n[1] < 0 and np.allclose(pos[idx[3:]][..., 1], 2.5)
    assert (t.width, t.height, t.samples_per_pixel) == (64, 48, 16)
    assert (t.min_bounces, t.max_bounces) == (3, 6)
    assert t.filter == "Gaussian"
    assert t.sampler == "Random" and not t.golden_ratio
    # sprout's camera: the entity's position and rotation rows, horizontal fov
    assert t.camera["fov_deg"] == 70
    assert np.allclose(t.camera["position"], [0, 1.2, -4])
    r = np.asarray(t.camera["rotation"]).reshape(3, 3)
    assert np.allclose(r @ r.T, np.eye(3), atol=1e-6)
    assert np.allclose(r[0], [1, 0, 0]) and r[2][1] < 0 < r[2][2]  # looks down toward +z
    # colours mapped sRGB -> ACEScg as sprout's default build loads them
    assert np.array_equal(mats["color"][1], take.srgb_to_ap1([0.9, 0.7, 0.3]))
    assert np.array_equal(mats["emission"][2], take.srgb_to_ap1([8, 7, 6]))


def test_unsupported_parts_are_named(tmp_path):
    import json

    from sprout_amd import take

    scene = {"materials": [{"name": "m", "rendering": {"Light": {"emission": [1, 1, 1]}}}],
             "entities": [{"type": "Light", "shape": {"type": "Sphere"}, "materials": ["m"]}]}
    path = tmp_path / "s.scene"
    path.write_text(json.dumps(scene))
    with pytest.raises(NotImplementedError, match="Sphere"):
        take.load_scene(str(path))
    bb = {"rendering": {"Light": {"emittance": {"quantity": "Radiance",
                                                "spectrum": {"temperature": 5000}}}}}
    with pytest.raises(NotImplementedError, match="temperature"):
        take._material(bb)
    # sprout has no Lambert material
    with pytest.raises(NotImplementedError, match="Lambert"):
        take._material({"rendering": {"Lambert": {"color": [0.5, 0.5, 0.5]}}})
    # Substitute keys this path does not render are named, not dropped
    for extra in ({"thickness": 0.1}, {"coating": {"color": [1, 1, 1]}}, {"two_sided": True},
                  {"metal_preset": "Gold"}, {"emission": [1, 0, 0]}, {"normal": {"file": "n.png"}},
                  {"textures": [{"file": "c.png", "usage": "Color"}]}, {"anisotropy": 0.5}):
        with pytest.raises(NotImplementedError):
            take._material({"rendering": {"Substitute": {"color": [0.5, 0.5, 0.5], **extra}}})
    with pytest.raises(NotImplementedError, match="texture"):
        take._material({"rendering": {"Substitute": {"color": {"file": "c.png"}}}})
    for q in ("Flux", "Intensity", "Exitance", "Luminance"):
        with pytest.raises(NotImplementedError, match=q):
            take._material({"rendering": {"Light": {"emittance": {"quantity": q, "value": 2}}}})


def test_colours_as_sprout_reads_them():
    """read_color's forms (material_provider.cpp:1030-1066) and Provider::
    load_light's emission rules (material_provider.cpp:299-363)."""
    from sprout_amd import take

    # defaults are constants, not mapped: Substitute colour 0.5, light emission 10
    sub = take._material({"rendering": {"Substitute": {}}})
    assert np.all(sub["color"] == np.float32(0.5)) and sub["roughness"] == np.float32(0.8)
    light = take._material({"rendering": {"Light": {}}})
    assert np.all(light["emission"] == 10)
    # radiance = value x mapped colour; no quantity: the emission is used
    rad = take._material({"rendering": {"Light": {"emittance": {
        "quantity": "Radiance", "value": 3, "spectrum": [1, 0.5, 0.25]}}}})
    assert np.array_equal(rad["emission"], np.float32(3) * take.srgb_to_ap1([1, 0.5, 0.25]))
    noq = take._material({"rendering": {"Light": {"emission": [2, 2, 2],
                                                   "emittance": {"value": 9}}}})
    assert np.array_equal(noq["emission"], take.srgb_to_ap1([2, 2, 2]))
    # hex colours and the texture-description form with a value
    assert np.array_equal(take.read_color("#ff8000"),
                          take.srgb_to_ap1([1.0, np.float32(128) / np.float32(255), 0.0]))
    c = take._material({"rendering": {"Substitute": {"color": {"value": [0.2, 0.3, 0.4]}}}})
    assert np.array_equal(c["color"], take.srgb_to_ap1([0.2, 0.3, 0.4]))
    # {"sRGB": ..., "linear": false}: read (mapped), gamma-decoded with the
    # reference's blue-from-green channel, mapped again
    g = take.read_color({"sRGB": [0.5, 0.2, 0.9], "linear": False})
    inner = take.srgb_to_ap1([0.5, 0.2, 0.9])
    dec = [take._gamma_to_linear(inner[0]), take._gamma_to_linear(inner[1]),
           take._gamma_to_linear(inner[1])]
    assert np.array_equal(g, take.srgb_to_ap1(dec))
    assert np.isclose(take._gamma_to_linear(np.float32(0.5)), 0.21404114, rtol=1e-6)


def test_golden_ratio_take(tmp_path):
    import json

    from sprout_amd import take

    doc = json.load(open(os.path.join(FIX, "small.take")))
    doc["sampler"] = {"Golden_ratio": {"samples_per_pixel": 32}}
    doc["scene"] = os.path.join(FIX, "small.scene")
    p = tmp_path / "g.take"
    p.write_text(json.dumps(doc))
    t = take.Take(str(p))
    assert t.golden_ratio and t.samples_per_pixel == 32
    doc["sampler"] = {"RD": {"samples_per_pixel": 4}}
    p.write_text(json.dumps(doc))
    with pytest.raises(NotImplementedError, match="RD"):
        take.Take(str(p))


@pytest.mark.skipif(not os.path.isdir(SAMPLE), reason="the reference is not here")
def test_reference_sample():
    from sprout_amd import take

    pos, idx, part = take.load_mesh(os.path.join(SAMPLE, "quad.json"))
    assert pos.shape == (4, 3) and idx.shape == (2, 3) and part.tolist() == [0, 0]
    # its lights are spheres with black-body spectra: named, not guessed
    with pytest.raises(NotImplementedError):
        take.Take(os.path.join(SAMPLE, "mis.take"))


def test_binary_mesh_written_by_the_reference():
    """tests/golden/mesh_parts.sub was written by the reference's own
    Exporter::write from mesh_parts.json (tests/golden/make_mesh_sub.py):
    both load to the same positions, triangles and parts."""
    from sprout_amd import take

    sub = take.load_mesh(os.path.join(HERE, "golden", "mesh_parts.sub"))
    src = take.load_mesh(os.path.join(HERE, "golden", "mesh_parts.json"))
    for a, b in zip(sub, src):
        assert a.dtype == b.dtype and np.array_equal(a, b)
    assert sub[2].tolist() == [1, 1, 1, 1, 0, 0, 0]


@pytest.mark.parametrize("case", ["int16_delta", "uint16", "int32_delta"])
def test_binary_mesh_index_encodings(tmp_path, case):
    """Each index encoding the exporter chooses (Int16 deltas, UInt16 when a
    delta needs 16 bits but every index fits, Int32 deltas above 65535),
    written by the reference's exporter at test time and read back."""
    import sys

    from oracle import oracle as O
    from sprout_amd import take

    lib = O.ref_lib_path()
    if not os.path.exists(lib):
        pytest.skip("oracle/_ref not built (the reference is absent)")
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import make_mesh_sub as M

    n = {"int16_delta": 40, "uint16": 40000, "int32_delta": 70000}[case]
    rng = np.random.default_rng(n)
    pos = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    idx = rng.integers(0, 40, (30, 3))
    if case != "int16_delta":
        idx[3] = [0, n - 1, 1]  # a delta of n - 1
    parts = [(0, 45, 2), (45, 45, 0)]
    path = M.export(lib, M.mesh_json(pos, idx, parts), str(tmp_path), "m")
    with open(path, "rb") as f:
        data = f.read()
    head = data[12:12 + int(np.frombuffer(data, np.uint64, 1, 4)[0])].decode()
    want = {"int16_delta": '"Int16"', "uint16": '"UInt16"', "int32_delta": '"Int32"'}[case]
    assert want in head
    p, i, part = take.load_mesh(path)
    assert np.array_equal(p, pos) and np.array_equal(i, idx.astype(np.uint32))
    assert part.tolist() == [2] * 15 + [0] * 15


def test_binary_mesh_rejects_bad_files(tmp_path):
    from sprout_amd import take

    src = open(os.path.join(HERE, "golden", "mesh_parts.sub"), "rb").read()
    bad = tmp_path / "short.sub"
    bad.write_bytes(src[:-4])  # index data cut short
    with pytest.raises(ValueError):
        take.load_mesh(str(bad))


def test_vertex_normals_as_sprout_gives_them():
    """a JSON mesh without normals: each triangle writes its geometric normal
    to its vertices in order, the last one stays (the floor's two back
    corners take the metallic triangle's normal); explicit normals are read
    and rotated without the scale"""
    from sprout_amd import take

    pos, idx, part, nrm = take.load_mesh(os.path.join(FIX, "plane.json"), with_normals=True)
    up = np.array([0, 1, 0], np.float32)
    assert np.allclose(nrm[[2, 3]], up)
    a, b, c = pos[idx[2]]
    g = np.cross(b - a, c - a)
    g /= np.linalg.norm(g)
    assert np.allclose(nrm[[0, 1, 4]], g, atol=1e-6)
    t = take.Take(os.path.join(HERE, "golden", "take_smooth", "smooth.take"))
    pos, idx, tmat, mats = t.scene
    assert t.normals.shape == pos.shape
    assert np.allclose(np.linalg.norm(t.normals, axis=1), 1, atol=1e-5)
    # the sphere (centre 0, 0.7, 0.4, radius 0.7): normals point away from the centre
    sphere = np.isin(np.arange(len(pos)), idx[(tmat == 1) | (tmat == 2)])
    radial = pos[sphere] - np.array([0, 0.7, 0.4], np.float32)
    radial /= np.linalg.norm(radial, axis=1, keepdims=True)
    assert np.allclose(t.normals[sphere], radial, atol=1e-4)


def test_non_uniform_scale_that_tilts_normals_is_refused(tmp_path):
    """sprout rotates normals without the scale (triangle_mesh.cpp:388-421):
    a tilted triangle under a non-uniform scale would shade differently, so
    it is refused; the lamp quad's scale (0.5, 0.5, 1) keeps its normal"""
    import json
    import shutil

    from sprout_amd import take

    for f in os.listdir(FIX):
        shutil.copy(os.path.join(FIX, f), tmp_path / f)
    doc = json.loads((tmp_path / "small.scene").read_text())
    doc["entities"][0]["transformation"]["scale"] = [2, 1, 2]
    (tmp_path / "small.scene").write_text(json.dumps(doc))
    with pytest.raises(NotImplementedError, match="non-uniform scale"):
        take.load_scene(str(tmp_path / "small.scene"))
    doc["entities"][0]["transformation"]["scale"] = [2, 2, 2]
    (tmp_path / "small.scene").write_text(json.dumps(doc))
    take.load_scene(str(tmp_path / "small.scene"))
