"""CPU checks of the C-ABI library: it loads, exports every declared entry
point, and its host-side generators and argument checks behave.  No call
here touches a GPU."""

import ctypes
import os
import re

import numpy as np
import pytest

from sprout_amd import _lib, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "sprout_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|char const\*)\s+(mi_\w+)\(", text, re.M)))


def test_library_exports_every_declared_entry_point():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_calls_without_context_fail_cleanly():
    lib = _lib.load()
    assert lib.mi_bvh_release(0) == -1
    assert lib.mi_sync(None) == -1
    assert lib.mi_release() == -1
    assert b"no context" in lib.mi_last_error() or lib.mi_last_error()
    info = _lib.MiBvhInfo()
    assert lib.mi_bvh_get_info(0, ctypes.byref(info)) == -1


def test_icosphere_counts_and_determinism():
    pos, idx = scenes.displaced_icosphere(subdivisions=3, num_random=100)
    assert idx.shape == (20 * 4 ** 3 + 100, 3)
    assert pos.shape == (10 * 4 ** 3 + 2 + 300, 3)
    assert idx.max() == len(pos) - 1
    pos2, idx2 = scenes.displaced_icosphere(subdivisions=3, num_random=100)
    assert np.array_equal(pos.view(np.uint32), pos2.view(np.uint32))
    assert np.array_equal(idx, idx2)
    # radial displacement of amplitude 0.1 around the unit sphere
    r = np.linalg.norm(pos[: 10 * 4 ** 3 + 2], axis=1)
    assert r.min() >= 0.9 - 1e-6 and r.max() <= 1.1 + 1e-6
    assert r.std() > 0.005
    # each icosphere edge is shared by exactly two triangles (closed surface)
    sphere = idx[: 20 * 4 ** 3]
    edges = np.sort(np.concatenate([sphere[:, [0, 1]], sphere[:, [1, 2]], sphere[:, [2, 0]]]), axis=1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert np.all(counts == 2)
    # random triangles: edges log-normal around exp(-5) by default
    pos_r, idx_r = scenes.displaced_icosphere(subdivisions=0, num_random=20000)
    tri = pos_r[idx_r[20:]]
    edge = np.linalg.norm(tri[:, 1] - tri[:, 0], axis=1)
    assert 0.004 < np.median(edge) < 0.015


def test_c1_mesh_matches_the_fixture_inputs():
    import hashlib
    import json

    stats = json.load(open(os.path.join(ROOT, "tests", "golden", "bvh_c1_stats.json")))
    pos, idx = scenes.displaced_icosphere()
    assert len(idx) == stats["num_triangles"] == 1410720
    assert hashlib.sha256(pos.tobytes()).hexdigest() == stats["positions_sha256"]
    assert hashlib.sha256(idx.tobytes()).hexdigest() == stats["indices_sha256"]


def test_camera_rays_are_unit_and_cover_the_view():
    rays = scenes.camera_rays(64, 48, fov_deg=60.0)
    assert rays.shape == (64 * 48, 8)
    d = rays[:, 4:7]
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)
    assert np.all(rays[:, 3] == 0) and np.all(rays[:, 7] == scenes.FLT_MAX)
    centre = d.mean(0) / np.linalg.norm(d.mean(0))
    fwd = -np.array([0.6, 0.9, 3.2]) / np.linalg.norm([0.6, 0.9, 3.2])
    assert np.dot(centre, fwd) > 0.99


def test_bounce_rays_leave_the_surface_on_the_incoming_side():
    from oracle import oracle as O

    pos, idx = scenes.displaced_icosphere(subdivisions=3, num_random=0)
    cam = scenes.camera_rays(32, 32)
    tree = O.build(pos, idx)
    hits = O.trace(tree, pos, idx, cam)
    b = scenes.bounce_rays(cam, hits, pos, idx, 4096, seed=7)
    assert np.allclose(np.linalg.norm(b[:, 4:7], axis=1), 1, atol=1e-6)
    # a closed outward surface: every bounce ray points away from the centre
    # side of its origin's tangent plane, so cosines with the origin are > 0
    cosang = np.sum(b[:, 0:3] * b[:, 4:7], axis=1) / np.linalg.norm(b[:, 0:3], axis=1)
    assert np.mean(cosang > -0.2) > 0.99
    # cosine distribution: E[cos] = 2/3 about the normal (radial here, approx.)
    assert 0.55 < np.mean(cosang) < 0.75


def test_generators_reject_bad_arguments():
    lib = _lib.load()
    assert lib.mi_gen_camera_rays(0, 0, None, None, None, 60.0, None) == -2
    n = ctypes.c_uint32(0)
    assert lib.mi_gen_displaced_icosphere(11, 4, 0.1, 1, 0, -5.0, 0.5, 2, None, ctypes.byref(n),
                                          None, ctypes.byref(n)) == -2
    with pytest.raises(_lib.MiError):
        rays = np.zeros((4, 8), np.float32)
        hits = np.full((4, 4), 0xFFFFFFFF, np.uint32)
        pos = np.zeros((3, 3), np.float32)
        idx = np.zeros((1, 3), np.uint32)
        scenes.bounce_rays(rays, hits, pos, idx, 4)
