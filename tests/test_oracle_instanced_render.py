"""The oracle's path tracer over mesh instances (pt.c with
oracle_pt_instances) pinned against the same scene baked into one triangle
soup, which the oracle's triangle path already matches to the reference.

With the same seeds both renders draw the same random numbers per path, and
the two forms of a hit differ only in rounding (object-space interpolation
then transform_point, against world-space vertices moved by transform_point),
so the images agree to a few ulps almost everywhere.  A wrong matrix, normal
transform, mesh or material lookup changes the image far beyond that.
"""

import numpy as np
import pytest

from oracle import oracle as O
from scene_util import baked, instanced_room, light_tree, oracle_instances


def rel_mse(x, r, eps=1e-2):
    return float(np.mean((x - r) ** 2 / (r * r + eps)))


@pytest.fixture(scope="module")
def room():
    sc = instanced_room()
    return sc, baked(sc)


def test_instanced_render_equals_the_baked_soup(room):
    from sprout_amd import scenes

    sc, (bpos, bidx, btm, bmats) = room
    w, h, spp = 32, 24, 8
    tree = light_tree(sc["positions"], sc["indices"], sc["triangle_materials"], sc["materials"])
    inst = O.render(sc["positions"], sc["indices"], sc["triangle_materials"], sc["materials"], w, h,
                    spp, scenes.ROOM_CAMERA, tree, seed=3, instances=oracle_instances(sc)) / spp
    btree = light_tree(bpos, bidx, btm, bmats)
    soup = O.render(bpos, bidx, btm, bmats, w, h, spp, scenes.ROOM_CAMERA, btree, seed=3) / spp
    assert np.all(np.isfinite(inst))
    diff = np.abs(inst - soup) / np.maximum(soup, 1e-3)
    assert np.mean(diff < 1e-3) > 0.95, np.mean(diff < 1e-3)
    assert rel_mse(inst, soup) < 1e-4
    # the instances are in view: without them the image differs
    plain = O.render(sc["positions"], sc["indices"], sc["triangle_materials"], sc["materials"], w,
                     h, spp, scenes.ROOM_CAMERA, tree, seed=3) / spp
    assert rel_mse(plain, soup) > 1e-2

