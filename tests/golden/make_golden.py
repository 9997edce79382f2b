"""Writes the golden fixtures of the BVH path from the reference's own code.

Runs only in a container that has the reference sources: it builds
oracle/_ref/libsprout_ref.so (oracle/ref.mk, the reference's Builder_SAH and
bvh::Tree compiled from their sources) and records, for seeded synthetic
inputs made by this repository's generators:

  bvh_small.npz     a 3.3 k-triangle mesh, the reference's serialized tree,
                    4096 rays (camera, bounce, random, axis-aligned, clipped,
                    aimed at vertices) and the reference's closest hits and
                    occlusion flags, plus this host's rcpss values so a test
                    can tell whether the native-reciprocal oracle can be
                    compared bit for bit on the host it runs on.
  bvh_c1_stats.json the 1.41 M-triangle micro-benchmark mesh: the
                    reference tree's node/slot counts, depth and SHA-256 of
                    its nodes (bounds by value) and slots.
  bsdf_small.npz    the reference's Substitute material sample (Disney +
                    GGX): 4096 evaluate cases and 4096 sample cases over random
                    frames, directions, colours, roughness, metallic, and the
                    32 x 32 GGX energy table it interpolates.

    python tests/golden/make_golden.py
"""

from __future__ import annotations

import hashlib
import json
import os
import platform
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from sprout_amd import scenes  # noqa: E402

FLT_MAX = np.float32(3.4028234663852886e38)


def canonical_node_bytes(nodes: np.ndarray) -> bytes:
    """Node array with bounds as values (+0 for -0) and leaf axis/pad cleared."""
    n = nodes.copy()
    n["lo"] = n["lo"] + np.float32(0)
    n["hi"] = n["hi"] + np.float32(0)
    n["axis"] = np.where(n["num"] == 0, n["axis"], 0)
    n["pad"] = 0
    return n.tobytes()


def tree_digest(tree: O.Tree) -> dict:
    return {
        "num_nodes": int(tree.num_nodes),
        "num_slots": int(len(tree.slots)),
        "max_depth": int(tree.depth()),
        "nodes_sha256": hashlib.sha256(canonical_node_bytes(tree.nodes)).hexdigest(),
        "slots_sha256": hashlib.sha256(tree.slots.astype("<u4").tobytes()).hexdigest(),
    }


def small_rays(pos: np.ndarray, idx: np.ndarray, tree: O.Tree, rng) -> np.ndarray:
    cam = scenes.camera_rays(32, 32, eye=(0.6, 0.9, 3.2), fov_deg=45.0)
    hits = O.trace(tree, pos, idx, cam)
    bounce = scenes.bounce_rays(cam, hits, pos, idx, 1024, seed=3)

    n = 1024
    o = (rng.random((n, 3)) * 3.0 - 1.5).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rand = np.concatenate([o, np.zeros((n, 1), np.float32), d, np.full((n, 1), FLT_MAX)], 1)

    # axis-aligned directions: zero (and negative zero) components make
    # reciprocal3 return NaN lanes, which the slab test filters
    m = 512
    o = (rng.random((m, 3)) * 2.4 - 1.2).astype(np.float32)
    axis = rng.integers(0, 3, m)
    sgn = rng.choice([-1.0, 1.0], m).astype(np.float32)
    d = np.zeros((m, 3), np.float32)
    d[np.arange(m), axis] = sgn
    negz = rng.random((m, 3)) < 0.5
    d = np.where((d == 0) & negz, np.float32(-0.0), d)
    aligned = np.concatenate([o, np.zeros((m, 1), np.float32), d, np.full((m, 1), FLT_MAX)], 1)

    # clipped intervals: min_t > 0 and finite max_t
    k = 256
    c = rand[:k].copy()
    c[:, 3] = rng.random(k).astype(np.float32) * 1.5
    c[:, 7] = c[:, 3] + rng.random(k).astype(np.float32) * 1.5

    # aimed exactly at vertices and edge midpoints: tie and edge cases
    j = 256
    targets = np.concatenate([pos[idx[rng.integers(0, len(idx), j // 2), 0]],
                              0.5 * (pos[idx[rng.integers(0, len(idx), j // 2), 0]] +
                                     pos[idx[rng.integers(0, len(idx), j // 2), 1]])]).astype(np.float32)
    eye = np.array([0.6, 0.9, 3.2], np.float32)
    d = targets - eye
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    aimed = np.concatenate([np.tile(eye, (j, 1)), np.zeros((j, 1), np.float32), d.astype(np.float32),
                            np.full((j, 1), FLT_MAX)], 1)

    return np.ascontiguousarray(
        np.concatenate([cam, bounce, rand, aligned, c, aimed]).astype(np.float32))


def bsdf_cases(rng, n):
    def unit(v):
        return v / np.linalg.norm(v, axis=1, keepdims=True)

    normal = unit(rng.normal(size=(n, 3))).astype(np.float32)
    frames = np.zeros((n, 12), np.float32)
    for i in range(n):
        t, b = O.orthonormal_basis(normal[i])
        frames[i, 0:3] = t
        frames[i, 3:6] = b
    frames[:, 6:9] = normal
    frames[:, 9:12] = normal
    wo = unit(rng.normal(size=(n, 3)))
    wo = np.where((wo * normal).sum(1, keepdims=True) < 0, -wo, wo).astype(np.float32)
    wi = unit(rng.normal(size=(n, 3)))
    wi = np.where((wi * normal).sum(1, keepdims=True) < 0, -wi, wi).astype(np.float32)
    params = np.zeros((n, 6), np.float32)
    params[:, 0:3] = rng.uniform(0.05, 0.95, (n, 3))
    params[:, 3] = rng.uniform(0.0, 1.0, n)
    params[:, 4] = (rng.random(n) < 0.3).astype(np.float32)
    params[: n // 4, 4] = rng.uniform(0, 1, n // 4)
    params[:, 5] = 1.46
    # a few back-facing wo (no contribution in the reference)
    wo[-16:] = -wo[-16:]
    xi = rng.random((n, 3)).astype(np.float32)
    return params, frames, wo, wi, xi


def main() -> None:
    subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "ref.mk")], check=True)

    rng = np.random.default_rng(20261202)

    pos, idx = scenes.displaced_icosphere(subdivisions=3, octaves=4, amplitude=0.1, noise_seed=1,
                                          num_random=2000, log_mu=-3.0, log_sigma=0.5,
                                          random_seed=2)
    ref = O.RefTree(pos, idx)
    tree = ref.tree()

    rays = small_rays(pos, idx, tree, rng)
    assert len(rays) == 4096, len(rays)

    hits, _ = ref.trace(rays)
    occluded, _ = ref.trace(rays, any_hit=True)

    probe = np.ascontiguousarray(rng.uniform(-4, 4, 256).astype(np.float32))
    native_rcp = np.array([O.lib().oracle_reciprocal(float(x), O.RCP_NATIVE) for x in probe],
                          np.float32)

    np.savez_compressed(
        os.path.join(HERE, "bvh_small.npz"),
        positions=pos, indices=idx, nodes=tree.nodes.view(np.uint8).reshape(-1, 32),
        slots=tree.slots, rays=rays, hits=hits, occluded=occluded, rcp_probe=probe,
        rcp_native=native_rcp,
        host=np.array(platform.processor() or platform.machine()),
    )

    ref_bsdf = O.ref_bsdf_lib()
    params, frames, wo, wi, xi = bsdf_cases(np.random.default_rng(7), 4096)
    np.savez_compressed(
        os.path.join(HERE, "bsdf_small.npz"),
        params=params, frames=frames, wo=wo, wi=wi, xi=xi,
        evaluate=O.substitute_evaluate(params, frames, wo, wi, ref_bsdf, "sref"),
        sample=O.substitute_sample(params, frames, wo, xi, ref_bsdf, "sref"),
        e_table=O.ggx_e_table(ref_bsdf, "sref"),
    )

    pos1, idx1 = scenes.displaced_icosphere()
    ref1 = O.RefTree(pos1, idx1)
    stats = tree_digest(ref1.tree())
    stats.update({
        "mesh": "displaced icosphere, 8 subdivisions, 4 octaves, amplitude 0.1, noise seed 1, "
                "+100000 random triangles log-normal(-5, 0.5), seed 2",
        "num_triangles": int(len(idx1)),
        "num_vertices": int(len(pos1)),
        "settings": [16, 64, 4],
        "positions_sha256": hashlib.sha256(pos1.tobytes()).hexdigest(),
        "indices_sha256": hashlib.sha256(idx1.tobytes()).hexdigest(),
    })
    with open(os.path.join(HERE, "bvh_c1_stats.json"), "w") as f:
        json.dump(stats, f, indent=1)
    print(json.dumps(stats, indent=1))


if __name__ == "__main__":
    main()
