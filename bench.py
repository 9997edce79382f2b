"""BVH traversal micro-benchmark on one MI355X: Grays/s closest-hit.

Workload (the reference's BVH micro-benchmark configuration):
  mesh    icosphere subdivided 8x (1 310 720 triangles) radially displaced by
          4-octave value noise (amplitude 0.1, seed 1) + 100 000 triangles in
          the unit cube with log-normal(-5, 0.5) edge lengths (seed 2)
  tree    sprout's SBVH, 16 slices / sweep 64 / 4 per leaf, built on the GPU
  rays    2048 x 2048 pinhole camera rays; 4 194 304 cosine bounce rays from
          their closest hits (offset_ray), min_t 0, max_t FLT_MAX
  step    one closest-hit query of all 4 194 304 bounce rays (rays resident
          in HBM before the timed region)

Prints one JSON line.  Secondary figures (coherent closest-hit, incoherent
any-hit, build time) go into the same line.  The roofline object uses the
reference loop's algorithmic bytes per ray (32 B per node fetched + 64 B per
triangle test, counted on the GPU by the reference-order kernel on the same
tree and rays) over the average kernel time measured with HIP events.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "trace_closest_c1.md")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_threads() -> int:
    """Threads for the CPU baselines: what `nproc` reports, i.e. the
    processing units this process may use (its CPU affinity, limited by
    OMP_NUM_THREADS where that is set)."""
    try:
        return max(1, int(subprocess.run(["nproc"], capture_output=True, text=True,
                                         check=True).stdout.strip()))
    except (OSError, ValueError, subprocess.CalledProcessError):
        return len(os.sched_getaffinity(0))


def host_info() -> dict:
    return {"nproc": host_threads(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "os_cpu_count": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def timed_launches(fn, steps: int, warmup: int):
    """Runs fn() warmup + steps times; returns (wall seconds of the timed
    steps, mean per-launch device time in ms from HIP events)."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()

    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for start, end in events:
        start.record()
        fn()
        end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = float(np.mean([s.elapsed_time(e) for s, e in events]))
    return wall, ms


def cpu_baseline(pos, idx, rays_np, threads: int):
    """sprout's own Tree::intersect (oracle/_ref, compiled from the reference
    sources) on a bounded sample of the same rays; the oracle restatement when
    that library is absent."""
    from oracle import oracle as O

    sample = rays_np
    passes = 2
    for arch in ("x86-64-v3", "native"):
        if os.path.exists(O.ref_lib_path(arch)):
            ref = O.RefTree(pos, idx, threads=threads, arch=arch)
            ref.trace(sample[:65536])
            secs = sum(ref.trace(sample)[1] for _ in range(passes))
            return {
                "value": round(passes * len(sample) / secs / 1e9, 5),
                "unit": "Grays/s",
                "cores": threads,
                "kind": "reference",
                "sample": f"{passes} passes over the {len(sample)} incoherent rays "
                          f"({secs * threads:.1f} thread-seconds), sprout bvh::Tree::intersect "
                          f"(built -O3 -march={arch}) on {threads} host threads; tree built by "
                          f"sprout's Builder_SAH in {ref.build_seconds:.2f} s",
            }
    tree = O.build(pos, idx)
    t0 = time.perf_counter()
    for _ in range(passes):
        O.trace(tree, pos, idx, sample, threads=threads)
    secs = time.perf_counter() - t0
    return {
        "value": round(passes * len(sample) / secs / 1e9, 5),
        "unit": "Grays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{passes} passes over the {len(sample)} incoherent rays, oracle restatement "
                  f"on {threads} threads",
    }


def room_render(frames: int, cpu: bool):
    """The box room (256 x 256, 16 spp, max depth 8, Russian roulette from
    depth 4, seed 1): one frame = one mi_render call of 16 samples of every
    pixel; returns figures for the JSON line."""
    import torch

    from sprout_amd import render, scenes

    pos, idx, tmat, mats = scenes.room()
    sc = render.Scene(pos, idx, tmat, mats)
    cam = render.camera(256, 256, **scenes.ROOM_CAMERA)
    acc = torch.zeros((256, 256, 3), dtype=torch.float64, device="cuda")
    render.render(sc, cam, 16, samples_per_wave=16, seed=1, accum=acc)
    torch.cuda.synchronize()
    paths = rays = 0.0
    t0 = time.perf_counter()
    for f in range(frames):
        _, st = render.render(sc, cam, 16, samples_per_wave=16, seed=1, first_sample=16 * (f + 1),
                              accum=acc, stats=True)
        paths += st.paths
        rays += st.extend_rays + st.shadow_rays
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = {
        "workload": "box room 256x256, 16 spp, max depth 8, RR from depth 4, MIS, Lambert",
        "ms_per_frame": round(wall / frames * 1e3, 3),
        "samples_per_s": round(paths / wall, 1),
        "grays_all_rays": round(rays / wall / 1e9, 4),
        "rays_per_sample": round(rays / paths, 3),
    }
    # the same frames with regenerating waves (a 4-spp path pool refilled
    # from the frame's samples)
    render.render(sc, cam, 16, samples_per_wave=4, seed=1, accum=acc, regenerate=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(frames):
        render.render(sc, cam, 16, samples_per_wave=4, seed=1, first_sample=16 * (f + 1), accum=acc,
                      regenerate=True)
    torch.cuda.synchronize()
    out["regenerating_ms_per_frame"] = round((time.perf_counter() - t0) / frames * 1e3, 3)
    # the same frames through the filtered sensor (Gaussian, radius 2, alpha
    # 1.8: the take default), on the 260 x 260 grid it samples
    sens = render.sensor()
    acc4 = torch.zeros((256, 256, 4), dtype=torch.float64, device="cuda")
    render.render_filtered(sc, cam, 16, sens, samples_per_wave=16, seed=1, accum=acc4)
    torch.cuda.synchronize()
    paths = 0.0
    t0 = time.perf_counter()
    for f in range(frames):
        _, st = render.render_filtered(sc, cam, 16, sens, samples_per_wave=16, seed=1,
                                       first_sample=16 * (f + 1), accum=acc4, stats=True)
        paths += st.paths
    torch.cuda.synchronize()
    fwall = time.perf_counter() - t0
    out["filtered_ms_per_frame"] = round(fwall / frames * 1e3, 3)
    out["filtered_samples_per_s"] = round(paths / fwall, 1)
    out["filtered_sensor"] = "Gaussian radius 2 alpha 1.8; 260x260 camera grid for the 256x256 image"
    sc.release()
    if cpu:
        from oracle import oracle as O

        threads = host_threads()
        t0 = time.perf_counter()
        from sprout_amd import lights

        tree = O.scene_light_tree(pos, idx, tmat, mats, lights.build_tree)
        O.render(pos, idx, tmat, mats, 256, 256, 4, scenes.ROOM_CAMERA, tree, seed=1,
                 threads=threads)
        dt = time.perf_counter() - t0
        out["cpu_port_samples_per_s"] = round(256 * 256 * 4 / dt, 1)
        out["cpu_port_threads"] = threads
        out["cpu_port_sample"] = "256x256 x 4 spp of the same room, oracle/pt.c"
    return out


def cloud_render(frames: int, cpu: bool):
    """Spectral heterogeneous volume as configured: 256^3 density grid of
    seeded 6-octave fBm (max density 20, scattering albedo 0.9,
    Henyey-Greenstein g 0.6) over a flat 32^3 majorant grid, in a box above a
    ground quad, lit by a 1000-triangle emissive quad; hero-wavelength
    transport (4 wavelengths over 380-720 nm), delta tracking for scattering
    and ratio tracking for transmittance; 1024 x 1024, 4 spp per frame, max
    depth 8.  Beside it: the same frames in rgb over sprout's gridtree and
    over the flat grid, and the guided render (the MLP at surface and medium
    vertices, trained online) in spectral mode."""
    import torch

    from sprout_amd import guide, medium, render, scenes

    pos, idx, tmat, mats = scenes.cloud_scene()
    grid = scenes.fbm_density(256, 6, 1, 20.0)
    sc = render.Scene(pos, idx, tmat, mats)
    w = 1024
    spp = 4
    cam = render.camera(w, w, **scenes.CLOUD_CAMERA)
    acc = torch.zeros((w, w, 3), dtype=torch.float64, device="cuda")

    def set_grid(kind):
        medium.set_medium(sc, grid, *scenes.CLOUD_BOX, scenes.CLOUD_CC_A, scenes.CLOUD_CC_S,
                          scenes.CLOUD_G, majorants=kind)

    def timed(spectral, stats=False):
        render.render(sc, cam, spp, samples_per_wave=spp, seed=1, accum=acc, spectral=spectral)
        torch.cuda.synchronize()
        paths = rays = 0.0
        t0 = time.perf_counter()
        for f in range(frames):
            _, st = render.render(sc, cam, spp, samples_per_wave=spp, seed=1,
                                  first_sample=spp * (f + 1), accum=acc, stats=stats,
                                  spectral=spectral)
            if stats:
                paths += st.paths
                rays += st.extend_rays + st.shadow_rays
        torch.cuda.synchronize()
        return time.perf_counter() - t0, paths, rays

    set_grid("flat")
    wall, paths, rays = timed(True, stats=True)
    out = {
        "workload": "fBm cloud 256^3 (max density 20, albedo 0.9, g 0.6), flat 32^3 majorant "
                    "grid, 1000-tri emissive quad, 1024x1024, 4 spp per frame, max depth 8, "
                    "hero-wavelength tracking (4 wavelengths, 380-720 nm)",
        "ms_per_frame": round(wall / frames * 1e3, 3),
        "samples_per_s": round(paths / wall, 1),
        "rays_per_sample": round(rays / paths, 3),
    }
    # the configuration's whole render, timed: 256 spp of the image in
    # 4-spp waves
    acc.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    render.render(sc, cam, 256, samples_per_wave=spp, seed=2, accum=acc, spectral=True)
    torch.cuda.synchronize()
    out["seconds_for_256_spp"] = round(time.perf_counter() - t0, 3)
    out["rgb_flat_majorants_ms_per_frame"] = round(timed(False)[0] / frames * 1e3, 3)
    set_grid("gridtree")
    out["rgb_gridtree_ms_per_frame"] = round(timed(False)[0] / frames * 1e3, 3)
    out["spectral_gridtree_ms_per_frame"] = round(timed(True)[0] / frames * 1e3, 3)
    # guided: the MLP at surface and medium vertices, one Adam step per wave
    set_grid("flat")
    g = guide.Guide(1)
    render.render_guided(sc, cam, spp, g, fraction=0.5, train=True, train_steps=1,
                         batch=262144, learning_rate=5e-3, samples_per_wave=spp, seed=1,
                         accum=acc, guide_medium=True, spectral=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(frames):
        render.render_guided(sc, cam, spp, g, fraction=0.5, train=True, train_steps=1,
                             batch=262144, learning_rate=5e-3, samples_per_wave=spp, seed=1,
                             first_sample=spp * (f + 1), accum=acc, guide_medium=True,
                             spectral=True)
    torch.cuda.synchronize()
    out["guided_spectral_ms_per_frame"] = round((time.perf_counter() - t0) / frames * 1e3, 3)
    g.release()
    sc.release()
    if cpu:
        from oracle import oracle as O
        from sprout_amd import lights

        threads = host_threads()
        om = O.Medium(grid, *scenes.CLOUD_BOX, scenes.CLOUD_CC_A, scenes.CLOUD_CC_S, scenes.CLOUD_G,
                      majorants="flat")
        tree = O.scene_light_tree(pos, idx, tmat, mats, lights.build_tree)
        t0 = time.perf_counter()
        O.render(pos, idx, tmat, mats, 256, 256, 16, scenes.CLOUD_CAMERA, tree, seed=1,
                 threads=threads, medium=om, spectral=True)
        dt = time.perf_counter() - t0
        out["cpu_port_samples_per_s"] = round(256 * 256 * 16 / dt, 1)
        out["cpu_port_threads"] = threads
        out["cpu_port_sample"] = ("256x256 x 16 spp of the same cloud, spectral, flat majorants, "
                                  "oracle/pt.c + oracle/medium.c")
    return out


def room_take_render(frames: int, cpu: bool, threads: int):
    """The box room as a sprout take (tests/golden/take_room: the walls as
    Substitutes of roughness 1, the ceiling light, the benchmark camera) at
    256 x 256, 16 spp per frame, Random sampler, no sensor filter, min / max
    bounces 4 / 8: the GPU against the reference's own renderer (oracle/_ref/
    sprout_render, compiled from its sources) on the host's threads, on a
    bounded sample of the same frames (128 spp of 256 x 256)."""
    import json
    import subprocess
    import tempfile

    import torch

    from sprout_amd import render, take

    here = os.path.join(ROOT, "tests", "golden", "take_room")
    t = take.Take(os.path.join(here, "room.take"))
    pos, idx, tmat, mats = t.scene
    sc = render.Scene(pos, idx, tmat, mats, normals=t.normals)
    cam = render.perspective(256, 256, **t.camera)
    acc = torch.zeros((256, 256, 3), dtype=torch.float64, device="cuda")
    kw = dict(samples_per_wave=16, seed=1, min_bounces=t.min_bounces, max_bounces=t.max_bounces)
    render.render(sc, cam, 16, accum=acc, **kw)
    torch.cuda.synchronize()
    paths = 0.0
    t0 = time.perf_counter()
    for f in range(frames):
        _, st = render.render(sc, cam, 16, first_sample=16 * (f + 1), accum=acc, stats=True, **kw)
        paths += st.paths
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sc.release()
    out = {"workload": "box room take 256x256 (Substitute walls), 16 spp per frame, "
                       "min/max bounces 4/8, MIS, shading normals",
           "ms_per_frame": round(wall / frames * 1e3, 3),
           "samples_per_s": round(paths / wall, 1)}
    if not cpu:
        return out
    for arch in ("x86_64_v3", None):
        exe = os.path.join(ROOT, "oracle", "_ref", "sprout_render" + (f"_{arch}" if arch else ""))
        if os.path.exists(exe):
            break
    else:
        out["cpu_reference"] = {"skipped": "oracle/_ref/sprout_render not built here"}
        return out
    with open(os.path.join(here, "room.take")) as f:
        doc = json.load(f)
    spp = 128
    doc["sampler"] = {"Random": {"samples_per_pixel": spp}}
    doc["camera"]["Perspective"]["sensor"] = {"resolution": [256, 256]}
    with tempfile.TemporaryDirectory() as tmp:
        t0 = time.perf_counter()
        subprocess.run([exe, here, json.dumps(doc), str(threads), os.path.join(tmp, "img.f32")],
                       check=True, capture_output=True)
        secs = time.perf_counter() - t0
    out["cpu_reference"] = {
        "samples_per_s": round(256 * 256 * spp / secs, 1), "cores": threads, "kind": "reference",
        "sample": f"{spp} spp of 256x256 by sprout's own renderer ({os.path.basename(exe)}: "
                  f"take and scene loading, Pathtracer_MIS, sensor) on {threads} host threads, "
                  f"{secs:.1f} s wall including loading"}
    out["gpu_over_cpu_reference"] = round(out["samples_per_s"] / out["cpu_reference"]["samples_per_s"], 1)
    return out


def cloud_guided_equal_time(seconds: float = 1.5, ref_spp: int = 65536):
    """The fBm cloud as configured (flat majorants, hero-wavelength
    tracking), at 256 x 256: plain spectral MIS against guided spectral MIS
    with the guide at surface and medium vertices (trained online from
    scratch in the first fifth of its time budget, then frozen), each for
    `seconds`; relMSE against a `ref_spp` plain spectral render made here.
    The plain side reaches about 3 600 spp in 1.5 s, so the reference's own
    variance is about 5 % of the measured error."""
    import torch

    from sprout_amd import guide, medium, render, scenes

    pos, idx, tmat, mats = scenes.cloud_scene()
    grid = scenes.fbm_density(256, 6, 1, 20.0)
    sc = render.Scene(pos, idx, tmat, mats)
    medium.set_medium(sc, grid, *scenes.CLOUD_BOX, scenes.CLOUD_CC_A, scenes.CLOUD_CC_S,
                      scenes.CLOUD_G, majorants="flat")
    size = 256
    cam = render.camera(size, size, **scenes.CLOUD_CAMERA)
    ref, _ = render.render(sc, cam, ref_spp, samples_per_wave=64, seed=99, spectral=True)
    torch.cuda.synchronize()
    ref = ref.cpu().numpy() / ref_spp
    acc = torch.zeros((size, size, 3), dtype=torch.float64, device="cuda")
    spw = 16
    render.render(sc, cam, spw, samples_per_wave=spw, seed=2, spectral=True)
    torch.cuda.synchronize()
    spp = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        render.render(sc, cam, spw, samples_per_wave=spw, seed=11, first_sample=spp, accum=acc,
                      spectral=True)
        torch.cuda.synchronize()
        spp += spw
    e_plain = rel_mse(acc.cpu().numpy() / spp, ref)
    out = {"workload": "fBm cloud 256^3, flat majorants, spectral, 256x256, max depth 8",
           "seconds": seconds, "reference_spp": ref_spp, "plain_spp": spp,
           "plain_relmse": e_plain}
    for depth in (1, 3):
        g = guide.Guide(8)
        acc.zero_()
        gspp = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < seconds:
            train = time.perf_counter() - t0 < 0.2 * seconds
            n = 4 if train else spw
            render.render_guided(sc, cam, n, g, fraction=0.5, train=train, train_steps=4,
                                 batch=65536, learning_rate=3e-3, record_depth=depth,
                                 guide_depth=depth, samples_per_wave=n, seed=12,
                                 first_sample=gspp, accum=acc, guide_medium=True, spectral=True)
            torch.cuda.synchronize()
            gspp += n
        e = rel_mse(acc.cpu().numpy() / gspp, ref)
        out[f"guided_depth{depth}_spp"] = gspp
        out[f"guided_depth{depth}_relmse"] = e
        out[f"guided_depth{depth}_equal_time_gain"] = round(e_plain / e, 3)
        g.release()
    sc.release()
    return out


def batch_render(frames: int, instanced: bool):
    """Batched path tracing: 64 placements of 4 distinct 20 480-triangle
    displaced spheres with per-instance Substitute materials, a ground plane
    and 10 000 emissive triangles in the light tree (1.32 M triangles);
    instanced (a tree per mesh under a top-level tree) or baked into one
    world-space soup; 1920 x 1080, 4 spp per frame in waves of 1 spp, Russian
    roulette from depth 3, max depth 16."""
    import torch

    from sprout_amd import render, scenes

    t0 = time.perf_counter()
    if instanced:
        b = scenes.batch_scene_instanced()
        sc = render.InstancedScene(b["meshes"], b["mesh_of"], b["world_to_object"],
                                   b["object_to_world"], b["instance_materials"], b["positions"],
                                   b["indices"], b["triangle_materials"], b["materials"])
    else:
        pos, idx, tmat, mats = scenes.batch_scene()
        sc = render.Scene(pos, idx, tmat, mats)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    cam = render.camera(1920, 1080, **scenes.BATCH_CAMERA)
    acc = torch.zeros((1080, 1920, 3), dtype=torch.float64, device="cuda")
    render.render(sc, cam, 1, samples_per_wave=1, min_bounces=3, max_bounces=16, seed=1, accum=acc)
    torch.cuda.synchronize()
    paths = rays = 0.0
    t0 = time.perf_counter()
    for f in range(frames):
        _, st = render.render(sc, cam, 4, samples_per_wave=1, min_bounces=3, max_bounces=16, seed=1,
                              first_sample=1 + 4 * f, accum=acc, stats=True)
        paths += st.paths
        rays += st.extend_rays + st.shadow_rays
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    extra = {}
    # the configuration's whole render, timed: 64 spp of 1080p, frame by frame
    acc.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(16):
        render.render(sc, cam, 4, samples_per_wave=1, min_bounces=3, max_bounces=16, seed=3,
                      first_sample=4 * f, accum=acc)
    torch.cuda.synchronize()
    extra["seconds_for_64_spp"] = round(time.perf_counter() - t0, 3)
    # the same frames with regenerating waves (a 1-spp pool of path slots
    # refilled from the frame's 4 spp)
    render.render(sc, cam, 4, samples_per_wave=1, min_bounces=3, max_bounces=16, seed=1, accum=acc,
                  regenerate=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(frames):
        render.render(sc, cam, 4, samples_per_wave=1, min_bounces=3, max_bounces=16, seed=1,
                      first_sample=1 + 4 * f, accum=acc, regenerate=True)
    torch.cuda.synchronize()
    extra["regenerating_ms_per_frame"] = round((time.perf_counter() - t0) / frames * 1e3, 2)
    if instanced:
        # an animated frame's top-level rebuild: every instance moved 0.01 up
        # per step (mi_scene_update_instances, host matrices in, tree on the device)
        o2w = b["object_to_world"].reshape(-1, 4, 4).astype(np.float64)
        times = []
        for k in range(1, 6):
            m = o2w.copy()
            m[:, 3, 1] += 0.01 * k
            w2o, o2w_k = np.linalg.inv(m).astype(np.float32), m.astype(np.float32)
            t1 = time.perf_counter()
            sc.update_instances(w2o, o2w_k)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t1)
        extra["top_level_rebuild_ms"] = round(float(np.median(times)) * 1e3, 3)
    else:
        # an animated frame of the soup: the 10 000 emitters moved (each part
        # translated), lights, light trees and triangle tree rebuilt on the
        # device (mi_scene_update_positions, device positions in)
        rng = np.random.default_rng(5)
        kinds = mats["kind"][tmat]
        em = kinds == 2
        verts = idx[em].reshape(-1)
        owner = np.repeat(tmat[em], 3)
        v, first = np.unique(verts, return_index=True)
        times, light = [], []
        for k in range(1, 4):
            moved = pos.copy()
            moved[v] += rng.uniform(-0.05, 0.05, (len(mats), 3)).astype(np.float32)[owner[first]]
            dev = torch.from_numpy(moved).cuda()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            light.append(sc.update_positions(dev))
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t1)
        extra["emitter_update_ms"] = round(float(np.median(times)) * 1e3, 2)
        extra["emitter_update_light_trees_ms"] = round(float(np.median(light)), 2)
    sc.release()
    form = "instanced: 4 mesh trees under a top-level tree" if instanced else "baked, one tree"
    return {
        **extra,
        "workload": f"64 instances x 20480-tri spheres ({form}, 1.32M tris), 10000 lights, "
                    "1920x1080, 4 spp per frame, RR from depth 3, max depth 16",
        "scene_setup_s": round(setup, 2),
        "ms_per_frame": round(wall / frames * 1e3, 2),
        "samples_per_s": round(paths / wall, 1),
        "grays_all_rays": round(rays / wall / 1e9, 4),
        "rays_per_sample": round(rays / paths, 3),
    }


def instance_trace(steps: int, warmup: int, cpu: bool = False):
    """Instanced meshes: 64 instances of 4 distinct 20 480-triangle displaced
    spheres (seeded similarity transforms in a 20-unit cube, 1.31 M instanced
    triangles); closest and any hit of 4 M rays from a sphere of radius 25
    toward random points in the cube."""
    import torch

    from sprout_amd import bvh, instances, scenes

    trees, meshes = [], []
    for k in range(4):
        pos, idx = scenes.displaced_icosphere(5, octaves=4, amplitude=0.15, noise_seed=7 + k,
                                              num_random=0)
        meshes.append((pos, idx))
        trees.append(bvh.Builder_SAH(16, 64, 4).build(torch.from_numpy(pos).cuda(),
                                                      torch.from_numpy(idx.view(np.int32)).cuda()))
    n_inst = 64
    w2o, o2w = instances.random_transforms(n_inst, 107, extent=10.0)
    mesh_of = np.arange(n_inst) % 4
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    inst = instances.Instances([trees[m] for m in mesh_of], w2o, o2w)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3

    rng = np.random.default_rng(8)
    n = 1 << 22
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * 25.0
    d = rng.uniform(-10, 10, (n, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 4:7] = d
    rays[:, 7] = 3.4e38
    d_rays = torch.from_numpy(rays).cuda()
    hits = torch.empty((n, 8), dtype=torch.int32, device="cuda")
    _, ms_closest = timed_launches(lambda: inst.intersect(d_rays), steps, warmup)
    _, ms_any = timed_launches(lambda: inst.intersect_p(d_rays), max(10, steps // 5), 2)
    hits = inst.intersect(d_rays)
    torch.cuda.synchronize()
    frac = float((hits.cpu().numpy().view(np.uint32)[:, 4] != 0xFFFFFFFF).mean())
    out = {
        "workload": "64 instances of 4 x 20480-tri displaced spheres, 4M rays into a 20-unit cube",
        "closest_grays": round(n / (ms_closest * 1e-3) / 1e9, 4),
        "any_grays": round(n / (ms_any * 1e-3) / 1e9, 4),
        "closest_kernel_ms": round(ms_closest, 4),
        "hit_fraction": round(frac, 4),
        "top_level_build_ms": round(build_ms, 2),
    }
    if cpu:
        from oracle import oracle as O

        oms = [(p, i, O.build(p, i)) for p, i in meshes]
        oi = O.Instances(oms, mesh_of.astype(np.uint32), w2o, o2w)
        m = 200000
        t0 = time.perf_counter()
        oi.trace(rays[:m])
        dt = time.perf_counter() - t0
        out["cpu_port_grays"] = round(m / dt / 1e9, 5)
        out["cpu_port_threads"] = 1
        out["cpu_port_sample"] = f"{m} of the rays, oracle/instances.c, one thread"
    inst.release()
    return out


SLIT_REF = os.path.join(ROOT, "tests", "golden", "slit_room_ref_1024.npz")
SLIT_REF_256 = os.path.join(ROOT, "tests", "golden", "slit_room_ref_256.npz")


def guide_kernels():
    """The guide's BF16 MFMA kernels on device-resident batches: forward of
    1 M queries and the loss gradient of 256 k records, mean kernel time from
    HIP events over 20 launches each.  MFMA work per sample: forward
    2 (224*64 + 64*64 + 64*64 + 64*32) = 49 152 flops; gradient = forward +
    back-propagation through three layers + the four weight gradients
    = 3 * 49 152 - 2 * 224 * 64 = 118 784 flops (the first layer's input
    gradient is not formed).  Utilisation = achieved / 2.5 PFLOP/s (dense
    BF16 on MI355X)."""
    import torch

    from sprout_amd import guide

    g = guide.Guide(5)
    n_fwd, n_grad = 1 << 20, 1 << 18
    x = torch.rand((n_fwd, 9), dtype=torch.float32, device="cuda")
    d = torch.nn.functional.normalize(torch.randn((n_grad, 3), device="cuda"), dim=1)
    w = torch.rand((n_grad, 1), dtype=torch.float32, device="cuda")
    xg = x[:n_grad].contiguous()
    _, ms_f = timed_launches(lambda: g.forward(x), 20, 3)
    _, ms_g = timed_launches(lambda: g.gradients(xg, d, w), 20, 3)
    g.release()
    flops_f = 2 * (224 * 64 + 64 * 64 + 64 * 64 + 64 * 32)
    flops_g = 3 * flops_f - 2 * 224 * 64
    tf_f = n_fwd * flops_f / (ms_f * 1e-3) / 1e12
    tf_g = n_grad * flops_g / (ms_g * 1e-3) / 1e12
    return {
        "forward_ms_per_1M_queries": round(ms_f, 4),
        "forward_tflops": round(tf_f, 1),
        "forward_mfma_utilisation": round(tf_f / 2500.0, 4),
        "gradient_ms_per_256k_records": round(ms_g, 4),
        "gradient_tflops": round(tf_g, 1),
        "gradient_mfma_utilisation": round(tf_g / 2500.0, 4),
        "note": "utilisation = MFMA flops of the layers / kernel time / 2.5 PFLOP/s dense BF16; "
                "the kernel time also holds the sin/cos features, the vMF terms and LDS staging",
    }


def rel_mse(x, r, eps=1e-2):
    """mean over pixels and channels of (x - r)^2 / (r^2 + eps), eps = 1e-2"""
    return float(np.mean((x - r) ** 2 / (r * r + eps)))


def slit_equal_time(seconds: float):
    """Slit room (light behind walls open through a 0.05-wide slit), max depth
    8, at 1024 x 1024 (256 x 256 when only that reference exists): plain MIS
    and guided MIS (trained online from scratch inside the same time budget:
    one-spp waves with 4 Adam steps of 65 536 records for the first fifth of
    the time, then 4-spp waves with the frozen guide at the first vertex)
    each render for `seconds`; relMSE of each against the 65536-spp FP64
    reference image."""
    import torch

    from sprout_amd import guide, render, scenes

    path = SLIT_REF if os.path.exists(SLIT_REF) else SLIT_REF_256
    if not os.path.exists(path):
        return {"skipped": "reference image missing (tests/golden/make_slit_reference.py)"}
    with np.load(path, allow_pickle=False) as z:
        ref = z["image"].astype(np.float64)
        ref_spp = int(z["spp"])
    size = ref.shape[0]
    spw = 4 if size >= 1024 else 16
    pos, idx, tmat, mats = scenes.room(0.05)
    sc = render.Scene(pos, idx, tmat, mats)
    cam = render.camera(size, size, **scenes.ROOM_CAMERA)

    # warm up both paths (kernels, buffers, graphs)
    g = guide.Guide(7)
    render.render(sc, cam, spw, samples_per_wave=spw, seed=2)
    render.render_guided(sc, cam, 1, g, train=True, batch=65536, samples_per_wave=1, seed=3,
                         guide_depth=1, record_depth=1)
    render.render_guided(sc, cam, spw, g, train=False, samples_per_wave=spw, seed=3, guide_depth=1)
    torch.cuda.synchronize()
    g.release()

    acc = torch.zeros((size, size, 3), dtype=torch.float64, device="cuda")
    spp = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        render.render(sc, cam, spw, samples_per_wave=spw, seed=11, first_sample=spp, accum=acc)
        torch.cuda.synchronize()
        spp += spw
    t_plain = time.perf_counter() - t0
    e_plain = rel_mse(acc.cpu().numpy() / spp, ref)

    g = guide.Guide(8)
    acc.zero_()
    gspp = 0
    trained = 0.0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        train = time.perf_counter() - t0 < 0.2 * seconds
        n = 1 if train else spw
        _, st = render.render_guided(sc, cam, n, g, fraction=0.5, train=train, train_steps=4,
                                     batch=65536, learning_rate=2e-3, record_depth=1, guide_depth=1,
                                     samples_per_wave=n, seed=12, first_sample=gspp, accum=acc,
                                     stats=True)
        torch.cuda.synchronize()
        gspp += n
        trained += st.trained
    t_guided = time.perf_counter() - t0
    e_guided = rel_mse(acc.cpu().numpy() / gspp, ref)
    g.release()
    sc.release()
    return {
        "workload": f"slit room {size}x{size}, max depth 8, RR from depth 4; reference {ref_spp} spp FP64",
        "seconds_plain": round(t_plain, 3),
        "seconds_guided": round(t_guided, 3),
        "spp_plain": spp,
        "spp_guided": gspp,
        "relmse_plain": round(e_plain, 8),
        "relmse_guided": round(e_guided, 8),
        "relmse_ratio_plain_over_guided": round(e_plain / e_guided, 3),
        "relmse_eps": 1e-2,
        "guide": "fraction 0.5 at the first vertex; trained the first 20 % of the time",
        "guide_samples_trained": trained,
        "reference_spp": ref_spp,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--width", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--guide-seconds", type=float, default=6.0,
                    help="time budget of each side of the guided vs plain comparison")
    ap.add_argument("--part", choices=["all", "room", "room_take", "volume", "volume_guiding",
                                       "guiding",
                                       "instances", "batch",
                                       "batch_instanced", "guide_kernels"],
                    default="all",
                    help="run one workload only and print its figures (for profiling)")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the traversal benchmark's line, without the other workloads' figures")
    ap.add_argument("--shape", default="",
                    help="traversal launch shape kernel,lds_entries,refill_lanes,claim_batch,"
                         "blocks_per_cu[,sort_rays] instead of autotuning (for profiles: the "
                         "timed launches then follow the one camera-ray launch)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="launch one 256 MiB device copy first (calibrates HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the ray batch "
              "shards across GPUs with a replicated tree, which this run does not measure)",
              file=sys.stderr)
        return 2

    import torch

    from sprout_amd import bvh, scenes

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    bvh.init(0)

    if args.copy_probe:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.part != "all":
        parts = {"room": lambda: room_render(5, False), "volume": lambda: cloud_render(3, False),
                 "volume_guiding": cloud_guided_equal_time,
                 "room_take": lambda: room_take_render(5, not args.no_cpu_baseline, host_threads()),
                 "batch": lambda: batch_render(2, False),
                 "batch_instanced": lambda: batch_render(2, True),
                 "guide_kernels": guide_kernels,
                 "guiding": lambda: slit_equal_time(args.guide_seconds),
                 "instances": lambda: instance_trace(args.steps, args.warmup)}
        print(json.dumps({"part": args.part, "figures": parts[args.part]()}), flush=True)
        return 0

    log("generating mesh")
    pos, idx = scenes.displaced_icosphere()
    d_pos = torch.from_numpy(pos).cuda()
    d_idx = torch.from_numpy(idx.view(np.int32)).cuda()

    log("building SBVH on the GPU")
    builder = bvh.Builder_SAH(16, 64, 4)
    tree = builder.build(d_pos, d_idx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tree = builder.build(d_pos, d_idx)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    info = tree.info()
    log(f"  {info.num_nodes} nodes, {info.num_slots} slots, depth {info.max_depth}, "
        f"{build_s * 1e3:.1f} ms")

    if args.shape:
        kernel, lds, lanes, batch, per_cu, sort = (int(x) for x in (args.shape + ",0").split(",")[:6])
        bvh.set_trace_config(256, per_cu, kernel, lanes, batch, lds, sort)
        shape = {"block_size": 256, "blocks_per_cu": per_cu, "kernel": kernel,
                 "refill_lanes": lanes, "claim_batch": batch, "lds_entries": lds,
                 "sort_rays": sort}
        log(f"launch shape {shape}")
    w = args.width
    log(f"camera rays {w}x{w}")
    cam = scenes.camera_rays(w, w)
    d_cam = torch.from_numpy(cam).cuda()
    cam_hits = tree.intersect(d_cam)
    torch.cuda.synchronize()
    cam_hits_np = cam_hits.cpu().numpy()
    log("bounce rays")
    bounce = scenes.bounce_rays(cam, cam_hits_np, pos, idx, w * w, seed=3)
    d_bounce = torch.from_numpy(bounce).cuda()
    n = len(bounce)

    if not args.shape:
        log("autotuning the traversal's launch shape")
        shape, tuned_ms = bvh.autotune(tree, d_bounce, 5)
        log(f"  {shape}: {tuned_ms:.3f} ms")

    hits = torch.empty((n, 4), dtype=torch.int32, device="cuda")
    occ = torch.empty((n,), dtype=torch.int32, device="cuda")
    hits_cam = torch.empty((n, 4), dtype=torch.int32, device="cuda")

    log(f"timing closest-hit, {args.steps} steps of {n} incoherent rays")
    wall, ms_incoh = timed_launches(lambda: tree.intersect(d_bounce, hits), args.steps, args.warmup)
    value = n * args.steps / wall / 1e9

    _, ms_coh = timed_launches(lambda: tree.intersect(d_cam, hits_cam), max(10, args.steps // 5), 2)
    _, ms_any = timed_launches(lambda: tree.intersect_p(d_bounce, occ), max(10, args.steps // 5), 2)

    log("counting reference-loop work per ray")
    _, cnt = tree.counts(d_bounce)
    torch.cuda.synchronize()
    cnt = cnt.cpu().numpy().astype(np.int64)
    alg_bytes = int(32 * cnt[:, 0].sum() + 64 * cnt[:, 1].sum())
    achieved = alg_bytes / (ms_incoh * 1e-3) / 1e9

    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        from sprout_amd.rocprof_summary import read_figures

        prof = read_figures(PROFILE_TRAFFIC)
        if prof.get("num_rays") == n and prof.get("num_nodes") == info.num_nodes:
            traffic = prof.get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline:
        threads = host_threads()
        log(f"CPU baseline on {threads} threads")
        cpu = cpu_baseline(pos, idx, bounce, threads)

    line = {
        "metric": "Grays/s closest-hit; relMSE at equal time, guided vs plain GPU MIS PT",
        "value": round(value, 4),
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "BVH micro-benchmark: SBVH(16,64,4) over a 1.41M-triangle displaced "
                        "icosphere + random triangles, closest hit of 4M incoherent cosine bounce rays",
            "triangles": int(len(idx)),
            "rays": int(n),
            "tree_nodes": int(info.num_nodes),
            "tree_depth": int(info.max_depth),
            "parallelism": "single GPU",
            "launch_shape": shape,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_ray": round(alg_bytes / n, 1),
            "kernel_ms": round(ms_incoh, 4),
        },
        "cpu_baseline": cpu,
        "extra": {
            "coherent_closest_grays": round(n / (ms_coh * 1e-3) / 1e9, 4),
            "incoherent_any_grays": round(n / (ms_any * 1e-3) / 1e9, 4),
            "incoherent_closest_kernel_grays": round(n / (ms_incoh * 1e-3) / 1e9, 4),
            "gpu_build_ms": round(build_s * 1e3, 2),
            "nodes_per_ray": round(float(cnt[:, 0].mean()), 2),
            "tri_tests_per_ray": round(float(cnt[:, 1].mean()), 2),
            "bounce_hit_fraction": None,
        },
    }
    hit_frac = float((hits.cpu().numpy().view(np.uint32)[:, 3] != 0xFFFFFFFF).mean())
    line["extra"]["bounce_hit_fraction"] = round(hit_frac, 4)
    line["extra"]["host"] = host_info()
    if args.no_extras:
        print(json.dumps(line), flush=True)
        return 0
    log("box room path tracing")
    line["extra"]["room_render"] = room_render(5, not args.no_cpu_baseline)
    log("box room take: GPU against sprout's own renderer")
    line["extra"]["room_take"] = room_take_render(5, not args.no_cpu_baseline, host_threads())
    log("batched path tracing")
    line["extra"]["batch_render"] = batch_render(2, False)
    line["extra"]["batch_render_instanced"] = batch_render(2, True)
    log("instanced meshes")
    line["extra"]["instances"] = instance_trace(args.steps, args.warmup, not args.no_cpu_baseline)
    log("fBm cloud path tracing")
    line["extra"]["volume_render"] = cloud_render(3, not args.no_cpu_baseline)
    log("fBm cloud: guided vs plain at equal time")
    line["extra"]["volume_guiding"] = cloud_guided_equal_time()
    log("guide kernels")
    line["extra"]["guide_kernels"] = guide_kernels()
    log("slit room: guided vs plain at equal time")
    line["extra"]["guiding"] = slit_equal_time(args.guide_seconds)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
