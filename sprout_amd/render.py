"""Host side of the wavefront MIS path tracer (mi_scene_create / mi_render).

Mirrors the reference's per-frame boundary: a scene made of triangles with
one material each (Lambert, Substitute or a one-sided constant light), and
a render call that adds every pixel's radiance sum over the requested
samples into an FP64 accumulator on the device, as
Driver::render_frame_forward (core/rendering/rendering_driver.cpp:295-360)
drives Pathtracer_MIS (core/rendering/integrator/surface/pathtracer_mis.cpp)
over all pixels.  Device memory and streams come from torch; every
computation runs in libsprout_gpu.so.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check
from .scenes import MATERIAL_DTYPE, ROOM_CAMERA


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Scene:
    """Triangles, per-triangle materials and materials (scenes.MATERIAL_DTYPE);
    normals: per-vertex shading normals (set_normals) or None."""

    def __init__(self, positions, indices, triangle_materials, materials, settings=(16, 64, 4),
                 normals=None):
        lib = _lib.load()
        pos = np.ascontiguousarray(positions, np.float32)
        idx = np.ascontiguousarray(indices, np.uint32)
        tmat = np.ascontiguousarray(triangle_materials, np.uint32)
        mats = np.ascontiguousarray(materials, MATERIAL_DTYPE)
        handle = ctypes.c_uint32(_lib.MI_NULL)
        s = _lib.MiBvhSettings(*settings)
        check(lib.mi_scene_create(pos.ctypes.data, len(pos), idx.ctypes.data, len(idx),
                                  tmat.ctypes.data, mats.ctypes.data, len(mats), ctypes.byref(s),
                                  ctypes.byref(handle)), "mi_scene_create")
        self.handle = handle.value
        self.num_triangles = len(idx)
        self.num_vertices = len(pos)
        if normals is not None:
            self.set_normals(normals)

    def set_normals(self, normals) -> None:
        """mi_scene_set_normals: (num_vertices, 3) shading normals, interpolated
        at each hit as sprout interpolates its vertex normals; None goes back
        to the geometric normal."""
        if normals is None:
            check(_lib.load().mi_scene_set_normals(self.handle, None, 0), "mi_scene_set_normals")
            return
        n = np.ascontiguousarray(normals, np.float32)
        if n.shape != (self.num_vertices, 3):
            raise ValueError(f"normals must be ({self.num_vertices}, 3)")
        check(_lib.load().mi_scene_set_normals(self.handle, n.ctypes.data, len(n)),
              "mi_scene_set_normals")

    def update_positions(self, positions) -> float:
        """mi_scene_update_positions: new vertex positions ((n, 3) float32,
        numpy or a device tensor) with the lights, light trees and triangle
        tree rebuilt on the device; returns the light part's milliseconds."""
        p = positions if isinstance(positions, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(positions, np.float32))
        p = p.to(device="cuda", dtype=torch.float32).contiguous()
        if p.dim() != 2 or p.shape[1] != 3:
            raise ValueError("positions must be (n, 3)")
        ms = ctypes.c_double(0.0)
        check(_lib.load().mi_scene_update_positions(self.handle, ctypes.c_void_p(p.data_ptr()),
                                                    p.shape[0], ctypes.byref(ms), _stream()),
              "mi_scene_update_positions")
        return ms.value

    def release(self) -> None:
        if self.handle != _lib.MI_NULL:
            check(_lib.load().mi_scene_release(self.handle), "mi_scene_release")
            self.handle = _lib.MI_NULL


class InstancedScene(Scene):
    """Mesh instances beside world-space triangles: meshes [(positions,
    indices)], instance i showing meshes[mesh_of[i]] with the row-vector 4x4
    matrices world_to_object[i] / object_to_world[i] and the non-emissive
    material instance_materials[i]; the triangles (positions, indices,
    triangle_materials) carry the emitters, as in Scene."""

    def __init__(self, meshes, mesh_of, world_to_object, object_to_world, instance_materials,
                 positions, indices, triangle_materials, materials, settings=(16, 64, 4)):
        lib = _lib.load()
        keep = []
        arr = (_lib.MiMesh * len(meshes))()
        for k, (mp, mi) in enumerate(meshes):
            mp = np.ascontiguousarray(mp, np.float32).reshape(-1, 3)
            mi = np.ascontiguousarray(mi, np.uint32).reshape(-1, 3)
            keep += [mp, mi]
            arr[k] = _lib.MiMesh(mp.ctypes.data, len(mp), mi.ctypes.data, len(mi))
        mesh_of = np.ascontiguousarray(mesh_of, np.uint32)
        w2o = np.ascontiguousarray(world_to_object, np.float32).reshape(-1, 16)
        o2w = np.ascontiguousarray(object_to_world, np.float32).reshape(-1, 16)
        imat = np.ascontiguousarray(instance_materials, np.uint32)
        if not (len(mesh_of) == len(w2o) == len(o2w) == len(imat)):
            raise ValueError("one mesh, two matrices and one material per instance")
        pos = np.ascontiguousarray(positions, np.float32)
        idx = np.ascontiguousarray(indices, np.uint32)
        tmat = np.ascontiguousarray(triangle_materials, np.uint32)
        mats = np.ascontiguousarray(materials, MATERIAL_DTYPE)
        handle = ctypes.c_uint32(_lib.MI_NULL)
        s = _lib.MiBvhSettings(*settings)
        check(lib.mi_scene_create_instanced(arr, len(meshes), mesh_of.ctypes.data, w2o.ctypes.data,
                                            o2w.ctypes.data, imat.ctypes.data, len(mesh_of),
                                            pos.ctypes.data, len(pos), idx.ctypes.data, len(idx),
                                            tmat.ctypes.data, mats.ctypes.data, len(mats),
                                            ctypes.byref(s), ctypes.byref(handle)),
              "mi_scene_create_instanced")
        del keep
        self.handle = handle.value
        self.num_instances = len(mesh_of)
        self.num_triangles = len(idx) + sum(len(np.asarray(m[1]).reshape(-1, 3))
                                            for m in (meshes[j] for j in mesh_of))
        self.num_vertices = len(np.asarray(positions).reshape(-1, 3))
        self.mesh_vertices = [len(np.asarray(m[0]).reshape(-1, 3)) for m in meshes]

    def set_mesh_normals(self, mesh: int, normals) -> None:
        """mi_scene_set_mesh_normals: mesh `mesh`'s vertex shading normals in
        object space ((n, 3)); None goes back to the geometric normal."""
        if normals is None:
            check(_lib.load().mi_scene_set_mesh_normals(self.handle, mesh, None, 0),
                  "mi_scene_set_mesh_normals")
            return
        n = np.ascontiguousarray(normals, np.float32)
        if n.shape != (self.mesh_vertices[mesh], 3):
            raise ValueError(f"normals must be ({self.mesh_vertices[mesh]}, 3)")
        check(_lib.load().mi_scene_set_mesh_normals(self.handle, mesh, n.ctypes.data, len(n)),
              "mi_scene_set_mesh_normals")

    def update_instances(self, world_to_object, object_to_world) -> None:
        """mi_scene_update_instances: new matrices for every instance, in
        creation order; the top-level tree is rebuilt on the device."""
        w2o = np.ascontiguousarray(world_to_object, np.float32).reshape(-1, 16)
        o2w = np.ascontiguousarray(object_to_world, np.float32).reshape(-1, 16)
        if len(w2o) != len(o2w):
            raise ValueError("two matrices per instance")
        check(_lib.load().mi_scene_update_instances(self.handle, w2o.ctypes.data, o2w.ctypes.data,
                                                    len(w2o)), "mi_scene_update_instances")


def perspective(width: int, height: int, position, rotation, fov_deg: float) -> _lib.MiCamera:
    """sprout's Perspective camera: world position, rotation rows r0 (image
    right), r1 (image up), r2 (view direction) as 9 floats, horizontal field
    of view in degrees (the take's "fov")."""
    rot = np.ascontiguousarray(rotation, np.float64).reshape(9)
    return _lib.MiCamera(width, height, (ctypes.c_float * 3)(*position),
                         (ctypes.c_float * 9)(*rot), float(fov_deg))


def look_at(width: int, height: int, eye, target, up, fov_deg: float):
    """A camera at eye looking at target, given a VERTICAL field of view in
    degrees, in sprout's camera form: (position, rotation rows r0 r1 r2 as 9
    float32, horizontal fov in degrees).  Rows (computed in double, rounded
    to float): r2 = normalize(target - eye), r0 = normalize(up x r2) (image
    right in sprout's left-handed camera frame), r1 = r2 x r0."""
    e = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - e
    if not np.linalg.norm(f) > 0:
        raise ValueError("eye and target coincide")
    f /= np.linalg.norm(f)
    r = np.cross(np.asarray(up, np.float64), f)
    if not np.linalg.norm(r) > 0:
        raise ValueError("up is parallel to the view direction")
    r /= np.linalg.norm(r)
    u = np.cross(f, r)
    rot = np.concatenate([r, u, f]).astype(np.float32)
    fov_h = np.degrees(2.0 * np.arctan(np.tan(0.5 * np.radians(float(fov_deg))) * width / height))
    return e.astype(np.float32), rot, np.float32(fov_h)


def camera(width: int, height: int, eye=ROOM_CAMERA["eye"], target=ROOM_CAMERA["target"],
           up=ROOM_CAMERA["up"], fov_deg=ROOM_CAMERA["fov_deg"]) -> _lib.MiCamera:
    """A Perspective camera at eye looking at target with a vertical field of
    view of fov_deg degrees (look_at)."""
    pos, rot, fov_h = look_at(width, height, eye, target, up, fov_deg)
    return perspective(width, height, pos, rot, fov_h)


def _flags(graph: bool, golden_ratio: bool, regenerate: bool = False,
           spectral: bool = False) -> int:
    return ((0 if graph else _lib.MI_RENDER_NO_GRAPH) | (_lib.MI_RENDER_GOLDEN_RATIO if golden_ratio else 0)
            | (_lib.MI_RENDER_REGENERATE if regenerate else 0)
            | (_lib.MI_RENDER_SPECTRAL if spectral else 0))


def render(scene: Scene, cam: _lib.MiCamera, spp: int, samples_per_wave: int = 1,
           min_bounces: int = 4, max_bounces: int = 8, seed: int = 1, first_sample: int = 0,
           graph: bool = True, accum: torch.Tensor | None = None, stats: bool = False,
           golden_ratio: bool = False, regenerate: bool = False, spectral: bool = False):
    """Adds the radiance sums of samples first_sample .. first_sample+spp-1 into
    accum ((height, width, 3) float64 on the device, made zero if None) and
    returns (accum, stats or None).  golden_ratio: pixel samples from sprout's
    Golden_ratio sampler (one frame per call, first_sample 0).  regenerate:
    ended paths hand their slots to the frame's next samples (the same image).
    spectral: hero-wavelength transport over four wavelengths per path."""
    if accum is None:
        accum = torch.zeros((cam.height, cam.width, 3), dtype=torch.float64, device="cuda")
    if accum.dtype != torch.float64 or not accum.is_contiguous() or \
            tuple(accum.shape) != (cam.height, cam.width, 3):
        raise ValueError("accum must be a contiguous (height, width, 3) float64 device tensor")
    st = _lib.MiRenderSettings(spp, samples_per_wave, first_sample, min_bounces, max_bounces,
                               _flags(graph, golden_ratio, regenerate, spectral), seed)
    out = _lib.MiRenderStats() if stats else None
    check(_lib.load().mi_render(scene.handle, ctypes.byref(cam), ctypes.byref(st),
                                ctypes.c_void_p(accum.data_ptr()), _stream(),
                                ctypes.byref(out) if out is not None else None), "mi_render")
    return accum, out


def ggx_e_table() -> torch.Tensor:
    t = torch.empty((32, 32), dtype=torch.float32, device="cuda")
    check(_lib.load().mi_ggx_e_table(ctypes.c_void_p(t.data_ptr()), _stream()), "mi_ggx_e_table")
    return t


def _batch(fn: str, width: int, params, frames, wo, w_or_xi) -> np.ndarray:
    n = len(params)
    dev = [torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
           for a in (params, frames, wo, w_or_xi)]
    out = torch.empty((n, width), dtype=torch.float32, device="cuda")
    check(getattr(_lib.load(), fn)(n, *[ctypes.c_void_p(d.data_ptr()) for d in dev],
                                   ctypes.c_void_p(out.data_ptr()), _stream()), fn)
    return out.cpu().numpy()


def substitute_evaluate(params, frames, wo, wi) -> np.ndarray:
    """(n, 4): reflection * n.wi (rgb), pdf; layouts as mi_substitute_evaluate."""
    return _batch("mi_substitute_evaluate", 4, params, frames, wo, wi)


def substitute_sample(params, frames, wo, xi) -> np.ndarray:
    """(n, 8): wi, reflection (rgb), pdf, type bits (as float32 bits)."""
    return _batch("mi_substitute_sample", 8, params, frames, wo, xi)


def sensor(kind: int = _lib.MI_FILTER_GAUSSIAN, radius: float = 2.0, alpha: float = 1.8,
           b: float = 1.0 / 3.0, c: float = 1.0 / 3.0, clamp_max=(-1.0, 0.0, 0.0)) -> _lib.MiSensor:
    """sprout's sensor: a filter (take defaults: Gaussian, radius 2, alpha 1.8;
    Mitchell b = c = 1/3) and a clamp (clamp_max[0] < 0: none; clamp_max[1]
    < 0: luminance to clamp_max[0]; else per channel)."""
    return _lib.MiSensor(kind, radius, alpha, b, c, (ctypes.c_float * 3)(*clamp_max))


def render_filtered(scene: Scene, cam: _lib.MiCamera, spp: int, sens: _lib.MiSensor,
                    samples_per_wave: int = 1, min_bounces: int = 4, max_bounces: int = 8,
                    seed: int = 1, first_sample: int = 0, graph: bool = True,
                    accum: torch.Tensor | None = None, stats: bool = False,
                    golden_ratio: bool = False, spectral: bool = False):
    """mi_render_filtered: adds the filtered samples into accum ((height,
    width, 4) float64 on the device: weighted rgb and weight; made zero if
    None) and returns (accum, stats or None); the image is rgb / weight."""
    if accum is None:
        accum = torch.zeros((cam.height, cam.width, 4), dtype=torch.float64, device="cuda")
    if accum.dtype != torch.float64 or not accum.is_contiguous() or \
            tuple(accum.shape) != (cam.height, cam.width, 4):
        raise ValueError("accum must be a contiguous (height, width, 4) float64 device tensor")
    st = _lib.MiRenderSettings(spp, samples_per_wave, first_sample, min_bounces, max_bounces,
                               _flags(graph, golden_ratio, spectral=spectral), seed)
    out = _lib.MiRenderStats() if stats else None
    check(_lib.load().mi_render_filtered(scene.handle, ctypes.byref(cam), ctypes.byref(st),
                                         ctypes.byref(sens), ctypes.c_void_p(accum.data_ptr()),
                                         _stream(), ctypes.byref(out) if out is not None else None),
          "mi_render_filtered")
    return accum, out


def render_guided(scene: Scene, cam: _lib.MiCamera, spp: int, guide, fraction: float = 0.5,
                  train: bool = True, train_steps: int = 4, batch: int = 0,
                  learning_rate: float = 1e-3, record_depth: int = 4, samples_per_wave: int = 1,
                  min_bounces: int = 4, max_bounces: int = 8, seed: int = 1,
                  first_sample: int = 0, accum: torch.Tensor | None = None, stats: bool = False,
                  graph: bool = True, guide_depth: int = 0, golden_ratio: bool = False,
                  guide_medium: bool = False, spectral: bool = False):
    """mi_render_guided: as render(), sampling directions from a one-sample
    mixture of the guide (probability fraction) and the BSDF at vertices of
    depth below guide_depth (0: all), and training the guide online after
    every wave when train is set.  guide_medium: scattering vertices in the
    scene's medium mix the guide with the phase function too."""
    if accum is None:
        accum = torch.zeros((cam.height, cam.width, 3), dtype=torch.float64, device="cuda")
    if accum.dtype != torch.float64 or not accum.is_contiguous() or \
            tuple(accum.shape) != (cam.height, cam.width, 3):
        raise ValueError("accum must be a contiguous (height, width, 3) float64 device tensor")
    st = _lib.MiRenderSettings(spp, samples_per_wave, first_sample, min_bounces, max_bounces,
                               _flags(graph, golden_ratio, spectral=spectral), seed)
    gs = _lib.MiGuideSettings(guide.handle, fraction, 1 if train else 0, train_steps, batch,
                              learning_rate, record_depth, guide_depth,
                              1 if guide_medium else 0)
    out = _lib.MiRenderStats() if stats else None
    check(_lib.load().mi_render_guided(scene.handle, ctypes.byref(cam), ctypes.byref(st),
                                       ctypes.byref(gs), ctypes.c_void_p(accum.data_ptr()),
                                       _stream(), ctypes.byref(out) if out is not None else None),
          "mi_render_guided")
    return accum, out
