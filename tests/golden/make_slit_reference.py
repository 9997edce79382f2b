"""Writes the converged image of the slit room that the equal-time guiding
comparison is measured against (needs a GPU).

Slit room (the box room with its light enclosed behind walls open only
through a 0.05-wide slit), 256 x 256 (or --size), plain MIS (mi_render), max
depth 8, Russian roulette from depth 4, 65536 samples per pixel accumulated
in FP64, seed 1000 (independent of every render compared with it).  The
image is stored as float32 radiance (float16 with --half: 2^-11 relative,
far below the errors it measures).

    python tests/golden/make_slit_reference.py [--spp 65536] [--size 256] [--half] [--out PATH]
"""

from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--spp", type=int, default=65536)
    ap.add_argument("--chunk", type=int, default=1024)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--half", action="store_true")
    ap.add_argument("--out", default=None, help="default: slit_room_ref_<size>.npz beside this file")
    ap.add_argument("--partial", default=None,
                    help="FP64 running sum kept here after every chunk; an existing one is resumed")
    ap.add_argument("--resume", default=None,
                    help="start from this running sum (written by --partial) instead")
    ap.add_argument("--seconds", type=float, default=1e9, help="stop (resumable) after this long")
    args = ap.parse_args()
    size = args.size
    out = args.out or os.path.join(HERE, f"slit_room_ref_{size}.npz")

    import torch

    from sprout_amd import bvh, render, scenes

    bvh.init(0)
    pos, idx, tmat, mats = scenes.room(0.05)
    sc = render.Scene(pos, idx, tmat, mats)
    cam = render.camera(size, size, **scenes.ROOM_CAMERA)
    acc = torch.zeros((size, size, 3), dtype=torch.float64, device="cuda")
    done = 0
    start = args.resume or args.partial
    if start and os.path.exists(start):
        with open(start, "rb") as f, np.load(f, allow_pickle=False) as z:
            acc.copy_(torch.from_numpy(z["sum"]))
            done = int(z["done"])
        print(f"resuming at {done} spp", flush=True)
    t0 = time.perf_counter()
    while done < args.spp and time.perf_counter() - t0 < args.seconds:
        n = min(args.chunk, args.spp - done)
        render.render(sc, cam, n, samples_per_wave=16 if n % 16 == 0 else 1, seed=1000,
                      first_sample=done, accum=acc)
        torch.cuda.synchronize()
        done += n
        if args.partial:
            with open(args.partial, "wb") as f:  # the name is kept as given
                np.savez(f, sum=acc.cpu().numpy(), done=np.int64(done))
        print(f"{done} spp, {time.perf_counter() - t0:.1f} s", flush=True)
    if done < args.spp:
        print(f"stopped at {done} of {args.spp} spp", flush=True)
        return
    img = (acc / args.spp).cpu().numpy().astype(np.float16 if args.half else np.float32)
    np.savez_compressed(out, image=img, spp=np.int64(args.spp), seed=np.int64(1000))
    print(f"wrote {out}: mean {img.astype(np.float64).mean():.5f}", flush=True)


if __name__ == "__main__":
    main()
