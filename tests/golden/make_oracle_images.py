#!/usr/bin/env python3
"""Writes the converged FP64 oracle images the RMSE measurements use.

For a workload, a fixed subset of pixels (drawn once with seed 99) is
rendered by the FP64 oracle (oracle/liboracle.so, t-culled traversal of the
reference BVH) with the workload's own lens strata (pw = ceil(sqrt(spp)))
and `--seeds` independent RNG keys (1000, 1001, ...), i.e. seeds x spp
samples per pixel. The result converges to the expectation of one
`spp`-sample render, so the RMSE of a GPU render against it is Monte Carlo
noise plus any bias of the GPU path, nothing else.

Output: tests/golden/oracle_<workload>.npz with pixels, mean rgb, valid
counts, seeds, spp and a hash of the scene arrays.

    python tests/golden/make_oracle_images.py --workload mesh --seeds 1024

`--batch-job K` renders job K of the batched uneven workload
(scenegen.batch_workload, at the job's own resolution and spp) instead, into
tests/golden/oracle_batch<K>.npz. `--crop` renders the 64 x 64 block of
pixels centred in the frame instead of the subset, into
tests/golden/oracle_<workload>_crop.npz; `--pixels 0` the full frame, into
tests/golden/oracle_<workload>_full.npz.
"""
import argparse
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "cuda-raytracer_amd"))

import oracle_lib as ol  # noqa: E402
from koishi_amd import scenegen as sg  # noqa: E402


def scene_hash(s):
    h = hashlib.sha256()
    for a in (s.positions, s.normals, s.indices):
        h.update(np.ascontiguousarray(a).tobytes())
    h.update(repr((s.meshes, s.spheres, sorted(s.shaders.items()), s.camera)).encode())
    return h.hexdigest()


def subset(s, n=4096, seed=99):
    """n pixels drawn once with `seed`; n = 0 means the full frame."""
    npix = s.width * s.height
    if n <= 0 or n >= npix:
        return np.arange(npix, dtype=np.uint32)
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(npix, n, replace=False)).astype(np.uint32)


def crop(s, size=64):
    """The size x size block of pixels centred in the frame, row by row."""
    x0, y0 = (s.width - size) // 2, (s.height - size) // 2
    ys, xs = np.meshgrid(np.arange(y0, y0 + size), np.arange(x0, x0 + size), indexing="ij")
    return (ys * s.width + xs).reshape(-1).astype(np.uint32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="mesh", choices=list(sg.WORKLOADS))
    ap.add_argument("--seeds", type=int, default=1024)
    ap.add_argument("--pixels", type=int, default=4096, help="0 = full frame")
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--batch-job", type=int, default=-1, help="job of the batched workload instead")
    ap.add_argument("--crop", action="store_true", help="the 64 x 64 block centred in the frame")
    args = ap.parse_args()
    if args.batch_job >= 0:
        s = sg.batch_workload()[0][args.batch_job]
        name = f"batch{args.batch_job}"
    else:
        s = sg.WORKLOADS[args.workload]()
        name = args.workload
    if args.crop:
        pix = crop(s)
        name += "_crop"
    else:
        pix = subset(s, args.pixels)
        if pix.size == s.width * s.height and args.batch_job < 0:
            name += "_full"
    orc = ol.OracleScene(s, "f64", traversal=1)
    acc = np.zeros((pix.size, 3))
    valid = np.zeros(pix.size, np.uint64)
    t0 = time.time()
    for k in range(args.seeds):
        p = s.params(seed=1000 + k)
        rgb, val, _ = orc.render(p, 0, p.spp, args.threads, pix)
        acc += rgb
        valid += val
        if (k + 1) % 64 == 0:
            print(f"{k + 1}/{args.seeds} seeds, {time.time() - t0:.0f} s", flush=True)
    mean = acc / np.maximum(valid, 1)[:, None]
    out = os.path.join(HERE, f"oracle_{name}.npz")
    np.savez_compressed(out, pixels=pix, mean=mean, valid=valid, seeds=args.seeds, spp=s.spp,
                        width=s.width, height=s.height, scene_hash=scene_hash(s))
    print(f"wrote {out}: {pix.size} px x {args.seeds * s.spp} samples in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
