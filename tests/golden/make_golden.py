"""Writes the golden vectors under tests/golden/ from the reference host code compiled from
its own sources (oracle/_ref/libnanoref.so, built by oracle/Makefile).

c1_sphere_zero_crossing_64x64.json:
    createLevelSetSphere<float>(100, 0, 1, 3); RayGenOp camera 512x512 at (0,0,400), FOV 45
    (examples/ex_raytrace_level_set/common.h:91-125), every 8th pixel in x and y (64x64 rays);
    Ray::worldToIndexF + math::zeroCrossing (math/HDDA.h:189-214): hit, ijk, t, v (t and v as
    float32 bit patterns).
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))


def main():
    from pyoracle import Reference
    ref = Reference()
    args = [100.0, [0.0, 0.0, 0.0], 1.0, 3.0]
    grid = ref.level_set_sphere(*args)
    eye, dir = ref.ray_gen(512, 512, 400.0)
    sel = (np.arange(0, 512, 8)[:, None] * 512 + np.arange(0, 512, 8)[None, :]).reshape(-1)
    eye, dir = eye[sel], dir[sel]
    hit, ijk, t, v, _ = ref.zero_crossing(grid, eye, dir)
    fx = {
        "source": "createLevelSetSphere<float>(100,(0,0,0),1,3) + RayGenOp 512x512 eye (0,0,400) FOV 45, every 8th pixel",
        "grid": {"fn": "level_set_sphere", "args": args, "bytes": int(grid.nbytes)},
        "eye": eye.tolist(),
        "dir": dir.tolist(),
        "hit": hit.tolist(),
        "ijk": ijk.tolist(),
        "t_bits": t.view(np.uint32).tolist(),
        "v_bits": v.view(np.uint32).tolist(),
    }
    with open(os.path.join(HERE, "c1_sphere_zero_crossing_64x64.json"), "w") as fh:
        json.dump(fx, fh, separators=(",", ":"))
    print("hits", int(hit.sum()), "of", len(hit))


if __name__ == "__main__":
    main()
