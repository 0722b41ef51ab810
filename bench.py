#!/usr/bin/env python3
"""Benchmark of the read-only NanoVDB query path on one MI355X.

Headline workload (the metric's point-query configuration): 2^26 query points per
grid on (a) createLevelSetTorus(300, 80, voxel 1, half-width 3) and (b) a 512^3
fBm fog volume (6 octaves, base frequency 1/64, seed 7, threshold 0.15); half of
the points uniform in the index bbox, half Gaussian (sigma 2 voxels) around the
surface; per point: getValue at the nearest voxel + trilinear sample + FD6
gradient.  One step = one pass over both batches.  Inputs are resident in HBM
before the timed region.  value = queries / second over the whole step.

Also reported (same JSON line, "extra"): level-set ray hits (512x512 camera,
sphere R=100) and the fog march (1024x1024 rays through the 512^3 fog) in rays/s.

The CPU baseline times the reference host code (oracle/_ref/libnanoref.so:
ReadAccessor + SampleFromVoxels<Acc,1> + the same FD6 gradient) on a bounded
sample of the same points on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rays/s + trilinear samples/s (%HBM peak); fwd+bwd/fwd time ratio at 1e-4 grad err"
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--log2-points", type=int, default=26)
    p.add_argument("--fog-size", type=int, default=512)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    p.add_argument("--no-extra", action="store_true", help="skip the ray workloads")
    p.add_argument("--cpu-log2-points", type=int, default=21)
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_points(torch, nvq, grid, n, gen, surface):
    """Half uniform in the index bbox, half N(0, 2^2) around surface points (index space)."""
    lo = torch.tensor(grid.index_bbox[:3], dtype=torch.float32, device="cuda")
    hi = torch.tensor(grid.index_bbox[3:], dtype=torch.float32, device="cuda") + 1.0
    nu = n // 2
    uni = torch.rand((nu, 3), generator=gen, device="cuda") * (hi - lo) + lo
    pick = surface[torch.randint(0, surface.shape[0], (n - nu,), generator=gen, device="cuda")]
    near = pick + 2.0 * torch.randn((n - nu, 3), generator=gen, device="cuda")
    return torch.cat([uni, near]).contiguous()


def torus_surface(torch, gen, m, R=300.0, r=80.0):
    th = torch.rand(m, generator=gen, device="cuda") * (2 * np.pi)
    ph = torch.rand(m, generator=gen, device="cuda") * (2 * np.pi)
    w = R + r * torch.cos(ph)
    return torch.stack([w * torch.cos(th), w * torch.sin(th), r * torch.sin(ph)], 1)


def fog_surface(torch, nvq, grid, gen, m):
    """Voxels on the 0.15 iso-contour: active voxels with an inactive +x/+y/+z neighbour."""
    lo = torch.tensor(grid.index_bbox[:3], device="cuda")
    hi = torch.tensor(grid.index_bbox[3:], device="cuda")
    out = []
    got = 0
    while got < m:
        c = (torch.rand((4 * m, 3), generator=gen, device="cuda") * (hi - lo + 1).float() + lo.float()).floor().int()
        _, a = nvq.ops.get_value(grid, c)
        edge = torch.zeros_like(a, dtype=torch.bool)
        for ax in range(3):
            d = torch.zeros(3, dtype=torch.int32, device="cuda")
            d[ax] = 1
            _, b = nvq.ops.get_value(grid, (c + d).contiguous())
            edge |= (a == 1) & (b == 0)
        sel = c[edge].float()
        out.append(sel)
        got += sel.shape[0]
    return torch.cat(out)[:m]


def leaf_bytes(grid):
    sizes = {1: 2144, 13: 352, 14: 608, 15: 1120}
    return grid.leaf_count * sizes.get(grid.info.grid_type, 2144)


def timed_launches(torch, fns, steps, warmup):
    """Run warmup steps, then `steps` timed steps; each step calls every fn once.  Returns the
    wall time of the timed region and the per-fn HIP-event kernel times (ms, summed)."""
    for _ in range(warmup):
        for f in fns:
            f()
    torch.cuda.synchronize()
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in fns]
          for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        for i, f in enumerate(fns):
            ev[s][i][0].record()
            f()
            ev[s][i][1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [sum(ev[s][i][0].elapsed_time(ev[s][i][1]) for s in range(steps)) for i in range(len(fns))]
    return wall, per


def cpu_baseline(grids_np, pts_np, log2n):
    """Reference host code on the host cores, bounded sample (2^log2n points per grid)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Reference, have_reference
    if not have_reference():
        return None
    ref = Reference()
    threads = max(1, min(16, os.cpu_count() or 1))
    n = 1 << log2n
    samples = [p[:n].copy() for p in pts_np]
    # warm-up, then the median of 5 timed runs over both grids
    for g, p in zip(grids_np, samples):
        ref.sample(g, p[: n // 8], grad_mode=1, nthreads=threads)
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        for g, p in zip(grids_np, samples):
            ref.sample(g, p, grad_mode=1, nthreads=threads)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": 2 * n / t, "unit": "samples/s", "cores": threads, "kind": "reference",
            "sample": f"2^{log2n} points per grid (first 2^{log2n} of the GPU batch) on the torus and the fog; "
                      f"ReadAccessor + SampleFromVoxels<Acc,1> + FD6 gradient + nearest getValue, g++ -O2, "
                      f"median of 5 runs after 1 warm-up; {threads} threads"}


def load_traffic():
    path = os.path.join(ROOT, "profiles", "sample_fd6_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("traffic_bytes_per_launch")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              f"(rays and queries shard by batch with a replicated grid, but no multi-GPU run is made here).")
        return 0
    import torch
    import openvdb_amd as nvq
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.init()

    t0 = time.time()
    torus_np = nvq.tools.create_level_set_torus(300.0, 80.0, (0, 0, 0), 1.0, 3.0)
    fs = args.fog_size - 1
    fog_np = nvq.tools.create_fog_fbm([0, 0, 0, fs, fs, fs], octaves=6, base_freq=1.0 / 64.0, seed=7, amplitude=1.0,
                                      threshold=0.15)
    log(f"grids built in {time.time() - t0:.1f}s: torus {torus_np.nbytes/1e6:.1f} MB, fog {fog_np.nbytes/1e6:.1f} MB")
    torus = nvq.DeviceGrid(torus_np)
    fog = nvq.DeviceGrid(fog_np)

    gen = torch.Generator(device="cuda")
    gen.manual_seed(1234)
    n = 1 << args.log2_points
    p_torus = make_points(torch, nvq, torus, n, gen, torus_surface(torch, gen, 1 << 20))
    p_fog = make_points(torch, nvq, fog, n, gen, fog_surface(torch, nvq, fog, gen, 1 << 20))
    outs = []
    for _ in range(2):
        outs.append((torch.empty(n, device="cuda"), torch.empty(n, device="cuda"), torch.empty((n, 3), device="cuda")))
    torch.cuda.synchronize()
    log(f"inputs ready: 2 x 2^{args.log2_points} points")

    lib = nvq.lib()
    import ctypes

    def call(grid, pts, o):
        def f():
            st = lib.nvq_sample(grid.handle, ctypes.c_void_p(pts.data_ptr()), n, 0, ctypes.c_void_p(o[0].data_ptr()),
                                ctypes.c_void_p(o[1].data_ptr()), ctypes.c_void_p(o[2].data_ptr()), 1,
                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
            if st != 0:
                raise RuntimeError(f"nvq_sample failed: {st}")
        return f

    fns = [call(torus, p_torus, outs[0]), call(fog, p_fog, outs[1])]
    wall, per = timed_launches(torch, fns, args.steps, args.warmup)
    ms_step = 1e3 * wall / args.steps
    value = 2 * n * args.steps / wall
    kernel_ms = [p / args.steps for p in per]
    b_io = n * (12 + 4 + 4 + 12)
    b_alg = [b_io + leaf_bytes(torus), b_io + leaf_bytes(fog)]
    achieved = (b_alg[0] + b_alg[1]) / (1e-3 * (kernel_ms[0] + kernel_ms[1])) / 1e9
    traffic = load_traffic()
    log(f"C2: {ms_step:.3f} ms/step, kernels {kernel_ms[0]:.3f} / {kernel_ms[1]:.3f} ms, {value/1e9:.2f} G samples/s, "
        f"{achieved:.0f} GB/s algorithmic")

    extra = {"c2_kernel_ms": {"torus": round(kernel_ms[0], 4), "fog": round(kernel_ms[1], 4)},
             "c2_alg_bytes_per_launch": {"torus": b_alg[0], "fog": b_alg[1]},
             "c2_samples_per_s": {"torus": n / (1e-3 * kernel_ms[0]), "fog": n / (1e-3 * kernel_ms[1])},
             "fog_grid": {"bytes": int(fog_np.nbytes), "leaves": fog.leaf_count}}

    if not args.no_extra:
        # C1: level-set ray hit, 512x512 camera at (0,0,400) FOV 45, sphere R=100
        sph = nvq.DeviceGrid(nvq.tools.create_level_set_sphere(100.0, (0, 0, 0), 1.0, 3.0))
        e, d = nvq.tools.camera_rays(512, 512, (0.0, 0.0, 400.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0)
        e, d = torch.from_numpy(e).cuda(), torch.from_numpy(d).cuda()
        hit = torch.empty(512 * 512, dtype=torch.uint8, device="cuda")
        ijk = torch.empty((512 * 512, 3), dtype=torch.int32, device="cuda")
        tt = torch.empty(512 * 512, device="cuda")
        vv = torch.empty(512 * 512, device="cuda")
        nr = torch.empty((512 * 512, 3), device="cuda")
        P = lambda t: ctypes.c_void_p(t.data_ptr())

        def c1():
            lib.nvq_zero_crossing(sph.handle, P(e), P(d), 512 * 512, 512, P(hit), P(ijk), P(tt), P(vv), P(nr),
                                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        w1, p1 = timed_launches(torch, [c1], 50, 5)
        extra["c1_rays_per_s"] = 512 * 512 * 50 / w1
        extra["c1_kernel_ms"] = p1[0] / 50
        extra["c1_hits"] = int(hit.sum().item())
        # C3: fog march, 1024x1024 rays, dt 0.5, sigma 0.05, camera outside the bbox looking at its centre
        c = args.fog_size / 2.0
        e3, d3 = nvq.tools.camera_rays(1024, 1024, (c + 0.35 * args.fog_size, c + 0.45 * args.fog_size,
                                                     c + 1.6 * args.fog_size), (c, c, c), (0.0, 1.0, 0.0), 45.0)
        e3, d3 = torch.from_numpy(e3).cuda(), torch.from_numpy(d3).cuda()
        L = torch.empty(1024 * 1024, device="cuda")
        T = torch.empty(1024 * 1024, device="cuda")

        def c3():
            lib.nvq_fog_march(fog.handle, P(e3), P(d3), 1024 * 1024, 1024, 0.5, 0.05, P(L), P(T),
                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        w3, p3 = timed_launches(torch, [c3], 10, 2)
        extra["c3_rays_per_s"] = 1024 * 1024 * 10 / w3
        extra["c3_kernel_ms"] = p3[0] / 10
        extra["c3_mean_transmittance"] = float(T.mean().item())
        log(f"C1 {extra['c1_rays_per_s']/1e9:.2f} G rays/s ({extra['c1_hits']} hits); "
            f"C3 {extra['c3_rays_per_s']/1e6:.1f} M rays/s")

    cpu = None
    if not args.no_cpu:
        k = 1 << args.cpu_log2_points
        pts = [p_torus[:k].cpu().numpy(), p_fog[:k].cpu().numpy()]
        cpu = cpu_baseline([torus_np, fog_np], pts, args.cpu_log2_points)
        if cpu:
            log(f"CPU reference: {cpu['value']/1e6:.1f} M samples/s on {cpu['cores']} threads")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "point queries (getValue + trilinear + FD6 gradient), 2^%d points on each of "
                               "levelset torus(300,80,dx 1,band 3) and fBm fog %d^3 (6 oct, f0 1/64, seed 7, thr 0.15); "
                               "50%% uniform in bbox, 50%% N(sigma 2) near the surface" % (args.log2_points, args.fog_size),
                   "points_per_grid": n, "grids": ["levelset_torus_300_80", f"fog_fbm_{args.fog_size}"],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
