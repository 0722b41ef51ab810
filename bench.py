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
sphere R=100) and the fog march (1024x1024 rays through the 512^3 fog) in rays/s;
the iso-surface example (sphere R=100 at (-20,0,0), 4096^2 pixels, persistent and
tiled launches) in rays/s; the differentiable march (256^3 fBm fog seed 7 against a seed-11 target, 64
cameras on a radius-600 shell x 256^2 rays, L2 image loss) as the time ratio of
forward+adjoint to forward; and the quantised batched query (16 level-set and fog
grids of bbox 64^3..1024^3, each as Fp16 / Fp8 / FpN(1e-3), ~2^24 rays per codec
in one persistent launch) in rays/s per codec.

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
    p.add_argument("--no-diff", action="store_true", help="skip the differentiable-march workload")
    p.add_argument("--no-multi", action="store_true", help="skip the quantised multi-grid workload")
    p.add_argument("--no-build", action="store_true", help="skip the device-side grid construction workload")
    p.add_argument("--no-adjoint", action="store_true", help="skip the point-query adjoint timing")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="also run one elementwise copy of MIB MiB (reads and writes MIB MiB) to scale HBM counters")
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
    """Points on the torus surface; createLevelSetTorus lays the ring in the xz-plane
    (tools/CreatePrimitives.h:161), y is its axis."""
    th = torch.rand(m, generator=gen, device="cuda") * (2 * np.pi)
    ph = torch.rand(m, generator=gen, device="cuda") * (2 * np.pi)
    w = R + r * torch.cos(ph)
    return torch.stack([w * torch.cos(th), r * torch.sin(ph), w * torch.sin(th)], 1)


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


def c2_points(torch, nvq, torus, fog, n, seed=1234):
    """The point-query batches of the headline workload, generated on the GPU from one seeded
    generator: 2^k points for the torus, then for the fog (half uniform, half near the surface)."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    p_torus = make_points(torch, nvq, torus, n, gen, torus_surface(torch, gen, 1 << 20))
    p_fog = make_points(torch, nvq, fog, n, gen, fog_surface(torch, nvq, fog, gen, 1 << 20))
    return p_torus, p_fog, gen


def c2_fog_grid(nvq, size=512):
    """The headline fog: fBm (6 octaves, base frequency 1/64, seed 7, amplitude 1) thresholded at 0.15."""
    s = size - 1
    return nvq.tools.create_fog_fbm([0, 0, 0, s, s, s], octaves=6, base_freq=1.0 / 64.0, seed=7, amplitude=1.0,
                                    threshold=0.15)


def leaf_bytes(grid):
    sizes = {1: 2144, 13: 352, 14: 608, 15: 1120}
    return grid.leaf_count * sizes.get(grid.info.grid_type, 2144)


def kernel_ms(kernel):
    """Summed HIP-event time (ms) and launch count of one library kernel since the last reset
    (the library's opt-in profiler records events on the launch stream)."""
    import ctypes
    import openvdb_amd as nvq
    ms = ctypes.c_double()
    n = ctypes.c_uint64()
    nvq.lib().nvq_profile_read(kernel.encode(), ctypes.byref(ms), ctypes.byref(n))
    return ms.value, n.value


def timed_launches(torch, fns, steps, warmup):
    """Run warmup steps, then `steps` timed steps; each step calls every fn once.  Returns the
    wall time of the timed region and the per-fn HIP-event times (ms, summed).  The library's
    kernel profiler is reset and enabled for the timed region only."""
    import openvdb_amd as nvq
    for _ in range(warmup):
        for f in fns:
            f()
    torch.cuda.synchronize()
    nvq.lib().nvq_profile_reset()
    nvq.lib().nvq_profile_enable(1)
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
    nvq.lib().nvq_profile_enable(0)
    per = [sum(ev[s][i][0].elapsed_time(ev[s][i][1]) for s in range(steps)) for i in range(len(fns))]
    return wall, per


def shell_cameras(nvq, count, res, centre, radius, seed, fov=45.0):
    """`count` pinhole cameras (RayGenOp convention) on a seeded sphere shell looking at `centre`."""
    r = np.random.default_rng(seed)
    v = r.normal(size=(count, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    eyes, dirs = [], []
    for k in range(count):
        eye = tuple(float(x) for x in np.asarray(centre) + radius * v[k])
        up = (0.0, 1.0, 0.0) if abs(v[k][1]) < 0.9 else (1.0, 0.0, 0.0)
        e, d = nvq.tools.camera_rays(res, res, eye, tuple(float(c) for c in centre), up, fov)
        eyes.append(e)
        dirs.append(d)
    return np.concatenate(eyes), np.concatenate(dirs)


def diff_march(torch, nvq, steps, cpu=True):
    """Differentiable march: forward time vs forward + L2-loss adjoint time (one step each); the
    rooflines of both kernels; with `cpu`, the gradient error and CPU timing on one camera."""
    box = [0, 0, 0, 255, 255, 255]
    fp_np = nvq.tools.create_fog_fbm(box, seed=7)
    fp = nvq.DeviceGrid(fp_np)
    ft = nvq.DeviceGrid(nvq.tools.create_fog_fbm(box, seed=11))
    e, d = shell_cameras(nvq, 64, 256, (128.0, 128.0, 128.0), 600.0, seed=2024)
    E, D = torch.from_numpy(e).cuda(), torch.from_numpy(d).cuda()
    n = E.shape[0]
    target, _ = nvq.ops.fog_march(ft, E, D, 0.5, 0.05, image_width=256)
    accum = torch.zeros(fp.leaf_count * 512, dtype=torch.int64, device="cuda")
    grad = torch.empty(fp.leaf_count * 512, dtype=torch.float32, device="cuda")

    def fwd():
        return nvq.ops.fog_march(fp, E, D, 0.5, 0.05, image_width=256)

    def fwd_bwd():
        L, T = fwd()
        gL = 2.0 * (L - target)
        accum.zero_()
        nvq.ops.fog_march_backward(fp, E, D, L, gL, None, None, 0.5, 0.05, image_width=256, accum=accum)
        nvq.ops.fixed_to_float(accum, out=grad)

    wf, pf = timed_launches(torch, [fwd], steps, 2)
    kf = kernel_ms("k_fog_march")[0] / steps
    wb, pb = timed_launches(torch, [fwd_bwd], steps, 2)
    kb = kernel_ms("k_fog_march_bwd")[0] / steps
    L, _ = fwd()
    nz = int((accum != 0).sum().item())
    leaves = int((accum.view(-1, 512) != 0).any(1).sum().item())
    U = leaves * 2144
    out = {"rays": n, "leaves": fp.leaf_count, "fwd_ms": 1e3 * wf / steps, "fwd_bwd_ms": 1e3 * wb / steps,
           "ratio": wb / wf, "fwd_rays_per_s": n * steps / wf, "fwd_kernel_ms": kf, "bwd_kernel_ms": kb,
           "loss": float(((L - target) ** 2).sum().item()), "grad_nonzero": nz, "leaves_touched": leaves,
           "roofline": roofline_entry("k_fog_march_bwd", n * (24 + 4 + 4) + 8 * nz + U, kb,
                                      "per ray 24 B eye/dir + 4 B radiance + 4 B dL/dradiance; 8 B per non-zero "
                                      "int64 accumulator; the float values of the leaves the rays touch (2144 B each)"),
           "fwd_roofline": roofline_entry("k_fog_march", n * (24 + 8) + U, kf,
                                          "per ray 24 B eye/dir + 8 B radiance/transmittance; the leaves the rays "
                                          "touch (2144 B each)")}
    if cpu:
        out["cpu_baseline"], out["grad_error"] = c4_cpu_leg(torch, nvq, fp_np, fp, E, D, target)
    return out


def c2_adjoint(torch, nvq, grids, steps):
    """Point-query adjoint on the C2 batches: forward (nearest + trilinear + FD6) vs forward +
    nvq_sample_bwd with seeded upstream gradients for every output, per grid."""
    out = {}
    for name, grid, pts in grids:
        n = pts.shape[0]
        gen = torch.Generator(device="cuda")
        gen.manual_seed(99)
        gn = torch.randn(n, generator=gen, device="cuda")
        gv = torch.randn(n, generator=gen, device="cuda")
        gg = torch.randn((n, 3), generator=gen, device="cuda")
        accum = torch.zeros(grid.grad_size, dtype=torch.int64, device="cuda")

        def fwd():
            nvq.ops.sample(grid, pts, grad_mode=1)

        def bwd():
            accum.zero_()
            nvq.ops.sample_backward(grid, pts, gn, gv, gg, grad_mode=1, accum=accum)

        wf, _ = timed_launches(torch, [fwd], steps, 1)
        wb, _ = timed_launches(torch, [fwd, bwd], steps, 1)
        kb, kn = kernel_ms("k_bin_sample_bwd")
        out[name] = {"fwd_ms": 1e3 * wf / steps, "fwd_bwd_ms": 1e3 * wb / steps, "ratio": wb / wf,
                     "bin_bwd_kernel_ms": kb / max(kn, 1), "values_touched": int((accum != 0).sum().item())}
        del accum
    return out


def c4_cpu_leg(torch, nvq, fp_np, fp, E, D, target, m=256 * 256):
    """Camera 0 of the differentiable march (256^2 rays): the GPU adjoint against the oracle's FP64
    adjoint over every leaf value (the gradient error of the metric), and the oracle's
    single-thread FP64 forward + adjoint timed as the CPU baseline."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle
    E0, D0 = E[:m].contiguous(), D[:m].contiguous()
    L0, _ = nvq.ops.fog_march(fp, E0, D0, 0.5, 0.05, image_width=256)
    gL = (2.0 * (L0 - target[:m])).contiguous()
    acc = nvq.ops.fog_march_backward(fp, E0, D0, L0, gL, dt=0.5, sigma_t=0.05, image_width=256)
    g = nvq.ops.fixed_to_float(acc).cpu().numpy().astype(np.float64)
    w = gL.cpu().numpy().astype(np.float64)
    o = Oracle()
    t0 = time.perf_counter()
    og = o.fog_march_adjoint(fp_np, E0.cpu().numpy(), D0.cpu().numpy(), 0.5, 0.05, w)
    t = time.perf_counter() - t0
    err = {"grad_rel_err": float(np.linalg.norm(g - og) / np.linalg.norm(og)),
           "grad_max_err_rel_to_max": float(np.abs(g - og).max() / np.abs(og).max()),
           "values": int(og.size), "nonzero": int((og != 0).sum()),
           "against": "FP64 oracle adjoint (or_fog_march_adjoint), camera 0 of 64, L2 loss weights 2(L - target)"}
    base = {"value": m / t, "unit": "rays/s (forward + adjoint)", "cores": 1, "kind": "port",
            "sample": "camera 0 of the 64 (256^2 rays): the oracle's FP64 re-march + adjoint, one thread"}
    return base, err


def multi_grid_set(nvq):
    """16 grids of uneven size: level-set spheres and tori with bboxes ~64^3..1024^3 and fBm fogs."""
    from concurrent.futures import ThreadPoolExecutor
    specs = [("sphere", 29.0), ("sphere", 60.0), ("sphere", 90.0), ("sphere", 180.0), ("sphere", 250.0),
             ("sphere", 508.0), ("torus", (24.0, 8.0)), ("torus", (50.0, 14.0)), ("torus", (100.0, 30.0)),
             ("torus", (200.0, 60.0)), ("torus", (380.0, 120.0)), ("fog", (64, 3)), ("fog", (96, 5)),
             ("fog", (128, 7)), ("fog", (192, 9)), ("fog", (256, 13))]

    def build(spec):
        kind, p = spec
        if kind == "sphere":
            return nvq.tools.create_level_set_sphere(p, (0.0, 0.0, 0.0), 1.0, 3.0)
        if kind == "torus":
            return nvq.tools.create_level_set_torus(p[0], p[1], (0.0, 0.0, 0.0), 1.0, 3.0)
        return nvq.tools.create_fog_fbm([0, 0, 0, p[0] - 1, p[0] - 1, p[0] - 1], seed=p[1])

    with ThreadPoolExecutor(8) as ex:
        return specs, list(ex.map(build, specs))


def multi_query_bench(torch, nvq, steps, total_log2=24, cpu=True):
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.time()
    specs, grids = multi_grid_set(nvq)
    log(f"C5: 16 float grids built in {time.time() - t0:.1f}s")
    # rays per grid proportional to the bbox surface area, as square images of side a multiple of 8
    boxes = []
    for g in grids:
        d = nvq.DeviceGrid(g)
        boxes.append(d.index_bbox)
        d.close()
    area = np.array([2 * ((b[3] - b[0] + 1) * (b[4] - b[1] + 1) + (b[4] - b[1] + 1) * (b[5] - b[2] + 1) +
                          (b[3] - b[0] + 1) * (b[5] - b[2] + 1)) for b in boxes], np.float64)
    share = area / area.sum() * (1 << total_log2)
    sides = [max(8, int(round(np.sqrt(s) / 8.0)) * 8) for s in share]
    eyes, dirs, offs = [], [], [0]
    r = np.random.default_rng(77)
    for b, side in zip(boxes, sides):
        lo, hi = np.array(b[:3], np.float64), np.array(b[3:], np.float64) + 1.0
        c = 0.5 * (lo + hi)
        diag = float(np.linalg.norm(hi - lo))
        v = r.normal(size=3)
        v /= np.linalg.norm(v)
        e, d = nvq.tools.camera_rays(side, side, tuple(c + 1.3 * diag * v), tuple(c), (0.0, 1.0, 0.0)
                                     if abs(v[1]) < 0.9 else (1.0, 0.0, 0.0), 45.0)
        eyes.append(e)
        dirs.append(d)
        offs.append(offs[-1] + side * side)
    e_np, d_np = np.concatenate(eyes), np.concatenate(dirs)
    E = torch.from_numpy(e_np).cuda()
    D = torch.from_numpy(d_np).cuda()
    n = E.shape[0]
    fog_rays = sum(s * s for (k, _), s in zip(specs, sides) if k == "fog")
    out = {"rays": n, "grid_rays": [s * s for s in sides], "grids": [f"{k}:{p}" for k, p in specs], "codecs": {}}
    for codec in ("fp16", "fp8", "fpn"):
        t0 = time.time()
        with ThreadPoolExecutor(8) as ex:
            q = list(ex.map(lambda g: nvq.tools.quantize(g, codec, 1e-3), grids))
        dg = [nvq.DeviceGrid(x) for x in q]
        log(f"C5 {codec}: quantised and uploaded in {time.time() - t0:.1f}s")
        res = {}

        def run():
            res["r"] = nvq.ops.multi_query(dg, E, D, offs, sides, 0.5, 0.05)

        w, p = timed_launches(torch, [run], steps, 1)
        hit = res["r"][0]
        kms = kernel_ms("k_multi_query")[0] / steps
        gbytes = int(sum(g.nbytes for g in dg))
        # zeroCrossing rays: 24 B in, hit 1 + ijk 12 + t 4 + v 4 + normal 12 out; march rays: 24 B in, 8 B out
        alg = (n - fog_rays) * (24 + 33) + fog_rays * (24 + 8) + gbytes
        c = {"rays_per_s": n * steps / w, "ms": 1e3 * w / steps, "kernel_ms": kms, "bytes": gbytes,
             "hits": int(hit.sum().item()),
             "roofline": roofline_entry("k_multi_query", alg, kms,
                                        "per zeroCrossing ray 24 B in + 33 B out, per march ray 24 B in + 8 B out, "
                                        "plus every byte of the 16 grids (an upper bound of the footprint)")}
        if cpu:
            c["cpu_baseline"] = c5_cpu_baseline(specs, q, e_np, d_np, offs)
            c["errors"] = c5_errors(torch, nvq, specs, grids, q, dg, e_np, d_np, offs, res["r"], codec)
        out["codecs"][codec] = c
        for g in dg:
            g.close()
        del q
        log(f"C5 {codec}: {c['rays_per_s']/1e6:.1f} M rays/s")
    return out


def c5_cpu_baseline(specs, quant, e_np, d_np, offs, stride=16):
    """The reference host code on the quantised grids (zeroCrossing for level sets, the march for
    fogs) over every `stride`-th ray of each grid, on the host threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Reference, have_reference
    if not have_reference():
        return None
    ref = Reference()
    nt = host_threads()
    sel = [(int(offs[j]), int(offs[j + 1])) for j in range(len(specs))]

    def run():
        for (kind, _), q, (lo, hi) in zip(specs, quant, sel):
            e, d = e_np[lo:hi:stride], d_np[lo:hi:stride]
            if kind == "fog":
                ref.fog_march(q, e, d, 0.5, 0.05, nthreads=nt, fp32=True, fp64=False)
            else:
                ref.zero_crossing(q, e, d, nthreads=nt)
    rays = sum(len(range(lo, hi, stride)) for lo, hi in sel)
    t = median_time(run, runs=3, warm=1)
    return {"value": rays / t, "unit": "rays/s", "cores": nt, "kind": "reference",
            "sample": f"every {stride}th ray of each of the 16 grids ({rays} rays); ReadAccessor + zeroCrossing "
                      f"(level sets) or the FP32 march over the reference HDDA and sampler (fogs), g++ -O2, "
                      f"median of 3 runs after 1 warm-up; {nt} threads"}


def c5_errors(torch, nvq, specs, sources, quant, dgrids, e_np, d_np, offs, gpu_out, codec, npts=1 << 14, stride=8):
    """Per grid, against the FP64 oracle: (i) arithmetic error, the GPU's FP32 decode + sample vs an
    FP64 evaluation of the same quantised codes; (ii) codec error, the FP64 evaluation of the
    quantised grid vs that of its FP32 source.  Point samples (uniform in the padded bbox) and the
    workload's own rays (every `stride`-th): relative radiance errors for fogs, hit-voxel and hit-time
    differences for level sets."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle
    o = Oracle()
    hit, ijk, a = (t.cpu().numpy() for t in gpu_out[:3])
    pts, gpu_val = [], []
    for j, dg in enumerate(dgrids):
        b = dg.index_bbox
        lo, hi = np.array(b[:3], np.float64) - 2, np.array(b[3:], np.float64) + 3
        x = (np.random.default_rng(900 + j).random((npts, 3)) * (hi - lo) + lo).astype(np.float32)
        _, v, _ = nvq.ops.sample(dg, torch.from_numpy(x).cuda(), nn=False, grad_mode=0)
        pts.append(x)
        gpu_val.append(v.cpu().numpy().astype(np.float64))

    def one(j):
        (kind, p), src, q = specs[j], sources[j], quant[j]
        vq, _ = o.sample_f64(q, pts[j])
        vs, _ = o.sample_f64(src, pts[j])
        r = {"grid": f"{kind}:{p}", "sample_arith_max_abs": float(np.abs(gpu_val[j] - vq).max()),
             "sample_codec_max_abs": float(np.abs(vq - vs).max())}
        lo, hi = int(offs[j]), int(offs[j + 1])
        e, d = e_np[lo:hi:stride], d_np[lo:hi:stride]
        if kind == "fog":
            Lq, _, _ = o.fog_march(q, e, d, 0.5, 0.05)
            Ls, _, _ = o.fog_march(src, e, d, 0.5, 0.05)
            lit = Ls > 1e-3
            g = a[lo:hi:stride].astype(np.float64)
            r["ray_arith_max_rel"] = float(np.max(np.abs(g - Lq)[lit] / Lq[lit])) if lit.any() else 0.0
            r["ray_codec_max_rel"] = float(np.max(np.abs(Lq - Ls)[lit] / Ls[lit])) if lit.any() else 0.0
            r["rays_checked"] = int(lit.sum())
        else:
            oh, oijk, ot, _, _ = o.zero_crossing(q, e, d)
            sh, sijk, st, _, _ = o.zero_crossing(src, e, d)
            gh, gijk = hit[lo:hi:stride], ijk[lo:hi:stride]
            r["ray_arith_mismatches"] = int((gh != oh).sum() + (np.any(gijk != oijk, axis=1) & (oh == 1)).sum())
            both = (oh == 1) & (sh == 1)
            r["ray_codec_moved_hits"] = int((np.any(oijk != sijk, axis=1) & both).sum() + (oh != sh).sum())
            r["ray_codec_max_dt"] = float(np.abs(ot[both] - st[both]).max()) if both.any() else 0.0
            r["rays_checked"] = int(len(oh))
        return r

    with ThreadPoolExecutor(8) as ex:
        rows = list(ex.map(one, range(len(specs))))
    log(f"C5 {codec} errors: sample (i) max {max(x['sample_arith_max_abs'] for x in rows):.2e}, "
        f"(ii) max {max(x['sample_codec_max_abs'] for x in rows):.2e}")
    return rows


def fog_voxels(torch, nvq, fog, size, gen):
    """Active voxels of the fog grid, found on the GPU by getValue over its dense lattice, shuffled."""
    ax = torch.arange(size, dtype=torch.int32, device="cuda")
    parts = []
    for x0 in range(0, size, 32):
        xs = ax[x0:x0 + 32]
        lattice = torch.stack(torch.meshgrid(xs, ax, ax, indexing="ij"), -1).reshape(-1, 3)
        _, act = nvq.ops.get_value(fog, lattice)
        parts.append(lattice[act.bool()])
    ijk = torch.cat(parts)
    return ijk[torch.randperm(len(ijk), device="cuda", generator=gen)].contiguous()


def build_bench(torch, nvq, ijk, steps, warmup):
    """Device-side construction (nvq_voxels_to_grid, index grid + per-voxel index) of the given
    voxels: wall time per build including its allocations and two host reads, plus the stages."""
    idx = torch.empty(len(ijk), dtype=torch.int64, device="cuda")
    import ctypes
    lib = nvq.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def build():
        h = ctypes.c_void_p()
        st = lib.nvq_voxels_to_grid(ctypes.c_void_p(ijk.data_ptr()), len(ijk), 20, 1.0, None,
                                    ctypes.c_void_p(idx.data_ptr()), stream, ctypes.byref(h))
        if st != 0:
            raise RuntimeError(f"nvq_voxels_to_grid failed: {st}")
        build.last = h

    def build_and_free():
        build()
        lib.nvq_grid_free(build.last)
    wall, _ = timed_launches(torch, [build_and_free], steps, warmup)
    stages = {k: kernel_ms(k)[0] / steps for k in ("v2g_sort", "v2g_scan", "k_v2g_heads", "k_v2g_leaves")}
    from openvdb_amd._lib import GridInfo
    build()
    gi = GridInfo()
    lib.nvq_grid_get_info(build.last, ctypes.byref(gi))
    lib.nvq_grid_free(build.last)
    return {"voxels": len(ijk), "ms": wall * 1e3 / steps, "voxels_per_s": len(ijk) * steps / wall,
            "stages_ms": stages, "grid_bytes": int(gi.bytes), "leaves": int(gi.node_count[0])}


def points_bench(torch, nvq, xyz, voxel_size, steps, warmup):
    """Device-side point grid (nvq_points_to_grid, World32 blind data) from float world points:
    wall time per build including its allocations, its host reads and the final stream wait."""
    import ctypes
    lib = nvq.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def build_and_free():
        h = ctypes.c_void_p()
        st = lib.nvq_points_to_grid(ctypes.c_void_p(xyz.data_ptr()), len(xyz), voxel_size, 3, None, stream,
                                    ctypes.byref(h))
        if st != 0:
            raise RuntimeError(f"nvq_points_to_grid failed: {st}")
        from openvdb_amd._lib import GridInfo
        gi = GridInfo()
        lib.nvq_grid_get_info(h, ctypes.byref(gi))
        build_and_free.info = (int(gi.bytes), int(gi.node_count[0]))
        lib.nvq_grid_free(h)
    wall, _ = timed_launches(torch, [build_and_free], steps, warmup)
    nbytes, leaves = build_and_free.info
    return {"points": len(xyz), "ms": wall * 1e3 / steps, "points_per_s": len(xyz) * steps / wall,
            "grid_bytes": nbytes, "leaves": leaves}


def points_cpu_baseline(xyz_np, voxel_size, log2n):
    """The oracle's single-thread restatement of pointsToGrid on the first 2^log2n points (the
    reference builds point grids only on the device), median of 3."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle
    o = Oracle()
    sample = np.ascontiguousarray(xyz_np[: 1 << log2n])
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        o.points_to_grid(sample, voxel_size, 3)
        times.append(time.perf_counter() - t0)
    return {"value": len(sample) / float(np.median(times)), "unit": "points/s", "cores": 1, "kind": "port",
            "sample": f"first 2^{log2n} of the same points, World32, median of 3"}


def inplace_bench(torch, nvq, host_grid, fn_name, steps, warmup):
    """Time one in-place device tool (nvq_update_grid_stats / nvq_signed_flood_fill) on a device
    copy of host_grid: wall ms per call (each call waits for its stream) and grid GB/s."""
    import ctypes
    g = nvq.DeviceGrid(host_grid)
    fn = getattr(nvq.lib(), fn_name)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def call():
        st = fn(g.handle, stream)
        if st != 0:
            raise RuntimeError(f"{fn_name} failed: {st}")
    wall, _ = timed_launches(torch, [call], steps, warmup)
    ms = wall * 1e3 / steps
    return {"ms": ms, "grid_bytes": int(host_grid.nbytes), "grid_gb_per_s": host_grid.nbytes / (ms * 1e6)}


def inplace_cpu_baselines(fog_np, torus_np):
    """Host counterparts: the reference's tools::updateGridStats (its util::forEach threads) on
    the fog, and the oracle's single-thread port of the device flood fill on the torus."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle, Reference, have_reference
    out = {}
    if have_reference():
        ref = Reference()
        t0 = time.perf_counter()
        ref.update_grid_stats(fog_np)
        out["grid_stats"] = {"ms": (time.perf_counter() - t0) * 1e3, "kind": "reference",
                             "cores": max(1, ref.hw_threads() // 2),
                             "sample": "tools::updateGridStats(StatsMode::Default) on the same fog grid"}
    t0 = time.perf_counter()
    Oracle().signed_flood_fill(torus_np)
    out["signed_flood_fill"] = {"ms": (time.perf_counter() - t0) * 1e3, "kind": "port", "cores": 1,
                                "sample": "oracle C port of cuda::signedFloodFill on the same torus"}
    return out


def build_cpu_baseline(ijk_np, log2n):
    """The reference host path for the same job (build::Grid<float>::setValue per voxel, then
    createNanoGrid<ValueOnIndex>), on the first 2^log2n voxels, one thread, median of 3."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Reference, have_reference
    if not have_reference():
        return None
    ref = Reference()
    sample = np.ascontiguousarray(ijk_np[: 1 << log2n])
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        ref.grid_from_voxels(sample, 20, 1.0)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    threads = max(1, ref.hw_threads() // 2)  # util::forEach's std::thread fallback (util/ForEach.h:47-48)
    return {"value": len(sample) / t, "unit": "voxels/s", "cores": threads, "kind": "reference",
            "sample": f"first 2^{log2n} of the shuffled fog voxels; build::Grid<float>::setValue (1 thread) + "
                      f"createNanoGrid<ValueOnIndex>(no stats, no tiles; {threads} threads), g++ -O2, median of 3"}


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


def host_threads():
    """Host threads for the reference CPU legs: every core of the host, capped at 16 (the host CPU
    share that goes with one GPU)."""
    return max(1, min(16, os.cpu_count() or 1))


def median_time(fn, runs=5, warm=1):
    for _ in range(warm):
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def roofline_entry(kernel, alg_bytes, avg_ms, note):
    """HBM roofline of one kernel: algorithmic bytes per launch over its average launch time."""
    gbs = alg_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    return {"bound": "hbm", "kernel": kernel, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": gbs / HBM_PEAK_GBS, "traffic": None, "alg_bytes_per_launch": int(alg_bytes),
            "kernel_avg_ms": avg_ms, "bytes": note}


def touched_leaves(torch, nvq, grid, E, D, width, dt=0.5, sigma=0.05):
    """Leaves whose values a march reads: those the adjoint (dL/dradiance = 1) writes to."""
    L, _ = nvq.ops.fog_march(grid, E, D, dt, sigma, image_width=width)
    acc = nvq.ops.fog_march_backward(grid, E, D, L, torch.ones_like(L), dt=dt, sigma_t=sigma, image_width=width)
    leaves = int((acc.view(-1, 512) != 0).any(1).sum().item())
    del acc
    return leaves


def binned_count(torch, nvq, grid, pts, chunk=1 << 22):
    """Queries of a batch that touch a leaf in the 8^3 slots of [cell-1, cell+2]^3: the ones the
    binned path sorts and stages (the others are answered by its classify pass).  Counted with
    getDim (1 inside a leaf) at the slot origins."""
    m = 0
    for s in range(0, pts.shape[0], chunk):
        c = torch.floor(pts[s:s + chunk]).to(torch.int32)
        lo, hi = torch.div(c - 1, 8, rounding_mode="floor"), torch.div(c + 2, 8, rounding_mode="floor")
        hit = torch.zeros(c.shape[0], dtype=torch.bool, device=c.device)
        for a in (0, 1):
            for b in (0, 1):
                for e in (0, 1):
                    sl = torch.stack([hi[:, 0] if a else lo[:, 0], hi[:, 1] if b else lo[:, 1],
                                      hi[:, 2] if e else lo[:, 2]], 1) * 8
                    _, _, dim = nvq.ops.get_value(grid, sl.contiguous(), want_dim=True)
                    hit |= dim == 1
        m += int(hit.sum().item())
    return m


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
    fog_np = c2_fog_grid(nvq, args.fog_size)
    log(f"grids built in {time.time() - t0:.1f}s: torus {torus_np.nbytes/1e6:.1f} MB, fog {fog_np.nbytes/1e6:.1f} MB")
    torus = nvq.DeviceGrid(torus_np)
    fog = nvq.DeviceGrid(fog_np)

    n = 1 << args.log2_points
    p_torus, p_fog, gen = c2_points(torch, nvq, torus, fog, n)
    outs = []
    for _ in range(2):
        outs.append((torch.empty(n, device="cuda"), torch.empty(n, device="cuda"), torch.empty((n, 3), device="cuda")))
    torch.cuda.synchronize()
    log(f"inputs ready: 2 x 2^{args.log2_points} points")

    if args.copy_calibration:
        src = torch.rand(args.copy_calibration << 18, device="cuda")
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calibration} MiB read + {args.copy_calibration} MiB written")
        del src, dst

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
    call_ms = [p / args.steps for p in per]
    # Algorithmic bytes per query: 12 B point in, 4 B nearest + 4 B value + 12 B gradient out; per
    # call also every leaf once (2^26 queries touch them all).  The binned path splits a call into
    # k_bin_classify (all n queries; marks the m that touch a leaf), the radix sort of those m,
    # k_bin_sample (those m, into records) and k_bin_finish (all n outputs: the m records copied,
    # the others answered from the slot table).
    leaves = [leaf_bytes(torus), leaf_bytes(fog)]
    b_call = [n * 32 + leaves[0], n * 32 + leaves[1]]
    binned = os.environ.get("NVQ_SAMPLE_DIRECT") != "1"
    stages = {}
    for k in (["k_bin_classify", "bin_sort", "k_bin_sample", "k_bin_finish"] if binned else ["k_sample"]):
        ms, cnt = kernel_ms(k)
        stages[k] = {"avg_ms": ms / max(cnt, 1), "launches": cnt}
    m = [binned_count(torch, nvq, torus, p_torus), binned_count(torch, nvq, fog, p_fog)] if binned else [n, n]
    per_launch = {  # algorithmic bytes one launch processes, mean of the two grids
        "k_bin_classify": 12 * n,
        "k_bin_sample": sum(32 * mi + lb for mi, lb in zip(m, leaves)) / 2,
        "k_bin_finish": sum(20 * n + 12 * (n - mi) for mi in m) / 2,
        "k_sample": sum(b_call) / 2,
    }
    dom = max((k for k in stages if k in per_launch), key=lambda k: stages[k]["avg_ms"])
    dom_avg = stages[dom]["avg_ms"]
    achieved = per_launch[dom] / (1e-3 * dom_avg) / 1e9 if dom_avg > 0 else 0.0
    call_gbs = sum(b_call) / (1e-3 * sum(call_ms)) / 1e9
    traffic = load_traffic() if dom == "k_bin_sample" else None
    log(f"C2: {ms_step:.3f} ms/step, calls {call_ms[0]:.3f} / {call_ms[1]:.3f} ms, stages "
        + ", ".join(f"{k} {v['avg_ms']:.3f}" for k, v in stages.items())
        + f" ms; binned {m[0]/n:.2f} / {m[1]/n:.2f} of the queries; {value/1e9:.2f} G samples/s; {dom} "
        f"{achieved:.0f} GB/s algorithmic; call {call_gbs:.0f} GB/s")

    extra = {"c2_call_ms": {"torus": round(call_ms[0], 4), "fog": round(call_ms[1], 4)},
             "c2_stages_ms": {k: round(v["avg_ms"], 4) for k, v in stages.items()},
             "c2_binned_queries": {"torus": m[0], "fog": m[1]},
             "c2_dominant_kernel": {"name": dom, "avg_ms": dom_avg, "launches": stages[dom]["launches"],
                                    "alg_bytes_per_launch": int(per_launch[dom])},
             "c2_alg_bytes_per_call": {"torus": b_call[0], "fog": b_call[1]},
             "c2_call_roofline": {"bound": "hbm", "achieved": call_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": call_gbs / HBM_PEAK_GBS,
                                  "bytes": "32 B per query (12 in, 20 out) + every leaf once, over the wall time "
                                           "of the whole nvq_sample call (all its kernels)"},
             "c2_samples_per_s": {"torus": n / (1e-3 * call_ms[0]), "fog": n / (1e-3 * call_ms[1])},
             "fog_grid": {"bytes": int(fog_np.nbytes), "leaves": fog.leaf_count}}

    if not args.no_adjoint:
        extra["c2_adjoint"] = c2_adjoint(torch, nvq, [("torus", torus, p_torus), ("fog", fog, p_fog)], 3)
        log("C2 adjoint: " + ", ".join(f"{k} fwd {v['fwd_ms']:.2f} ms, fwd+bwd {v['fwd_bwd_ms']:.2f} ms "
                                        f"(ratio {v['ratio']:.2f}, k_bin_sample_bwd {v['bin_bwd_kernel_ms']:.2f} ms)"
                                        for k, v in extra["c2_adjoint"].items()))

    if not args.no_extra:
        # C1: level-set ray hit, 512x512 camera at (0,0,400) FOV 45, sphere R=100
        sph_np = nvq.tools.create_level_set_sphere(100.0, (0, 0, 0), 1.0, 3.0)
        sph = nvq.DeviceGrid(sph_np)
        e_np, d_np = nvq.tools.camera_rays(512, 512, (0.0, 0.0, 400.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0)
        e, d = torch.from_numpy(e_np).cuda(), torch.from_numpy(d_np).cuda()
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
        k1 = kernel_ms("k_zero_crossing")[0] / 50
        hit_leaves = int(torch.unique(torch.div(ijk[hit.bool()], 8, rounding_mode="floor"), dim=0).shape[0])
        c1r = {"rays_per_s": 512 * 512 * 50 / w1, "kernel_ms": k1, "hits": int(hit.sum().item()),
               "roofline": roofline_entry("k_zero_crossing", 512 * 512 * (24 + 33) + hit_leaves * 2144, k1,
                                          "per ray 24 B eye/dir in, hit 1 + ijk 12 + t 4 + v 4 + normal 12 B out; "
                                          "the leaves holding hit voxels (2144 B each, a lower bound of the "
                                          "footprint)")}
        if not args.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            from pyoracle import Reference, have_reference
            if have_reference():
                ref, nt = Reference(), host_threads()
                t = median_time(lambda: ref.zero_crossing(sph_np, e_np, d_np, nthreads=nt))
                c1r["cpu_baseline"] = {"value": 512 * 512 / t, "unit": "rays/s", "cores": nt, "kind": "reference",
                                       "sample": f"the whole 512^2 camera; Ray::worldToIndexF + math::zeroCrossing "
                                                 f"+ trilinear-gradient normal, g++ -O2, median of 5 runs after 1 "
                                                 f"warm-up; {nt} threads"}
        extra["c1"] = c1r
        # C3: fog march, 1024x1024 rays, dt 0.5, sigma 0.05, camera outside the bbox looking at its centre
        c = args.fog_size / 2.0
        e3_np, d3_np = nvq.tools.camera_rays(1024, 1024, (c + 0.35 * args.fog_size, c + 0.45 * args.fog_size,
                                                           c + 1.6 * args.fog_size), (c, c, c), (0.0, 1.0, 0.0), 45.0)
        e3, d3 = torch.from_numpy(e3_np).cuda(), torch.from_numpy(d3_np).cuda()
        L = torch.empty(1024 * 1024, device="cuda")
        T = torch.empty(1024 * 1024, device="cuda")

        def c3():
            lib.nvq_fog_march(fog.handle, P(e3), P(d3), 1024 * 1024, 1024, 0.5, 0.05, P(L), P(T),
                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        w3, p3 = timed_launches(torch, [c3], 10, 2)
        k3 = kernel_ms("k_fog_march")[0] / 10
        mean_t = float(T.mean().item())
        leaves3 = touched_leaves(torch, nvq, fog, e3, d3, 1024)
        c3r = {"rays_per_s": 1024 * 1024 * 10 / w3, "kernel_ms": k3, "mean_transmittance": mean_t,
               "leaves_touched": leaves3,
               "roofline": roofline_entry("k_fog_march", 1024 * 1024 * (24 + 8) + leaves3 * 2144, k3,
                                          "per ray 24 B eye/dir in, 8 B radiance/transmittance out; the leaves the "
                                          "rays read (2144 B each)")}
        if not args.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            from pyoracle import Reference, have_reference
            if have_reference():
                ref, nt = Reference(), host_threads()
                es, ds = e3_np[::16].copy(), d3_np[::16].copy()
                t = median_time(lambda: ref.fog_march(fog_np, es, ds, 0.5, 0.05, nthreads=nt, fp32=True, fp64=False),
                                runs=3)
                c3r["cpu_baseline"] = {"value": len(es) / t, "unit": "rays/s", "cores": nt, "kind": "reference",
                                       "sample": f"every 16th ray of the 1024^2 camera ({len(es)} rays); the FP32 "
                                                 f"march over the reference HDDA, ReadAccessor and trilinear sampler "
                                                 f"(oracle/ref_wrap.cpp), g++ -O2, median of 3; {nt} threads"}
        extra["c3"] = c3r
        log(f"C1 {c1r['rays_per_s']/1e9:.2f} G rays/s ({c1r['hits']} hits); C3 {c3r['rays_per_s']/1e6:.1f} M rays/s")
        # iso-surface example (examples/ex_raytrace_iso_surface): createLevelSetSphere(100, (-20,0,0)),
        # 4096^2 pixels, iso = voxel size, RenderOp camera; persistent and tiled launches
        iso_np = nvq.tools.create_level_set_sphere(100.0, (-20.0, 0.0, 0.0), 1.0, 3.0)
        iso_g = nvq.DeviceGrid(iso_np)
        from openvdb_amd._lib import IsoCamera
        cd = nvq.tools.iso_camera(iso_np, 4096, 4096)
        cam = IsoCamera(cd["width"], cd["height"], cd["iso"], cd["voxel_size"], cd["wbbox_dim_z"], cd["tan_half_fov"],
                        (ctypes.c_float * 3)(*cd["origin"]))
        img = torch.empty(4096 * 4096, device="cuda")
        for persistent, kname in ((1, "k_iso_render_persistent"), (0, "k_iso_render")):
            def ci(p=persistent):
                st = lib.nvq_render_iso_surface(iso_g.handle, ctypes.byref(cam), p, P(img), None, None,
                                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
                if st != 0:
                    raise RuntimeError(f"nvq_render_iso_surface failed: {st}")
            wi, _ = timed_launches(torch, [ci], 10, 2)
            key = "iso_render_persistent" if persistent else "iso_render_tiled"
            extra[key] = {"rays_per_s": 4096 * 4096 * 10 / wi, "kernel_ms": kernel_ms(kname)[0] / 10}
        log(f"iso-surface 4096^2: persistent {extra['iso_render_persistent']['rays_per_s']/1e9:.2f} G rays/s, "
            f"tiled {extra['iso_render_tiled']['rays_per_s']/1e9:.2f} G rays/s")
        if not args.no_diff:
            extra["c4"] = diff_march(torch, nvq, 5, cpu=not args.no_cpu)
            ge = extra["c4"].get("grad_error", {}).get("grad_rel_err")
            log(f"C4: fwd {extra['c4']['fwd_ms']:.2f} ms, fwd+bwd {extra['c4']['fwd_bwd_ms']:.2f} ms, "
                f"ratio {extra['c4']['ratio']:.2f}, gradient error vs FP64 {ge}")
        if not args.no_multi:
            extra["c5"] = multi_query_bench(torch, nvq, 3, cpu=not args.no_cpu)
        if not args.no_build:
            vox = fog_voxels(torch, nvq, fog, args.fog_size, gen)
            extra["voxels_to_grid"] = build_bench(torch, nvq, vox, 5, 2)
            if not args.no_cpu:
                extra["voxels_to_grid"]["cpu_baseline"] = build_cpu_baseline(vox.cpu().numpy(), 21)
            b = extra["voxels_to_grid"]
            log(f"voxelsToGrid: {b['voxels']/1e6:.1f} M voxels in {b['ms']:.1f} ms "
                f"({b['voxels_per_s']/1e9:.2f} G voxels/s); stages {b['stages_ms']}")
            del vox
            extra["grid_stats"] = inplace_bench(torch, nvq, fog_np, "nvq_update_grid_stats", 5, 1)
            extra["signed_flood_fill"] = inplace_bench(torch, nvq, torus_np, "nvq_signed_flood_fill", 5, 1)
            if not args.no_cpu:
                for k, v in inplace_cpu_baselines(fog_np, torus_np).items():
                    extra[k]["cpu_baseline"] = v
            log(f"updateGridStats fog: {extra['grid_stats']['ms']:.2f} ms; "
                f"signedFloodFill torus: {extra['signed_flood_fill']['ms']:.2f} ms")
            pts = torch.randn((1 << 24, 3), generator=gen, device="cuda") * 100.0
            extra["points_to_grid"] = points_bench(torch, nvq, pts, 0.5, 5, 2)
            if not args.no_cpu:
                extra["points_to_grid"]["cpu_baseline"] = points_cpu_baseline(pts.cpu().numpy(), 0.5, 21)
            b = extra["points_to_grid"]
            log(f"pointsToGrid: {b['points']/1e6:.1f} M points in {b['ms']:.1f} ms "
                f"({b['points_per_s']/1e9:.2f} G points/s, {b['leaves']} leaves)")
            del pts

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
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
