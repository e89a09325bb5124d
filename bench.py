"""Benchmark: fwd+bwd iterations per second of the Gaussian rasterizer at
1M Gaussians, 800x800, on one MI355X.

Workload (default ``--workload single``): 1M free Gaussians in FP32 — means a
60/40 mixture of the unit-ball interior and a radius-3 shell, log-scales
N(-3.5, 0.4), normalised N(0,1)^4 quaternions, sigmoid(N(0,1)) opacity, SH
degree 3 (DC N(0, 0.5), rest N(0, 0.1)) — seen by one 800x800 pinhole camera
(FoV 50 deg, distance 4, white background).  One step = drop-in ``render`` +
0.8 L1 + 0.2 (1 - SSIM) against a seeded target (fused HIP loss kernels, same
objective as utils/loss_utils.py) + ``backward`` into all six parameter
tensors.  Inputs are resident in HBM before timing starts.

Prints ONE JSON line.  ``roofline`` is for the dominant kernel, timed with
device events on the launching stream over the timed region; ``cpu_baseline``
times the FP64 CPU oracle on a bounded sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

HBM_PEAK_GBS = 8000.0
METRIC = "fwd+bwd it/s @1M Gaussians 800² vs plain; render p50/p99 ms; HBM GB/s"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--gaussians", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=800)
    ap.add_argument("--height", type=int, default=800)
    ap.add_argument("--cpu-baseline-seconds", type=float, default=20.0,
                    help="time budget of the sampled CPU-oracle baseline (0 disables)")
    ap.add_argument("--render-samples", type=int, default=100, help="forward-only renders for p50/p99")
    return ap.parse_args()


class StageTimer:
    """Device events around every C-ABI stage, recorded on the launching (current) stream."""

    def __init__(self):
        self.events = {}
        self.enabled = True

    def __call__(self, name):
        timer = self

        class _Ctx:
            def __enter__(self):
                if timer.enabled:
                    self.a = torch.cuda.Event(enable_timing=True)
                    self.a.record()
                return self

            def __exit__(self, *exc):
                if timer.enabled:
                    b = torch.cuda.Event(enable_timing=True)
                    b.record()
                    timer.events.setdefault(name, []).append((self.a, b))
                return False
        return _Ctx()

    def averages_ms(self):
        return {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in self.events.items()}

    def reset(self):
        self.events = {}


def algorithmic_bytes(P, V, K, X, T):
    """Compulsory HBM bytes per stage for one view (each tensor read or written once)."""
    return {
        "geometry": 236 * P + 4 * P + 48 * V + 8 * P,
        "binning": 20 * V + 12 * K + 24 * K + 8 * K + 8 * T,
        "render_forward": 8 * T + 40 * K + 20 * X,
        "render_backward": 8 * T + 40 * K + 20 * X + 36 * V,
        "geometry_backward": 276 * V + 248 * P,
    }


def cpu_baseline(model, cam, target, budget_s):
    """FP64 CPU oracle on a bounded sample: full preprocess, binning and
    geometry backward; blend forward+backward on random tiles until the budget
    is spent, extrapolated by instance count.  Returns the baseline dict."""
    from oracle import raster as OR
    threads = min(16, len(os.sched_getaffinity(0)))
    torch.set_num_threads(threads)
    cpu = lambda t: t.detach().cpu()
    means, opac = cpu(model.get_xyz), cpu(model.get_opacity)
    shs, scales, rots = cpu(model.get_features), cpu(model.get_scaling), cpu(model.get_rotation)
    st = OR.settings_for(_cpu_cam(cam), torch.ones(3), 3)
    W, H = cam.image_width, cam.image_height
    t0 = time.perf_counter()
    pre = OR.preprocess(means, opac, st, shs=shs, scales=scales, rotations=rots)
    t1 = time.perf_counter()
    keys, vals, ranges = OR.bin_gaussians(pre["xy"].float().numpy(), pre["radii"].numpy(),
                                          pre["depth"].float().numpy(), W, H)
    t2 = time.perf_counter()
    K = int(keys.size)
    rng = np.random.default_rng(0)
    tiles = rng.permutation(np.nonzero(ranges[:, 1] > ranges[:, 0])[0])
    dpix = torch.sign(torch.randn(3, H, W, dtype=torch.float64)) / (3 * H * W)
    done_k, t_tiles, n_tiles = 0, 0.0, 0
    deadline = time.perf_counter() + budget_s
    for tile in tiles:
        a = time.perf_counter()
        OR.render_forward(ranges, vals, pre["xy"], pre["conic"], pre["opacity"], pre["rgb"], st.bg, W, H,
                          tiles=[tile])
        OR.render_backward(ranges, vals, pre["xy"], pre["conic"], pre["opacity"], pre["rgb"], st.bg, W, H, dpix,
                           tiles=[tile])
        t_tiles += time.perf_counter() - a
        done_k += int(ranges[tile, 1] - ranges[tile, 0])
        n_tiles += 1
        if time.perf_counter() > deadline:
            break
    P = means.shape[0]
    dm2 = torch.zeros(P, 3, dtype=torch.float64)
    dcon = torch.zeros(P, 4, dtype=torch.float64)
    dcol = torch.zeros(P, 3, dtype=torch.float64)
    a = time.perf_counter()
    OR.geometry_backward(pre, st, means, dm2, dcon, dcol, shs=shs, scales=scales, rotations=rots)
    t_gb = time.perf_counter() - a
    t_blend = t_tiles * K / max(done_k, 1)
    total = (t1 - t0) + (t2 - t1) + t_blend + t_gb
    return {"value": 1.0 / total, "unit": "it/s", "cores": threads, "kind": "port",
            "sample": (f"FP64 torch oracle, {P} Gaussians {W}x{H}: preprocess {t1 - t0:.2f}s, binning "
                       f"{t2 - t1:.2f}s and geometry backward {t_gb:.2f}s over all Gaussians; blend fwd+bwd on "
                       f"{n_tiles} random tiles ({done_k} of {K} instances, {t_tiles:.1f}s) extrapolated by "
                       f"instance count to {t_blend:.1f}s; loss excluded")}


def _cpu_cam(cam):
    class C:
        pass
    c = C()
    for k in ("image_width", "image_height", "FoVx", "FoVy"):
        setattr(c, k, getattr(cam, k))
    c.world_view_transform = cam.world_view_transform.cpu()
    c.full_proj_transform = cam.full_proj_transform.cpu()
    c.camera_center = cam.camera_center.cpu()
    return c


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the rasterizer path here is "
              "single-GPU and no multi-GPU figure is produced.", file=sys.stderr)
        sys.exit(2)
    from dmgs_amd import cameras, scenes
    from dmgs_amd import rasterizer as RZ
    from dmgs_amd.gaussian_renderer import PipelineParams, render
    from dmgs_amd.loss import fused_dssim_l1

    dev = "cuda"
    torch.cuda.set_device(0)
    W, H, N = args.width, args.height, args.gaussians
    model = scenes.free_scene(N, seed=0, layout="mixture", shell_radius=3.0, device=dev).requires_grad_(True)
    cam = cameras.front_camera(W, H, fov_deg=50.0, distance=4.0, device=dev)
    target = scenes.target_image(H, W, seed=1, device=dev)
    bg = torch.ones(3, device=dev)
    pipe = PipelineParams()

    def step():
        for p in model.parameters():
            p.grad = None
        out = render(cam, model, pipe, bg)
        loss = fused_dssim_l1(out["render"], target, 0.2)
        loss.backward()
        return out

    log(f"scene ready: {N} Gaussians, {W}x{H}; warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    timer = StageTimer()
    RZ.set_stage_hook(timer)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    RZ.set_stage_hook(None)
    stage_ms = timer.averages_ms()
    ms_step = 1000.0 * elapsed / args.steps
    it_s = args.steps / elapsed

    # sizes of the workload (from the last step's outputs)
    with torch.no_grad():
        radii = out["radii"]
        V = int((radii > 0).sum().item())
        _, _, st = RZ.forward_raw(model.get_xyz.detach(), model.get_opacity.detach(), model.get_features.detach(),
                                  None, model.get_scaling.detach(), model.get_rotation.detach(), None,
                                  _settings(cam, bg))
        K = st["K"]
    X = W * H
    T = ((W + 15) // 16) * ((H + 15) // 16)
    ab = algorithmic_bytes(N, V, K, X, T)
    kern = max(("render_forward", "render_backward"), key=lambda k: stage_ms.get(k, 0.0))
    achieved = ab[kern] / (stage_ms[kern] * 1e-3) / 1e9
    traffic = _read_traffic(kern)

    # forward-only render latency (plain path)
    lat = []
    with torch.no_grad():
        for _ in range(args.render_samples):
            torch.cuda.synchronize()
            a = time.perf_counter()
            render(cam, model, pipe, bg)
            torch.cuda.synchronize()
            lat.append(1000.0 * (time.perf_counter() - a))
    lat = np.array(lat)

    total_bytes = sum(ab.values())
    result = {
        "metric": METRIC,
        "value": round(it_s, 3),
        "unit": "it/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{N} free Gaussians (60/40 unit ball / r=3 shell), {W}x{H}, one view, "
                               "fwd+bwd of 0.8 L1 + 0.2 (1-SSIM)",
                   "gaussians": N, "width": W, "height": H, "path": "plain", "visible": V, "instances": K},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "kernel": kern,
                     "kernel_ms": round(stage_ms[kern], 4),
                     "algorithmic_bytes": ab[kern]},
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "render_ms_p50": round(float(np.percentile(lat, 50)), 3),
        "render_ms_p99": round(float(np.percentile(lat, 99)), 3),
        "step_algorithmic_GBs": round(total_bytes / (ms_step * 1e-3) / 1e9, 2),
    }
    log(f"{it_s:.2f} it/s; stages {stage_ms}")
    if args.cpu_baseline_seconds > 0:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(model, cam, target, args.cpu_baseline_seconds)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


def _settings(cam, bg):
    from dmgs_amd.rasterizer import GaussianRasterizationSettings
    return GaussianRasterizationSettings(
        image_height=cam.image_height, image_width=cam.image_width, tanfovx=math.tan(cam.FoVx * 0.5),
        tanfovy=math.tan(cam.FoVy * 0.5), bg=bg, scale_modifier=1.0, viewmatrix=cam.world_view_transform,
        projmatrix=cam.full_proj_transform, sh_degree=3, campos=cam.camera_center, prefiltered=False, debug=False)


def _read_traffic(kernel):
    """HBM bytes per launch of ``kernel`` measured with rocprofv3 PMC counters
    (committed under profiles/), or None when no measurement exists."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(kernel)
    except (OSError, ValueError):
        return None


if __name__ == "__main__":
    main()
