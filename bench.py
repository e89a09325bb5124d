"""Benchmark: fwd+bwd view-iterations per second of the Gaussian rasterizer at
1M Gaussians, 800x800, on one MI355X — batched (all views in one captured
HIP graph) beside the plain one-view-per-launch pipeline.

Default workload (``--workload views``): 1M free Gaussians in FP32 — means a
60/40 mixture of the unit-ball interior and a radius-3 shell, log-scales
N(-3.5, 0.4), N(0,1)^4 quaternions (normalised by the model), sigmoid(N(0,1))
opacity, SH degree 3 (DC N(0, 0.5), rest N(0, 0.1)) — seen by 8 seeded
cameras on a radius-4 sphere looking at the origin, 800x800 each, FoV 50 deg,
white background.  One iteration = render one view + 0.8 L1 + 0.2 (1 - SSIM)
against a seeded target + backward into all raw parameters.

* plain:   the drop-in ``render`` + fused loss + autograd backward, one view
           per call (host read of the instance count every view);
* batched: ``dmgs_amd.batched`` — every view in one launch sequence, summed
           loss, captured once in a HIP graph and replayed.

``--workload uneven``: 3M Gaussians (shell radius 5), 16 orbit cameras at
radius 4 with resolutions drawn from {512^2, 800^2, 1280x720, 1600x900,
1920x1080}.

Prints ONE JSON line.  ``value`` is the batched throughput in view
iterations per second; ``plain`` holds the one-view-per-launch figure on the
same views.  ``roofline`` is for the dominant kernel of the batched step,
timed with device events on the launching stream over eager steps of the
same workload; ``cpu_baseline`` times the FP64 CPU oracle on a bounded
sample of one 800x800 view.
"""
import argparse
import gc
import json
import math
import os
import sys
import time

import numpy as np
import torch

HBM_PEAK_GBS = 8000.0
METRIC = "fwd+bwd it/s @1M Gaussians 800² vs plain; render p50/p99 ms; HBM GB/s"
UNEVEN_SIZES = [(512, 512), (800, 800), (1280, 720), (1600, 900), (1920, 1080)]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=("views", "uneven"), default="views")
    ap.add_argument("--gaussians", type=int, default=None, help="default 1M (views) / 3M (uneven)")
    ap.add_argument("--views", type=int, default=None, help="default 8 (views) / 16 (uneven)")
    ap.add_argument("--width", type=int, default=800)
    ap.add_argument("--height", type=int, default=800)
    ap.add_argument("--cpu-baseline-seconds", type=float, default=20.0,
                    help="time budget of the sampled CPU-oracle baseline (0 disables)")
    ap.add_argument("--render-samples", type=int, default=50, help="forward-only renders for p50/p99")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain one-view-per-launch comparison")
    ap.add_argument("--serving-requests", type=int, default=200,
                    help="forward render requests for the serving figures (0 disables)")
    ap.add_argument("--serving-rate", type=float, default=200.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--pair-tiles", type=int, default=64,
                    help="tiles of view 0 on which the pair evaluations are counted")
    ap.add_argument("--uneven-steps", type=int, default=8,
                    help="graph replays of the uneven 3M / 16-view workload (0 disables; default workload only)")
    ap.add_argument("--probe", action="store_true",
                    help="counter-collection mode: a few eager batched steps plus a 256 MiB device copy of known "
                         "size (the scale for FETCH_SIZE / WRITE_SIZE), then exit; run under rocprofv3 --pmc")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _time_ms(fn, reps=20, warmup=3):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        t.append(a.elapsed_time(b))
    return float(np.median(t))


def binding_run(dev):
    """Mesh -> Gaussian binding at the sizes of the reference's only
    published timings (proto&effi_stage2.ipynb:74-76,161: 98 528 faces x 6 =
    591 168 Gaussians, stage-2 forward pieces 1.25 + 0.28 + 0.21 + 1.25 ms on
    an unstated NVIDIA GPU) and at the mesh-bound workload (icosphere of
    81 920 faces, 6 Dirichlet-barycentric Gaussians per face, stage 3)."""
    from dmgs_amd import binding as PB
    from dmgs_amd import scenes as S
    v7, f7 = S.icosphere(7)
    faces = f7[:98528].to(dev)
    verts = v7.to(dev).requires_grad_(True)
    bc, rad = PB.barycentric_v2(6, device=dev)
    sf = torch.tensor([0.4], device=dev, requires_grad=True)

    def s2_fwd():
        with torch.no_grad():
            PB.bind_stage2(verts, faces, bc, sf, rad, spatial_lr_scale=2.5)

    def s2_fwd_bwd():
        xyz, cov, _ = PB.bind_stage2(verts, faces, bc, sf, rad, spatial_lr_scale=2.5)
        torch.autograd.backward([xyz, cov], [gx, gc])

    with torch.no_grad():
        xyz0, cov0, _ = PB.bind_stage2(verts, faces, bc, sf, rad, spatial_lr_scale=2.5)
    gx, gc = torch.ones_like(xyz0), torch.ones_like(cov0)
    m = PB.mesh_scene(subdivisions=6, seed=0, device=dev).requires_grad_(True)

    def s3_fwd():
        with torch.no_grad():
            m.renew_gaussian()

    def s3_fwd_bwd():
        m.renew_gaussian()
        torch.autograd.backward([m._xyz, m._rot, m._scale], [torch.ones_like(m._xyz), torch.ones_like(m._rot),
                                                             torch.ones_like(m._scale)])

    return {"stage2_faces": int(faces.shape[0]), "stage2_gaussians": int(xyz0.shape[0]),
            "stage2_fwd_ms": round(_time_ms(s2_fwd), 4), "stage2_fwd_bwd_ms": round(_time_ms(s2_fwd_bwd), 4),
            "reference_stage2_fwd_ms": 2.99,
            "reference_note": "sum of proto&effi_stage2.ipynb:74-76,161 (face frame + means, M, L, Sigma + "
                              "expand; unstated NVIDIA GPU, PyTorch ops)",
            "stage3_faces": int(m.faces.shape[0]), "stage3_gaussians": int(m.bc_coords.shape[0]),
            "stage3_fwd_ms": round(_time_ms(s3_fwd), 4), "stage3_fwd_bwd_ms": round(_time_ms(s3_fwd_bwd), 4)}


class StageTimer:
    """Device events around named stages, recorded on the current stream."""

    def __init__(self):
        self.events = {}

    def __call__(self, name):
        timer = self

        class _Ctx:
            def __enter__(self):
                self.a = torch.cuda.Event(enable_timing=True)
                self.a.record()
                return self

            def __exit__(self, *exc):
                b = torch.cuda.Event(enable_timing=True)
                b.record()
                timer.events.setdefault(name, []).append((self.a, b))
                return False
        return _Ctx()

    def averages_ms(self):
        return {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in self.events.items()}


def algorithmic_bytes(P, V, K, X, T):
    """Compulsory HBM bytes per stage (each tensor read or written once).
    P Gaussian entries, V visible entries, K instances, X pixels, T tiles,
    each summed over the views of one launch."""
    return {
        "geometry": 236 * P + 4 * P + 48 * V + 8 * P,
        "binning": 20 * V + 12 * K + 24 * K + 8 * K + 8 * T,
        "render_forward": 8 * T + 40 * K + 20 * X,
        "render_backward": 8 * T + 40 * K + 20 * X + 36 * V,
        "geometry_backward": 276 * V + 248 * P,
    }


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def build_workload(args, dev):
    from dmgs_amd import cameras, scenes
    if args.workload == "views":
        N = args.gaussians or 1_000_000
        B = args.views or 8
        model = scenes.free_scene(N, seed=0, layout="mixture", shell_radius=3.0, device=dev)
        pos = cameras.sphere_positions(B, 4.0, seed=2)
        sizes = [(args.width, args.height)] * B
        desc = (f"{N} free Gaussians (60/40 unit ball / r=3 shell), {B} orbit views (r=4) of "
                f"{args.width}x{args.height}, fwd+bwd of 0.8 L1 + 0.2 (1-SSIM) per view")
    else:
        N = args.gaussians or 3_000_000
        B = args.views or 16
        model = scenes.free_scene(N, seed=0, layout="mixture", shell_radius=5.0, device=dev)
        pos = cameras.sphere_positions(B, 4.0, seed=2)
        g = np.random.default_rng(2)
        sizes = [UNEVEN_SIZES[i] for i in g.integers(0, len(UNEVEN_SIZES), size=B)]
        desc = (f"{N} free Gaussians (60/40 unit ball / r=5 shell), {B} orbit views (r=4), resolutions drawn "
                f"from 512^2..1920x1080, fwd+bwd of summed 0.8 L1 + 0.2 (1-SSIM)")
    cams = [cameras.camera_at(p, w, h, fov_deg=50.0, device=dev) for p, (w, h) in zip(pos, sizes)]
    targets = [scenes.target_image(h, w, seed=1 + i, device=dev) for i, (w, h) in enumerate(sizes)]
    return model, cams, targets, sizes, desc


def run_plain(model, cams, targets, bg, steps, warmup):
    """Drop-in render + fused loss + autograd backward, one view per call."""
    from dmgs_amd import rasterizer as RZ
    from dmgs_amd.gaussian_renderer import PipelineParams, render
    from dmgs_amd.loss import fused_dssim_l1
    model.requires_grad_(True)
    pipe = PipelineParams()

    def step():
        for p in model.parameters():
            p.grad = None
        for cam, tgt in zip(cams, targets):
            out = render(cam, model, pipe, bg)
            fused_dssim_l1(out["render"], tgt, 0.2).backward()

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    timer = StageTimer()
    RZ.set_stage_hook(timer)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    RZ.set_stage_hook(None)
    model.requires_grad_(False)
    B = len(cams)
    return {"value": round(B * steps / elapsed, 3), "unit": "it/s", "ms_per_view": round(1000 * elapsed / (B * steps), 4),
            "stage_ms": {k: round(v, 4) for k, v in timer.averages_ms().items()}}


def serving_run(n_requests, rate, sweep_rates=(100.0, 400.0, 1600.0, 6400.0)):
    """Four resident scenes (0.5M / 1M / 2M / 3M Gaussians, same
    distributions), Poisson arrivals at ``rate``, random scene, camera on a
    sphere of radius U[3, 6], resolution drawn from 256^2 .. 1920x1080;
    continuous batching over 4 streams with HBM-budget admission."""
    from dmgs_amd import batched as BT
    from dmgs_amd import scenes as S
    from dmgs_amd import serving as SV
    scenes = {}
    for i, n in enumerate((500_000, 1_000_000, 2_000_000, 3_000_000)):
        scenes[f"s{n}"] = BT.pack_model(S.free_scene(n, seed=10 + i, layout="mixture", shell_radius=3.0,
                                                     device="cuda"))
    names = sorted(scenes)
    srv = SV.RenderServer(scenes, max_views=8, num_streams=4)
    SV.serve_once(srv, SV.poisson_requests(names, 16, None, seed=99))  # warm-up
    # warm-up at a moderate rate: the first batches of each (scene, size class) time both blend shapes
    SV.serve_once(srv, SV.poisson_requests(names, 150, 400.0, seed=98))
    gc.collect()
    gc.freeze()
    stats = SV.serve_once(srv, SV.poisson_requests(names, n_requests, rate, seed=3))
    stats.update(rate_per_s=rate, requests=n_requests, budget_GB=round(srv.budget / 1e9, 1),
                 peak_reserved_GB=round(srv.peak_reserved / 1e9, 3), max_in_flight=srv.max_active)
    sweep = []
    for r in sweep_rates:  # the same request mix at other Poisson rates
        st = SV.serve_once(srv, SV.poisson_requests(names, n_requests, r, seed=int(r)))
        sweep.append({k: st[k] for k in ("p50_ms", "p99_ms", "p50_ms_1080p", "p99_ms_1080p", "renders_per_s",
                                         "mean_batch") if k in st})
        sweep[-1]["rate_per_s"] = r
    stats["sweep"] = sweep
    # capacity: the same number of requests all arriving at once (the queue
    # is never empty, so renders/s is what the server sustains)
    st = SV.serve_once(srv, SV.poisson_requests(names, n_requests, None, seed=77))
    stats["capacity"] = {k: st[k] for k in ("p50_ms", "p99_ms", "renders_per_s", "mean_batch") if k in st}
    stats["latency_clock"] = "from each request's scheduled arrival"
    return stats


def plain_latency(model, cam, bg, samples):
    from dmgs_amd.gaussian_renderer import PipelineParams, render
    pipe = PipelineParams()
    lat = []
    with torch.no_grad():
        for _ in range(3):  # warm the allocator's pools for this path first
            render(cam, model, pipe, bg)
        for _ in range(samples):
            torch.cuda.synchronize()
            a = time.perf_counter()
            render(cam, model, pipe, bg)
            torch.cuda.synchronize()
            lat.append(1000.0 * (time.perf_counter() - a))
    return lat


def cpu_baseline(model, cam, budget_s):
    """FP64 CPU oracle on a bounded sample of one view: full preprocess,
    binning and geometry backward; blend forward+backward on random tiles
    until the budget is spent, extrapolated by instance count."""
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
    return {"value": round(1.0 / total, 6), "unit": "it/s", "cores": threads, "kind": "port",
            "sample": (f"FP64 torch oracle, one {W}x{H} view of {P} Gaussians: preprocess {t1 - t0:.2f}s, binning "
                       f"{t2 - t1:.2f}s and geometry backward {t_gb:.2f}s over all Gaussians; blend fwd+bwd on "
                       f"{n_tiles} random tiles ({done_k} of {K} instances, {t_tiles:.1f}s) extrapolated by "
                       f"instance count to {t_blend:.1f}s; loss excluded")}


def cpu_small_view(repeats=5):
    """FP64 CPU oracle, full forward + backward of the small workload (10k
    free Gaussians, one 200x200 view, L1 against a seeded target): median of
    ``repeats`` runs."""
    from dmgs_amd import cameras, scenes
    from oracle import raster as OR
    m = scenes.free_scene(10_000, seed=0, layout="uniform")
    cam = cameras.front_camera(200, 200, device="cpu")
    st = OR.settings_for(cam, torch.ones(3), 3)
    target = scenes.target_image(200, 200, seed=1).double()
    kw = dict(shs=m.get_features, scales=m.get_scaling, rotations=m.get_rotation)
    times = []
    for _ in range(repeats):
        a = time.perf_counter()
        col, _, state = OR.rasterize(m.get_xyz, m.get_opacity, st, **kw)
        dL = torch.sign(col - target) / col.numel()
        OR.rasterize_backward(state, st, m.get_xyz, dL, **kw)
        times.append(time.perf_counter() - a)
    t = float(np.median(times))
    return {"value": round(1.0 / t, 4), "unit": "it/s", "seconds_per_it": round(t, 3), "cores": torch.get_num_threads(),
            "kind": "port", "sample": f"10000 Gaussians, one 200x200 view, full fwd+bwd of L1, median of {repeats}"}


def cpu_mesh_view():
    """FP64 CPU oracle, one full forward + backward of the mesh-bound
    workload (81 920 faces x 6 Gaussians, one 800x800 view) into the vertices."""
    from dmgs_amd import binding as PB
    from dmgs_amd import cameras
    from oracle import binding as OB
    from oracle import raster as OR
    gm = PB.mesh_scene(subdivisions=6, seed=0)
    st = OR.settings_for(cameras.front_camera(800, 800, device="cpu"), torch.ones(3), 3)
    feats = torch.cat([gm._features_dc, gm._features_rest], 1)
    a = time.perf_counter()
    v = gm.verts.double().requires_grad_(True)
    xyz, q, s, _ = OB.stage3(v, gm.faces, gm.bc_coords.double(), gm._rotation.double(), gm._scaling.double(),
                             gm.thin_z_scale)
    kw = dict(shs=feats, scales=s.detach(), rotations=q.detach())
    col, _, state = OR.rasterize(xyz.detach(), torch.sigmoid(gm._opacity), st, **kw)
    dL = torch.full_like(col, 1.0 / col.numel())
    g = OR.rasterize_backward(state, st, xyz.detach(), dL, **kw)
    torch.autograd.backward([xyz, q], [g[0], g[6]])
    t = time.perf_counter() - a
    return {"value": round(1.0 / t, 5), "unit": "it/s", "seconds_per_it": round(t, 2), "cores": torch.get_num_threads(),
            "kind": "port", "sample": f"{xyz.shape[0]} mesh-bound Gaussians, one 800x800 view, one full fwd+bwd "
                                      "through the binding into the vertices"}


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


def _read_valu(kernel):
    """Counter figures of ``kernel`` at the default workload, committed under
    profiles/: wave64 VALU instructions per launch (SQ_INSTS_VALU), VALU
    busy quad-cycles per launch (SQ_ACTIVE_INST_VALU) and the launch's
    duration in that counter run; or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "valu.json")
    try:
        with open(path) as f:
            v = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    return v if isinstance(v, dict) else None


def _valu_issue(kernel, ms, args):
    """VALU issue of ``kernel`` on MI355X (256 CUs x 4 SIMDs, one wave64
    instruction per SIMD per 4 cycles at 2.4 GHz): the busy fraction from the
    committed counter run (its busy cycles over its own duration, so both
    come from one run), and the instruction rate at the live time (which
    includes the stage's two small tile-order kernels); only for the default
    workload."""
    v = _read_valu(kernel)
    if v is None or args.workload != "views" or args.gaussians or args.views or (args.width, args.height) != (800, 800):
        return None
    simds, clock = 256 * 4, 2.4e9
    peak = simds * clock / 4  # wave instructions per second
    n = v["instructions"]
    return {"instructions": n, "profiled_us": v["profiled_us"],
            "frac": round(v["active_quad_cycles"] * 4 / (simds * clock * v["profiled_us"] * 1e-6), 4),
            "achieved_per_s_live": round(n / (ms * 1e-3), 1), "peak_per_s": peak}


def _read_traffic(kernel):
    """HBM bytes per launch of ``kernel`` measured with rocprofv3 PMC counters
    (committed under profiles/), or None when no measurement exists."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(kernel)
    except (OSError, ValueError):
        return None


def probe(args, model, cams, targets, bg):
    """Eager batched steps (every kernel of the step, one dispatch each per
    step) and a copy of exactly 256 MiB: under rocprofv3 --pmc, the copy
    kernel's FETCH_SIZE / WRITE_SIZE per dispatch against its 256 MiB gives
    the factor that turns the counters into bytes."""
    from dmgs_amd import batched as BT
    params = BT.pack_model(model)
    r = BT.BatchedRenderer(cams, bg, model.num, sh_degree=3, variant=4)
    r.plan(params)
    step = BT.BatchedStep(r, BT.BatchedLoss(r, targets, 0.2), params)
    for _ in range(max(1, args.steps)):
        step.run(check=False)
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"probe done: {args.steps} eager steps, instances {int(r.instance_count().item())}, copy 268435456 B x3")


def measure_batched(model, cams, targets, bg, steps, warmup, autotune=True):
    """Batched fwd+bwd of every view in one HIP graph replay per step:
    throughput, stage times (device events on the launching stream over
    eager steps of the same work) and the step's sizes."""
    from dmgs_amd import batched as BT
    params = BT.pack_model(model)
    r = BT.BatchedRenderer(cams, bg, model.num, sh_degree=3)
    K_cap = r.plan(params)
    loss = BT.BatchedLoss(r, targets, 0.2)
    step = BT.BatchedStep(r, loss, params)
    tune = r.autotune(params, reps=2) if autotune else {}
    log(f"planned K_cap={K_cap}; autotune ms {tune} -> variant {r.variant}")
    step.capture()
    for _ in range(warmup):
        step.replay()
    torch.cuda.synchronize()
    log(f"timing {steps} batched steps ({len(cams)} views each)")
    t0 = time.perf_counter()
    for _ in range(steps):
        step.replay(check=False)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if step.check() or r.overflowed():
        raise RuntimeError("instance capacity overflowed during the timed region")
    timer = StageTimer()
    for _ in range(max(3, steps // 2)):
        with timer("geometry+binning"):
            r.forward(params, step.images, stages=BT.STAGE_GEOMETRY | BT.STAGE_BINNING)
        with timer("render_forward"):
            r.forward(params, step.images, stages=BT.STAGE_BLEND)
        with timer("loss"):
            loss.forward(step.images)
            loss.backward(step.images, step.dimages)
        with timer("render_backward"):
            r.backward(params, step.dimages, step.grads, stages=BT.STAGE_BLEND)
        with timer("geometry_backward"):
            r.backward(params, step.dimages, step.grads, stages=BT.STAGE_GEOMETRY)
    torch.cuda.synchronize()
    K_cell, K_tile = list_counts(r)
    return dict(r=r, params=params, step=step, loss=loss, tune=tune, K_cap=K_cap, elapsed=elapsed,
                it_s=len(cams) * steps / elapsed, ms_step=1000.0 * elapsed / steps, stage_ms=timer.averages_ms(),
                K=K_cell, K_tile=K_tile, V=int((r.radii() > 0).sum().item()), X=r.total_pixels, T=r.total_tiles)


def list_counts(r):
    """(cell instances, tile instances) of the last forward: a cell
    instance's tile mask holds one bit per (splat, tile) pair of the
    reference's per-tile lists."""
    import ctypes
    from dmgs_amd import _native as N
    ko, vo = N.c_size_t(0), N.c_size_t(0)
    N.check(r.lib.dmgs_batched_list_offsets(ctypes.byref(r._desc()), ctypes.byref(ko), ctypes.byref(vo)), "offsets")
    K = min(int(r.instance_count().item()) & 0xFFFFFFFF, r.K_cap)
    mask = (r.lists[ko.value:ko.value + 4 * K].view(torch.int32) >> 16) & 0xFFFF
    bits = torch.zeros_like(mask)
    for b in range(16):
        bits += (mask >> b) & 1
    return K, int(bits.sum().item())


def _reach_strips(cx, cy, a, b, c, o, x0, y0):
    """Which of the four 16x4 pixel strips of the tile at (x0, y0) each
    splat's alpha >= 1/255 ellipse reaches, by the blend kernels' staging
    test (same margins); (4, L) bool."""
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        tau = 2.0 * np.log(255.0 * o)
    t = tau * 1.001 + 0.001
    out = np.zeros((4, cx.size), dtype=bool)
    m = 0.01
    for k in range(4):
        rx0, rx1 = x0 - m, x0 + 15.0 + m
        ry0, ry1 = y0 + 4.0 * k - m, y0 + 4.0 * k + 3.0 + m
        inside = (cx >= rx0) & (cx <= rx1) & (cy >= ry0) & (cy <= ry1)
        best = np.full(cx.size, np.inf)
        with np.errstate(divide="ignore", invalid="ignore"):
            for xe in (rx0, rx1):
                dx = xe - cx
                dy = np.clip(-b * dx / c, ry0 - cy, ry1 - cy)
                best = np.minimum(best, a * dx * dx + 2 * b * dx * dy + c * dy * dy)
            for ye in (ry0, ry1):
                dy = ye - cy
                dx = np.clip(-b * dy / a, rx0 - cx, rx1 - cx)
                best = np.minimum(best, a * dx * dx + 2 * b * dx * dy + c * dy * dy)
        out[k] = (tau > 0) & (inside | (best <= t) | ~((a > 0) & (c > 0)))
    return out


def pair_counts(r, view, n_tiles, seed=0):
    """(pixel, splat) pair evaluations on sampled tiles of one view of the
    last batched forward, from the device's own splat records and lists
    (float64 restatement of the per-pixel rules):
      reference:        a per-pixel loop, every list entry up to and including
                        the pixel's stopping entry (the reference kernel's work);
      strips_forward:   the one-wave-per-tile forward, 64 lanes per (splat,
                        16x4 strip) the splat's footprint reaches while a pixel
                        of the strip still blends;
      strips_backward:  the strip backward, 64 lanes per (splat, reached strip)
                        in front of the strip's last contributor;
      contributing:     pairs that pass the alpha test before the pixel stops.
    Also the sample's tile instances (list entries)."""
    import ctypes
    from dmgs_amd import _native as N
    d = r.views[view]
    W, H, gx, gy = d.width, d.height, d.grid_x, d.grid_y
    cgx = (gx + 3) // 4
    P = r.P
    o5 = r._offsets[5]
    rec = r.state[o5 + 48 * view * P:o5 + 48 * (view + 1) * P].view(torch.float32).view(P, 12).cpu().numpy()
    rec = rec.astype(np.float64)
    ko, vo = N.c_size_t(0), N.c_size_t(0)
    N.check(r.lib.dmgs_batched_list_offsets(ctypes.byref(r._desc()), ctypes.byref(ko), ctypes.byref(vo)), "offsets")
    o2 = r._offsets[2]
    rng = np.random.default_rng(seed)
    counts = dict(reference=0, strips_forward=0, strips_backward=0, contributing=0, tile_instances=0, tiles=0)
    for t in rng.permutation(gx * gy)[:n_tiles]:
        tx, ty = int(t) % gx, int(t) // gx
        cell = d.cell_base + (ty // 4) * cgx + tx // 4
        s0, s1 = r.state[o2 + 8 * cell:o2 + 8 * cell + 8].view(torch.int32).cpu().tolist()
        keys = r.lists[ko.value + 4 * s0:ko.value + 4 * s1].view(torch.int32).cpu().numpy().view(np.uint32)
        ids = r.lists[vo.value + 4 * s0:vo.value + 4 * s1].view(torch.int32).cpu().numpy().view(np.uint32)
        bit = 1 << ((ty & 3) * 4 + (tx & 3))
        g = ids[((keys >> 16) & bit) != 0].astype(np.int64) - view * P
        L = g.size
        counts["tiles"] += 1
        counts["tile_instances"] += L
        if L == 0:
            continue
        cx, cy, a, b, c, o = (rec[g, i] for i in (0, 1, 2, 3, 4, 5))
        x0, y0 = 16.0 * tx, 16.0 * ty
        reach = _reach_strips(cx, cy, a, b, c, o, x0, y0)
        py, px = np.mgrid[0:16, 0:16]
        px, py = (px + x0).reshape(-1), (py + y0).reshape(-1)
        dx = cx[None, :] - px[:, None]
        dy = cy[None, :] - py[:, None]
        power = -0.5 * (a * dx * dx + c * dy * dy) - b * dx * dy
        alpha = np.minimum(0.99, o * np.exp(power))
        valid = (power <= 0) & (alpha >= 1.0 / 255.0)
        Tinc = np.cumprod(np.where(valid, 1.0 - alpha, 1.0), axis=1)
        stop = valid & (Tinc < 1e-4)
        sidx = np.where(stop.any(1), stop.argmax(1), L)
        j = np.arange(L)[None, :]
        include = valid & (j < sidx[:, None])
        ncontrib = np.where(include, j + 1, 0).max(1)
        inside = (px < W) & (py < H)
        counts["reference"] += int(np.minimum(sidx + 1, L)[inside].sum())
        counts["contributing"] += int(include[inside].sum())
        strip = (py - int(y0)) // 4
        for k in range(4):
            sel = inside & (strip == k)
            if not sel.any():
                continue
            alive_to = sidx[sel].max()          # the strip blends up to its last stopping entry
            counts["strips_forward"] += 64 * int((reach[k] & (np.arange(L) <= alive_to)).sum())
            behind = ncontrib[sel].max()       # entries in front of the strip's last contributor
            counts["strips_backward"] += 64 * int((reach[k] & (np.arange(L) < behind)).sum())
    return counts


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the rasterizer path here is "
              "single-GPU and no multi-GPU figure is produced.", file=sys.stderr)
        sys.exit(2)
    from dmgs_amd import batched as BT

    dev = "cuda"
    torch.cuda.set_device(0)
    model, cams, targets, sizes, desc = build_workload(args, dev)
    bg = torch.ones(3, device=dev)
    B, N = len(cams), model.num
    log(f"workload ready: {desc}")
    if args.probe:
        probe(args, model, cams, targets, bg)
        return

    m = measure_batched(model, cams, targets, bg, args.steps, args.warmup)
    r, params, step, loss = m["r"], m["params"], m["step"], m["loss"]
    it_s, ms_step, stage_ms = m["it_s"], m["ms_step"], m["stage_ms"]
    K, V, X, T = m["K"], m["V"], m["X"], m["T"]
    ab = algorithmic_bytes(B * N, V, K, X, T)
    kern = max(("render_forward", "render_backward"), key=lambda k: stage_ms[k])
    achieved = ab[kern] / (stage_ms[kern] * 1e-3) / 1e9
    step_bytes = ab["geometry"] + ab["binning"] + ab["render_forward"] + ab["render_backward"] + \
        ab["geometry_backward"]
    losses = loss.losses[:, 0].sum().item()
    # pair evaluations on sampled tiles of the first view, scaled to the step by tile instances
    pc = pair_counts(r, 0, args.pair_tiles, seed=1)
    scale = m["K_tile"] / max(pc["tile_instances"], 1)
    pairs = {k: int(round(pc[k] * scale)) for k in ("reference", "strips_forward", "strips_backward", "contributing")}
    valu = _valu_issue(f"batched_{kern}", stage_ms[kern], args)
    pairs_note = {"sample": f"{pc['tiles']} random tiles of view 0 ({pc['tile_instances']} tile instances), "
                            f"scaled by the step's {m['K_tile']} tile instances",
                  "forward_evaluated_per_contributing": round(pc["strips_forward"] / max(pc["contributing"], 1), 3),
                  "backward_evaluated_per_contributing": round(pc["strips_backward"] / max(pc["contributing"], 1), 3),
                  "reference_evaluated_per_contributing": round(pc["reference"] / max(pc["contributing"], 1), 3)}
    if valu:
        pairs_note["backward_valu_per_64_evaluated"] = round(valu["instructions"] / (pairs["strips_backward"] / 64), 2)

    # ---- forward-only latency: batch, and the plain single view ----
    # the objects made so far move to the permanent GC generation, so a full
    # collection of them cannot land inside a timed render (as in a server)
    gc.collect()
    gc.freeze()
    blat = []
    for _ in range(max(5, args.render_samples // 5)):
        torch.cuda.synchronize()
        a = time.perf_counter()
        r.forward(params, step.images)
        torch.cuda.synchronize()
        blat.append(1000.0 * (time.perf_counter() - a))
    plat = plain_latency(model, cams[0], bg, args.render_samples)
    # one view through the batched engine (resident packed parameters), forward only
    r1 = BT.BatchedRenderer(cams[:1], bg, model.num, sh_degree=3, variant=3)
    r1.plan(params, backward=False)
    img1 = r1.new_images()
    vlat = []
    for k in range(3 + args.render_samples):
        torch.cuda.synchronize()
        a = time.perf_counter()
        r1.forward(params, img1)
        torch.cuda.synchronize()
        if k >= 3:
            vlat.append(1000.0 * (time.perf_counter() - a))
    del r1, img1

    plain = None
    if not args.no_plain:
        log("timing the plain one-view-per-launch path on the same views")
        plain = run_plain(model, cams, targets, bg, args.steps, args.warmup)

    result = {
        "metric": METRIC,
        "value": round(it_s, 3),
        "unit": "it/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc, "gaussians": N, "views": B, "sizes": [f"{w}x{h}" for w, h in sizes],
                   "path": "batched",
                   "step": f"{B} views fwd+bwd in one HIP graph replay; value = views/s",
                   "visible_entries": V, "instances": K, "tile_instances": m["K_tile"],
                   "instance_capacity": m["K_cap"], "render_variant": r.variant, "summed_loss": round(losses, 6)},
        "plain": plain,
        "vs_plain": round(it_s / plain["value"], 3) if plain else None,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": _read_traffic(f"batched_{kern}"),
                     "kernel": kern, "kernel_ms": round(stage_ms[kern], 4), "algorithmic_bytes": ab[kern],
                     "K_cell": K, "K_tile": m["K_tile"],
                     "note": "algorithmic bytes = 8T + 40K + 20X (+36V backward) with K = K_cell, the (splat, "
                             "64x64-pixel cell) instances the lists hold; with K = K_tile, the reference's "
                             "(splat, 16x16 tile) pairs, the figure would be ~5.7x higher than the bytes that "
                             "exist to move.  Measured traffic (rocprofv3 FETCH/WRITE, copy-calibrated) is below "
                             "the algorithmic bytes: the blend kernels are bound by VALU issue",
                     "valu_issue": valu,
                     "pairs": dict(pairs, **pairs_note)},
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "autotune_ms": {str(k): round(v, 4) for k, v in m["tune"].items()},
        "render_ms_p50": round(float(np.percentile(plat, 50)), 3),
        "render_ms_p99": round(float(np.percentile(plat, 99)), 3),
        "batch_render_ms_p50": round(float(np.percentile(blat, 50)), 3),
        "batch_render_ms_p99": round(float(np.percentile(blat, 99)), 3),
        "view_render_ms_p50": round(float(np.percentile(vlat, 50)), 3),
        "view_render_ms_p99": round(float(np.percentile(vlat, 99)), 3),
        "step_algorithmic_GBs": round(step_bytes / (ms_step * 1e-3) / 1e9, 2),
    }
    log(f"batched {it_s:.2f} it/s; plain {plain['value'] if plain else 'skipped'}; stages {stage_ms}")
    log(f"pairs {result['roofline']['pairs']}")
    del step, loss, r, m
    torch.cuda.empty_cache()
    if args.workload == "views" and args.uneven_steps > 0:
        result["uneven"] = uneven_run(args, dev, bg)
        log(f"uneven {result['uneven']}")
    result["binding"] = binding_run(dev)
    log(f"binding {result['binding']}")
    if args.serving_requests > 0:
        log(f"serving: {args.serving_requests} requests at {args.serving_rate}/s over 4 resident scenes")
        torch.cuda.empty_cache()
        result["serving"] = serving_run(args.serving_requests, args.serving_rate)
        log(f"serving {result['serving']}")
    if args.cpu_baseline_seconds > 0:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(model, cams[0], args.cpu_baseline_seconds)
        log(f"cpu baseline (1M, 800x800, sampled) {result['cpu_baseline']}")
        result["cpu_baseline"]["small_view"] = cpu_small_view()
        log(f"cpu baseline (10k, 200x200, full) {result['cpu_baseline']['small_view']}")
        result["cpu_baseline"]["mesh_view"] = cpu_mesh_view()
        log(f"cpu baseline (mesh-bound, 800x800, full) {result['cpu_baseline']['mesh_view']}")
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


def uneven_run(args, dev, bg):
    """The uneven multi-resolution workload (3M Gaussians, 16 orbit views of
    512^2 .. 1920x1080) as one HIP graph per step, beside its plain
    one-view-per-launch figure."""
    import copy
    ua = copy.copy(args)
    ua.workload, ua.gaussians, ua.views = "uneven", None, None
    model, cams, targets, sizes, desc = build_workload(ua, dev)
    log(f"uneven workload ready: {desc}")
    m = measure_batched(model, cams, targets, bg, args.uneven_steps, 2)
    out = {"value": round(m["it_s"], 3), "unit": "it/s", "ms_per_step": round(m["ms_step"], 3),
           "steps": args.uneven_steps, "workload": desc, "sizes": [f"{w}x{h}" for w, h in sizes],
           "instances": m["K"], "tile_instances": m["K_tile"], "visible_entries": m["V"],
           "render_variant": m["r"].variant, "stage_ms": {k: round(v, 3) for k, v in m["stage_ms"].items()}}
    del m
    torch.cuda.empty_cache()
    if not args.no_plain:
        plain = run_plain(model, cams, targets, bg, max(2, args.uneven_steps // 4), 1)
        out["plain"] = plain
        out["vs_plain"] = round(out["value"] / plain["value"], 3)
    del model, cams, targets
    torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    main()
