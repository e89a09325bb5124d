"""Render-request server: resident scenes, HBM-budget admission, continuous
batching over several HIP streams, latency statistics.

The reference serves one viewer at a time, blocking: gaussian_renderer/
network_gui.py:26-85 receives a camera as length-prefixed JSON, renders it
with ``render`` and sends raw RGB bytes back before the next request is read.
Here requests are admitted while others are on the GPU:

* scenes stay resident in HBM as component-major parameter tensors (the
  batched pipeline's layout, ``dmgs_amd.batched``);
* pending requests of one scene are grouped into a multi-view batch (up to
  ``max_views``; views may differ in resolution) when a stream slot is free
  and the batch's memory fits the free part of the HBM budget;
* each batch runs on its own stream in two phases — (1) preprocess + depth
  sort + scan, then an asynchronous read of the instance count; (2) once the
  count is on the host, lists sized exactly, binning and blend — so several
  batches' kernels overlap on the device and the host never blocks on one;
* memory is reserved at admission (state + images exact, lists estimated
  from the scene's observed instances per pixel) and corrected to the exact
  size after the count; a batch that cannot fit waits for memory to free
  up, and a request that can never fit the budget is rejected.

Every image is computed by the same per-view arithmetic whether a request
runs alone or batched with others (the batched pipeline gives each view the
single-view result), so results do not depend on the mix.
"""
import collections
import itertools
import statistics
import time

import numpy as np
import torch

from . import batched as BT
from . import cameras as C


class Request:
    """One forward render: scene name, camera (reference Camera fields) and
    background, optionally a timeout in seconds from submission.  Filled in
    by the server: ``image`` (3, H, W), ``status`` ('done', 'rejected',
    'cancelled' or 'timed_out'; only 'done' carries an image), timing and the
    batch it ran in."""

    _ids = itertools.count()

    def __init__(self, scene, camera, bg=(1.0, 1.0, 1.0), arrival=None, scaling_modifier=1.0, timeout_s=None):
        self.id = next(self._ids)
        self.scene, self.camera = scene, camera
        self.scaling_modifier = float(scaling_modifier)
        self.bg = torch.as_tensor(bg, dtype=torch.float32)
        self.arrival = arrival
        self.timeout_s = None if timeout_s is None else float(timeout_s)
        self.image = None
        self.status = "new"
        self.error = None
        self.t_submit = self.t_done = None
        self.batch_size = 0
        self.cancel_requested = False

    @property
    def deadline(self):
        if self.timeout_s is None or self.t_submit is None:
            return None
        return self.t_submit + self.timeout_s

    @property
    def latency_ms(self):
        return None if self.t_done is None else 1000.0 * (self.t_done - self.t_submit)

    @property
    def resolution(self):
        return (int(self.camera.image_width), int(self.camera.image_height))


class _Batch:
    """A multi-view batch in flight on one stream."""

    def __init__(self, server, scene, requests, stream):
        self.server, self.scene, self.requests, self.stream = server, scene, requests, stream
        params, sh = server.scenes[scene]
        self.params = params
        with torch.cuda.stream(stream):
            self.r = BT.BatchedRenderer([q.camera for q in requests], [q.bg for q in requests], params.shape[1],
                                        sh_degree=sh, scale_modifier=requests[0].scaling_modifier,
                                        device=server.device)
            self.r.variant, self.vkey = server._choose_variant(scene, self.r.total_tiles)
            self.r._alloc_state()
            self.images = self.r.new_images()
            self.count = torch.empty(2, dtype=torch.int32, pin_memory=True)  # instances, status word
        self.reserved = 0
        self.K = None
        self.phase = "geometry"
        self.event = torch.cuda.Event()
        self.t_start = torch.cuda.Event(enable_timing=True)
        self.t_end = torch.cuda.Event(enable_timing=True)

    def fixed_bytes(self):
        return self.r.state.numel() + self.images.numel() * 4

    def launch_geometry(self):
        with torch.cuda.stream(self.stream):
            self.r.count_instances_async(self.params, self.count)
            self.event.record(self.stream)

    def launch_render(self):
        stages = BT.STAGE_BINNING | BT.STAGE_BLEND
        if self.r.cell_mode == 1 and int(self.count[1]) & 1:
            # a depth the view-major keys cannot order: redo the geometry with global cell lists
            self.r.cell_mode = 0
            stages = BT.STAGES_ALL
        with torch.cuda.stream(self.stream):
            self.r.plan(self.params, K_cap=max(self.K, 1), backward=False)
            self.t_start.record(self.stream)
            self.r.forward(self.params, self.images, stages=stages)
            self.t_end.record(self.stream)
            self.event.record(self.stream)
        self.phase = "render"


# Blend shapes the server chooses between when no variant is fixed: quarter
# tiles of 8x8 pixels (fastest for small views and short lists) and one wave
# per 16x16 tile (fastest for large views and long lists).  All shapes give
# bit-identical images, so the choice changes only the time.
AUTO_VARIANTS = (3, 4)


class RenderServer:
    """Continuous-batching render server over resident scenes.

    scenes: {name: params (59, P) float32 on ``device``} or {name: (params,
    sh_degree)}.  hbm_budget: bytes the server may hold for in-flight work
    (default 80 % of the free HBM at start-up).  variant: the blend shape of
    every batch, or None to time each of AUTO_VARIANTS on the first batches
    of a (scene, size class) and keep the faster (size class: the bit length
    of the batch's tile count)."""

    def __init__(self, scenes, device="cuda", hbm_budget=None, max_views=8, num_streams=4, variant=None,
                 initial_instances_per_pixel=32.0):
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.scenes = {}
        for name, v in scenes.items():
            params, sh = v if isinstance(v, tuple) else (v, 3)
            if params.shape[0] != BT.CHANNELS or params.dtype != torch.float32 or not params.is_contiguous():
                raise ValueError(f"scene {name}: params must be a contiguous float32 (59, P) tensor")
            self.scenes[name] = (params.to(self.device), int(sh))
        self.budget = int(hbm_budget if hbm_budget is not None else 0.8 * torch.cuda.mem_get_info(self.device)[0])
        self.max_views = int(max_views)
        self.variant = None if variant is None else int(variant)
        self.variant_ms = {}  # (scene, size class) -> {variant: [render ms of its batches]}
        self.streams = [torch.cuda.Stream(self.device) for _ in range(num_streams)]
        self.free_streams = collections.deque(self.streams)
        self.pending = collections.deque()
        self.active = []
        self.completed = []
        self.reserved = 0
        self.peak_reserved = 0
        self.ipp = {name: float(initial_instances_per_pixel) for name in self.scenes}
        self.batches_run = 0
        self.requeued = 0
        self.max_active = 0

    # ---- sizing -----------------------------------------------------------
    def _lists_bytes(self, K):
        return int(BT.N.load().dmgs_batched_lists_size(max(int(K), 1)))

    def _state_bytes(self, P, cams):
        lib = BT.N.load()
        tiles = sum(((c.image_width + 15) // 16) * ((c.image_height + 15) // 16) for c in cams)
        pixels = sum(c.image_width * c.image_height for c in cams)
        return int(lib.dmgs_batched_state_size(len(cams), P, tiles, pixels)) + 12 * pixels, pixels

    def footprint(self, scene, cams):
        """Estimated bytes of a batch (state + images exact, lists from the
        scene's instances-per-pixel estimate)."""
        P = self.scenes[scene][0].shape[1]
        fixed, pixels = self._state_bytes(P, cams)
        return fixed + self._lists_bytes(self.ipp[scene] * pixels)

    # ---- blend shape ------------------------------------------------------
    def _choose_variant(self, scene, tiles):
        """(variant, timing key or None) for a new batch."""
        if self.variant is not None:
            return self.variant, None
        key = (scene, int(tiles).bit_length())
        seen = self.variant_ms.setdefault(key, {v: [] for v in AUTO_VARIANTS})
        for v in AUTO_VARIANTS:  # two samples of each first
            if len(seen[v]) < 2:
                return v, key
        return min(AUTO_VARIANTS, key=lambda v: statistics.median(seen[v][-8:])), key

    # ---- request flow -----------------------------------------------------
    def submit(self, req):
        req.t_submit = time.perf_counter()
        if req.scene not in self.scenes:
            req.status, req.error = "rejected", f"unknown scene {req.scene!r}"
            self.completed.append(req)
            return req
        need = self.footprint(req.scene, [req.camera])
        if need > self.budget:
            req.status, req.error = "rejected", f"needs ~{need} B of HBM, budget is {self.budget} B"
            self.completed.append(req)
            return req
        req.status = "pending"
        self.pending.append(req)
        return req

    def _form_batch(self):
        head = self.pending[0]
        group = [head]
        for q in itertools.islice(self.pending, 1, None):
            if len(group) >= self.max_views:
                break
            if q.scene == head.scene and q.scaling_modifier == head.scaling_modifier:
                group.append(q)
        # shrink until the batch fits the free budget (the head alone always fits an empty server)
        while group:
            need = self.footprint(head.scene, [q.camera for q in group])
            if self.reserved + need <= self.budget:
                return group, need
            if len(group) == 1:
                return None, need
            group = group[:max(1, len(group) // 2)]
        return None, 0

    def _admit(self):
        while self.pending and self.free_streams:
            group, need = self._form_batch()
            if group is None:
                if not self.active:  # cannot fit even an idle server
                    q = self.pending.popleft()
                    q.status, q.error = "rejected", f"needs ~{need} B of HBM, budget is {self.budget} B"
                    q.t_done = time.perf_counter()
                    self.completed.append(q)
                    continue
                break  # wait for memory
            for q in group:
                self.pending.remove(q)
                q.status = "running"
                q.batch_size = len(group)
            stream = self.free_streams.popleft()
            b = _Batch(self, group[0].scene, group, stream)
            b.reserved = need
            self.reserved += need
            self.peak_reserved = max(self.peak_reserved, self.reserved)
            b.launch_geometry()
            self.active.append(b)
            self.batches_run += 1
            self.max_active = max(self.max_active, len(self.active))

    def _advance(self, b):
        if not b.event.query():
            return False
        if b.phase == "geometry":
            if b.K is None:
                b.K = int(b.count[0]) & 0xFFFFFFFF
                self.ipp[b.scene] = max(0.5 * self.ipp[b.scene], b.K / b.r.total_pixels)
            exact = b.fixed_bytes() + self._lists_bytes(b.K)
            extra = exact - b.reserved
            if extra > 0 and self.reserved + extra > self.budget:
                if exact > self.budget:
                    self._finish(b, error=f"needs {exact} B of HBM, budget is {self.budget} B")
                    return True
                b.waiting = True
                return False  # wait for another batch to release memory
            b.waiting = False
            self.reserved += extra
            b.reserved = exact
            self.peak_reserved = max(self.peak_reserved, self.reserved)
            b.phase = "count_read"
            b.launch_render()
            return False
        self._finish(b)
        return True

    def _finish(self, b, error=None):
        now = time.perf_counter()
        if error is None and b.vkey is not None and b.phase == "render":
            self.variant_ms[b.vkey][b.r.variant].append(b.t_start.elapsed_time(b.t_end))
        for v, q in enumerate(b.requests):
            q.t_done = now
            if error is not None:
                q.status, q.error = "rejected", error
            elif q.cancel_requested:
                q.status, q.error = "cancelled", "cancelled while rendering"
            elif q.deadline is not None and now > q.deadline:
                q.status, q.error = "timed_out", f"finished {1000.0 * (now - q.deadline):.3f} ms after its deadline"
            else:
                q.image = b.r.image(b.images, v)
                q.status = "done"
            self.completed.append(q)
        self.reserved -= b.reserved
        self.free_streams.append(b.stream)

    # ---- cancellation and timeouts -------------------------------------------
    def cancel(self, req):
        """Cancel a request: a pending one leaves the queue at once; one whose
        batch is on the GPU finishes with the batch (its batch-mates are
        unaffected) and is then reported as cancelled, without an image.
        Returns False if the request had already completed."""
        if req.status == "pending":
            self.pending.remove(req)
            req.status, req.error = "cancelled", "cancelled before it was scheduled"
            req.t_done = time.perf_counter()
            self.completed.append(req)
            return True
        if req.status == "running":
            req.cancel_requested = True
            return True
        return False

    def _expire(self, now):
        """Requests still queued past their deadline leave the queue as timed
        out (requests already on the GPU are judged when they finish)."""
        if not any(q.timeout_s is not None for q in self.pending):
            return
        keep = collections.deque()
        for q in self.pending:
            if q.deadline is not None and now > q.deadline:
                q.status, q.error = "timed_out", "deadline passed while queued"
                q.t_done = now
                self.completed.append(q)
            else:
                keep.append(q)
        self.pending = keep

    def _requeue(self, b):
        """Give back a counted batch that cannot get its list memory while
        every other batch waits too: its requests return to the head of the
        queue (their geometry pass is repeated later)."""
        for q in reversed(b.requests):
            if q.cancel_requested:
                q.status, q.error = "cancelled", "cancelled while waiting for list memory"
                q.t_done = time.perf_counter()
                self.completed.append(q)
                continue
            q.status = "pending"
            self.pending.appendleft(q)
        self.reserved -= b.reserved
        self.free_streams.append(b.stream)
        self.requeued += 1

    def poll(self):
        """Advance in-flight batches, admit new ones; returns the number of
        batches still in flight."""
        still = []
        for b in self.active:
            if not self._advance(b):
                still.append(b)
        self.active = still
        if len(self.active) > 1 and all(getattr(b, "waiting", False) for b in self.active):
            self._requeue(self.active.pop())  # newest waiting batch yields its memory
        self._expire(time.perf_counter())
        self._admit()
        return len(self.active)

    def drain(self):
        """Serve everything submitted so far."""
        while self.active or self.pending:
            self.poll()
        return self.completed

    def take_completed(self):
        out, self.completed = self.completed, []
        return out

    # ---- open-loop driver ---------------------------------------------------
    def run(self, requests, rate=None):
        """Submit ``requests`` at their ``arrival`` offsets (seconds from the
        start; all at once when arrivals are unset) and serve until all are
        done.  Returns the completed requests.  A request's latency counts
        from its scheduled arrival, so time the host loop spends before it
        notices an arrival is part of the latency."""
        t0 = time.perf_counter()
        order = sorted(requests, key=lambda q: q.arrival or 0.0)
        i = 0
        while i < len(order) or self.active or self.pending:
            now = time.perf_counter() - t0
            while i < len(order) and (order[i].arrival or 0.0) <= now:
                q = self.submit(order[i])
                q.t_submit = t0 + (q.arrival or 0.0)
                i += 1
            self.poll()
        torch.cuda.synchronize(self.device)
        return self.take_completed()


def latency_stats(done, wall_s=None):
    ok = [q for q in done if q.status == "done"]
    lat = np.array([q.latency_ms for q in ok]) if ok else np.zeros(0)
    out = {"completed": len(ok), "rejected": sum(q.status == "rejected" for q in done)}
    for st in ("cancelled", "timed_out"):
        n = sum(q.status == st for q in done)
        if n:
            out[st] = n
    if ok:
        out.update(p50_ms=round(float(np.percentile(lat, 50)), 3), p99_ms=round(float(np.percentile(lat, 99)), 3),
                   mean_batch=round(float(np.mean([q.batch_size for q in ok])), 2))
        hd = [q.latency_ms for q in ok if q.resolution == (1920, 1080)]
        if hd:
            out["p99_ms_1080p"] = round(float(np.percentile(hd, 99)), 3)
            out["p50_ms_1080p"] = round(float(np.percentile(hd, 50)), 3)
    if wall_s:
        out["renders_per_s"] = round(len(ok) / wall_s, 2)
    return out


SERVING_SIZES = [(256, 256), (512, 512), (800, 800), (1280, 720), (1920, 1080)]


def poisson_requests(scene_names, n, rate, seed=3, radius=(3.0, 6.0), sizes=SERVING_SIZES, device="cpu"):
    """n requests with Poisson arrivals at ``rate`` per second, a uniformly
    chosen scene, a camera looking at the origin from a sphere of radius
    U[radius], and a resolution drawn from ``sizes``."""
    g = np.random.default_rng(seed)
    gaps = g.exponential(1.0 / rate, size=n) if rate else np.zeros(n)
    arrivals = np.cumsum(gaps) - gaps[0]
    pos = C.sphere_positions(n, radius, seed)
    reqs = []
    for k in range(n):
        w, h = sizes[g.integers(0, len(sizes))]
        cam = C.camera_at(pos[k], w, h, fov_deg=50.0, device=device)
        reqs.append(Request(scene_names[g.integers(0, len(scene_names))], cam, arrival=float(arrivals[k])))
    return reqs


def serve_once(server, requests):
    t0 = time.perf_counter()
    done = server.run(requests)
    wall = time.perf_counter() - t0
    return latency_stats(done, wall)


def main():
    """Sweep of arrival rates over four resident scenes (0.5M/1M/2M/3M);
    prints one JSON line per rate."""
    import argparse
    import json
    from . import scenes as S
    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", default="50,100,200,400")
    ap.add_argument("--requests", type=int, default=300)
    ap.add_argument("--sizes", default="500000,1000000,2000000,3000000")
    ap.add_argument("--max-views", type=int, default=8)
    ap.add_argument("--streams", type=int, default=4)
    args = ap.parse_args()
    dev = "cuda"
    scenes = {}
    for i, n in enumerate(int(x) for x in args.sizes.split(",")):
        m = S.free_scene(n, seed=10 + i, layout="mixture", shell_radius=3.0, device=dev)
        scenes[f"s{n}"] = BT.pack_model(m)
    names = sorted(scenes)
    server = RenderServer(scenes, device=dev, max_views=args.max_views, num_streams=args.streams)
    serve_once(server, poisson_requests(names, 16, None, seed=99))  # warm-up
    for rate in (float(r) for r in args.rates.split(",")):
        stats = serve_once(server, poisson_requests(names, args.requests, rate, seed=int(rate)))
        print(json.dumps({"rate_per_s": rate, **stats, "peak_reserved_GB": round(server.peak_reserved / 1e9, 3)}),
              flush=True)


if __name__ == "__main__":
    main()
