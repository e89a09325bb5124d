"""Batched multi-view rendering: B views of one scene in one launch sequence.

This replaces the per-camera loop of the training and evaluation scripts
(train_geo_stage3.py:104-124 renders one camera per iteration through
``gaussian_renderer.render`` and reads ``num_rendered`` back to the host on
every call).  Here:

* the scene is one component-major tensor ``params`` of 59 x P raw
  parameters (``pack_params``); the exp / normalise / sigmoid activations of
  scene/gaussian_model.py:95-118 run inside the kernels, and the backward
  returns raw-parameter gradients in the same layout (``unpack_grads``);
* all views go through one preprocess, one depth sort, one duplicate, one
  tile sort and one blend launch (views may differ in resolution);
* the instance count stays on the device: the list capacity is planned once
  from a counting pass and checked after the fact (``overflowed``), so a
  steady-state step has no host synchronisation and can be captured in a
  HIP graph (``BatchedStep``);
* the blend work-group shape / LDS batch depth is picked by ``autotune``.

Images and gradients equal those of B single-view ``render`` calls followed
by autograd through the activations (same per-view arithmetic).
"""
import ctypes
import math
import time

import torch

from . import _native as N

CHANNELS = 59
XYZ, SCALE, ROT, OPACITY, SH = 0, 3, 6, 10, 11
STAGE_GEOMETRY, STAGE_BINNING, STAGE_BLEND, STAGES_ALL = 1, 2, 4, 7
RENDER_VARIANTS = ("16x16 px, 256-deep LDS batch", "16x16 px, 128-deep", "8x8 px, 64-deep", "8x8 px, 128-deep",
                   "one wave per tile, 64-deep", "one wave per tile, 128-deep",
                   "16x16 px 256-deep forward, one-wave-per-tile 64-deep backward")


def pack_params(xyz, log_scale, rotation, opacity_logit, features_dc, features_rest, out=None):
    """Raw model parameters -> (59, P) component-major float32 tensor.

    Channels: 0-2 xyz, 3-5 log-scale, 6-9 raw quaternion (w, x, y, z),
    10 opacity logit, 11 + 3k + c SH coefficient k of colour c (k = 0 is the
    DC term; degrees below 3 are zero-padded to 16 coefficients)."""
    P = xyz.shape[0]
    if out is None:
        out = torch.zeros(CHANNELS, P, dtype=torch.float32, device=xyz.device)
    with torch.no_grad():
        out[XYZ:XYZ + 3] = xyz.detach().t()
        out[SCALE:SCALE + 3] = log_scale.detach().t()
        out[ROT:ROT + 4] = rotation.detach().t()
        out[OPACITY] = opacity_logit.detach().reshape(P)
        sh = torch.cat((features_dc.detach(), features_rest.detach()), dim=1).reshape(P, -1)
        out[SH:SH + sh.shape[1]] = sh.t()
        out[SH + sh.shape[1]:] = 0
    return out


def pack_model(model, out=None):
    """``pack_params`` of a model with the reference's raw attributes."""
    return pack_params(model._xyz, model._scaling, model._rotation, model._opacity, model._features_dc,
                       model._features_rest, out=out)


def unpack_grads(grads, sh_coeffs=16):
    """(59, P) gradients -> dict shaped like the model's raw parameters;
    ``sh_coeffs`` is the model's SH coefficients per channel,
    (max_sh_degree + 1)^2 (16 for degree 3, 1 for degree 0)."""
    P = grads.shape[1]
    sh = grads[SH:SH + 3 * sh_coeffs].t().reshape(P, sh_coeffs, 3)
    return {
        "_xyz": grads[XYZ:XYZ + 3].t(),
        "_scaling": grads[SCALE:SCALE + 3].t(),
        "_rotation": grads[ROT:ROT + 4].t(),
        "_opacity": grads[OPACITY].reshape(P, 1),
        "_features_dc": sh[:, :1],
        "_features_rest": sh[:, 1:],
    }


def _view_record(cam, bg):
    v = N.DmgsView()
    wv = cam.world_view_transform.detach().float().cpu().contiguous().reshape(-1).tolist()
    fp = cam.full_proj_transform.detach().float().cpu().contiguous().reshape(-1).tolist()
    cc = cam.camera_center.detach().float().cpu().reshape(-1).tolist()
    b = torch.as_tensor(bg).detach().float().cpu().reshape(-1).tolist()
    for i in range(16):
        v.viewmatrix[i] = wv[i]
        v.projmatrix[i] = fp[i]
    for i in range(3):
        v.campos[i] = cc[i]
        v.bg[i] = b[i]
    v.tanfovx = math.tan(cam.FoVx * 0.5)
    v.tanfovy = math.tan(cam.FoVy * 0.5)
    v.width, v.height = int(cam.image_width), int(cam.image_height)
    return v


class BatchedRenderer:
    """Forward / backward of B views of a P-Gaussian scene.

    ``cameras`` carry the reference Camera fields the renderer reads
    (world_view_transform, full_proj_transform, camera_center, FoVx, FoVy,
    image_width, image_height); ``bg`` is one colour or one per view."""

    def __init__(self, cameras, bg, num_gaussians, sh_degree=3, scale_modifier=1.0, device="cuda",
                 headroom=1.2, variant=0, cell_mode=None):
        if not cameras:
            raise ValueError("BatchedRenderer needs at least one camera")
        self.lib = N.load()
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.B, self.P = len(cameras), int(num_gaussians)
        self.sh_degree, self.scale_modifier = int(sh_degree), float(scale_modifier)
        self.headroom, self.variant = float(headroom), int(variant)
        bgs = bg if isinstance(bg, (list, tuple)) else [bg] * self.B
        self.views = (N.DmgsView * self.B)(*[_view_record(c, b) for c, b in zip(cameras, bgs)])
        tt, tc, mc, tp = N.c_int32(0), N.c_int32(0), N.c_int32(0), N.c_int64(0)
        N.check(self.lib.dmgs_batched_prepare_views(self.views, self.B, ctypes.byref(tt), ctypes.byref(tc),
                                                    ctypes.byref(mc), ctypes.byref(tp)), "dmgs_batched_prepare_views")
        self.total_tiles, self.total_cells, self.total_pixels = tt.value, tc.value, tp.value
        self.max_view_cells = mc.value
        # view-major lists (one cell-sort pass for views of up to 256 cells) where they apply
        fits = self.B <= 16 and self.max_view_cells <= 4096
        self.cell_mode = (1 if fits else 0) if cell_mode is None else int(cell_mode)
        # both tables go through pinned host memory and asynchronous copies on
        # the current stream: building a renderer never waits for the device
        # (a render server builds one per batch while other batches run);
        # the host buffers live as long as the renderer
        pin = self.device.type == "cuda"
        tv = torch.empty(self.total_tiles, dtype=torch.int32, pin_memory=pin)
        N.check(self.lib.dmgs_batched_tile_view(self.views, self.B, ctypes.c_void_p(tv.data_ptr()),
                                                self.total_tiles), "dmgs_batched_tile_view")
        nbytes = ctypes.sizeof(self.views)
        vh = torch.empty(nbytes, dtype=torch.uint8, pin_memory=pin)
        ctypes.memmove(vh.data_ptr(), ctypes.addressof(self.views), nbytes)
        self._host_tables = (tv, vh)
        self.tile_view = tv.to(self.device, non_blocking=pin)
        self.views_dev = vh.to(self.device, non_blocking=pin)
        self._tables_ready = None
        if pin:  # launches on another stream wait for the copies (see _after_tables)
            self._tables_ready = torch.cuda.Event()
            self._tables_ready.record()
        self.sizes = [(v.height, v.width) for v in self.views]
        self.pixel_base = [v.pixel_base for v in self.views]
        self.K_cap = 0
        self.state = self.lists = self.bwd_ws = None
        self.dens = None
        self._offsets = None

    # ---- planning ---------------------------------------------------------
    def _alloc_state(self):
        n = self.lib.dmgs_batched_state_size(self.B, self.P, self.total_tiles, self.total_pixels)
        if n == 0:
            raise RuntimeError("dmgs_batched_state_size rejected the batch")
        offs = (N.c_size_t * 9)()
        N.check(self.lib.dmgs_batched_offsets(self.B, self.P, self.total_tiles, self.total_pixels, 0, offs),
                "dmgs_batched_offsets")
        self.state = torch.empty(n, dtype=torch.uint8, device=self.device)
        self._offsets = list(offs)

    def _after_tables(self):
        """Order the current stream after the tables' copies (a no-op on the
        stream that built the renderer; skipped under graph capture, which
        starts from a synchronised device)."""
        if self._tables_ready is not None and not torch.cuda.is_current_stream_capturing():
            torch.cuda.current_stream(self.device).wait_event(self._tables_ready)

    def _desc(self, capacity=None):
        self._after_tables()
        d = N.DmgsBatch()
        d.views, d.tile_view = self.views_dev.data_ptr(), self.tile_view.data_ptr()
        d.num_views, d.sh_degree, d.num_gaussians = self.B, self.sh_degree, self.P
        d.scale_modifier, d.variant = self.scale_modifier, self.variant
        d.total_tiles, d.total_cells, d.total_pixels = self.total_tiles, self.total_cells, self.total_pixels
        d.cell_mode, d.max_view_cells = self.cell_mode, self.max_view_cells
        d.capacity = self.K_cap if capacity is None else capacity
        d.state = self.state.data_ptr()
        d.lists = None if self.lists is None else self.lists.data_ptr()
        d.bwd_ws = None if self.bwd_ws is None else self.bwd_ws.data_ptr()
        d.dens = None if self.dens is None else self.dens.data_ptr()
        return d

    def count_instances(self, params):
        """Run preprocess + depth sort + scan only and return the number of
        (Gaussian, tile) instances of the batch (one host read)."""
        if self.state is None:
            self._alloc_state()
        self._check_params(params)
        N.check(self.lib.dmgs_batched_forward(ctypes.byref(self._desc(0)), N.ptr(params), None, STAGE_GEOMETRY,
                                              N.stream_handle(self.device)), "dmgs_batched_forward")
        return int(self.instance_count().item()) & 0xFFFFFFFF

    def count_instances_async(self, params, host_out):
        """Launch preprocess + depth sort + scan and copy the instance count
        into ``host_out[0]`` (pinned int32 tensor) without waiting, and the
        status word (bit 0: see ``depth_clamped``) into ``host_out[1]`` when
        it has two elements; valid once the current stream has reached it."""
        if self.state is None:
            self._alloc_state()
        self._check_params(params)
        N.check(self.lib.dmgs_batched_forward(ctypes.byref(self._desc(0)), N.ptr(params), None, STAGE_GEOMETRY,
                                              N.stream_handle(self.device)), "dmgs_batched_forward")
        host_out[0:1].copy_(self.instance_count(), non_blocking=True)
        if host_out.numel() > 1:
            o = self._offsets[8]
            host_out[1:2].copy_(self.state[o:o + 4].view(torch.int32), non_blocking=True)

    def plan(self, params, K_cap=None, backward=True):
        """Size the instance lists for ``params`` (with headroom) and allocate
        the workspaces.  Returns the capacity."""
        if self.state is None:
            self._alloc_state()
        if K_cap is None:
            K = self.count_instances(params)
            K_cap = max(1024, int(math.ceil(K * self.headroom)))
        self.K_cap = int(K_cap)
        self.lists = torch.empty(self.lib.dmgs_batched_lists_size(self.K_cap), dtype=torch.uint8, device=self.device)
        if backward and self.bwd_ws is None:  # the forward blend clears it before each backward
            self.bwd_ws = torch.empty(self.lib.dmgs_batched_backward_size(self.B, self.P), dtype=torch.uint8,
                                      device=self.device)
            self._acc_clear = False
        return self.K_cap

    def _require_plan(self):
        if self.lists is None:
            raise RuntimeError("BatchedRenderer.plan() must run before forward/backward")

    def _check_params(self, params):
        if params.dtype != torch.float32 or not params.is_contiguous() or tuple(params.shape) != (CHANNELS, self.P):
            raise ValueError(f"params must be a contiguous float32 ({CHANNELS}, {self.P}) tensor")
        if params.device != self.device:
            raise ValueError("params must live on the renderer's device")

    # ---- launches ---------------------------------------------------------
    def new_images(self):
        return torch.empty(3 * self.total_pixels, dtype=torch.float32, device=self.device)

    def forward(self, params, images=None, stages=STAGES_ALL):
        """Render every view into one packed buffer (view v at 3*pixel_base[v],
        laid out (3, H, W)); returns it.  ``stages`` may split the call
        (geometry, binning, blend; run in that order)."""
        self._require_plan()
        self._check_params(params)
        if images is None:
            images = self.new_images()
        N.check(self.lib.dmgs_batched_forward(ctypes.byref(self._desc()), N.ptr(params), N.ptr(images), stages,
                                              N.stream_handle(self.device)), "dmgs_batched_forward")
        if stages & STAGE_BLEND and self.bwd_ws is not None:
            self._acc_clear = True  # the blend cleared the backward accumulator
        return images

    def backward(self, params, dL_dimages, grads=None, stages=STAGE_BLEND | STAGE_GEOMETRY):
        """Raw-parameter gradients (59, P) of sum over views of <dL/dimage, image>
        (``stages``: blend backward, then the per-Gaussian backward)."""
        self._require_plan()
        self._check_params(params)
        if self.bwd_ws is None:
            raise RuntimeError("planned without backward")
        if dL_dimages.numel() != 3 * self.total_pixels or not dL_dimages.is_contiguous():
            raise ValueError("dL_dimages must be a contiguous buffer packed like the images")
        if grads is None:
            grads = torch.empty(CHANNELS, self.P, dtype=torch.float32, device=self.device)
        if stages & STAGE_BLEND:
            if not getattr(self, "_acc_clear", False):  # a second backward after one forward
                self.bwd_ws.zero_()
            self._acc_clear = False
        N.check(self.lib.dmgs_batched_backward(ctypes.byref(self._desc()), N.ptr(params), N.ptr(dL_dimages),
                                               N.ptr(grads), stages, N.stream_handle(self.device)),
                "dmgs_batched_backward")
        return grads

    # ---- densification statistics ------------------------------------------
    def track_densification(self, enable=True):
        """Let every backward accumulate, per Gaussian over the views where it
        is visible and in view order, the statistics the reference's training
        loop gathers one view at a time (gaussian_model.py:405-407,
        train.py:116-118): |dL/dmean2D| in NDC, the number of views and the
        largest screen radius.  Starts from zero."""
        if enable and self.dens is None:
            self.dens = torch.zeros(3, self.P, dtype=torch.float32, device=self.device)
        elif not enable:
            self.dens = None

    def densification_stats(self):
        """(xyz_gradient_accum (P, 1), denom (P, 1), max_radii2D (P,)) views
        of the accumulated statistics."""
        if self.dens is None:
            raise RuntimeError("track_densification() is off")
        return self.dens[0].view(-1, 1), self.dens[1].view(-1, 1), self.dens[2]

    def reset_densification_stats(self):
        if self.dens is not None:
            self.dens.zero_()

    # ---- results ----------------------------------------------------------
    def image(self, packed, v):
        """(3, H, W) view of image ``v`` inside a packed buffer."""
        H, W = self.sizes[v]
        o = 3 * self.pixel_base[v]
        return packed[o:o + 3 * H * W].view(3, H, W)

    def instance_count(self):
        """Device tensor (1,) holding the instance count of the last forward."""
        o = self._offsets[1]
        return self.state[o:o + 4].view(torch.int32)

    def depth_clamped(self):
        """True when the last view-major forward met a visible depth of 2^27
        or more (its lists may misorder such far splats).  Reads back."""
        o = self._offsets[8]
        return self.cell_mode == 1 and (int(self.state[o:o + 4].view(torch.int32).item()) & 1) != 0

    def overflowed(self):
        """True when the last forward must be redone: it needed more
        instances than planned (its lists were truncated; re-plan and render
        again) or, in view-major mode, met a depth the view-major keys cannot
        order (the renderer then switches to global cell lists; render
        again).  Reads back."""
        if self.depth_clamped():
            self.cell_mode = 0
            return True
        return (int(self.instance_count().item()) & 0xFFFFFFFF) > self.K_cap

    def cell_lists(self):
        """(ranges (cells, 2), sorted keys, entry ids) of the last forward as
        CPU tensors.  A key is cell | tile mask << 16 (cells of 4x4 tiles,
        ``ranges`` indexed by global cell id); the cell is the global id in
        cell mode 0 and (cell within the view | view << 12) in mode 1."""
        ko, vo = N.c_size_t(0), N.c_size_t(0)
        N.check(self.lib.dmgs_batched_list_offsets(ctypes.byref(self._desc()), ctypes.byref(ko), ctypes.byref(vo)),
                "dmgs_batched_list_offsets")
        K = min(int(self.instance_count().item()) & 0xFFFFFFFF, self.K_cap)
        o = self._offsets[2]
        ranges = self.state[o:o + 8 * self.total_cells].view(torch.int32).view(-1, 2).cpu()
        keys = self.lists[ko.value:ko.value + 4 * K].view(torch.int32).cpu()
        vals = self.lists[vo.value:vo.value + 4 * K].view(torch.int32).cpu()
        return ranges, keys, vals

    def radii(self):
        """(B, P) int32 radii of the last forward."""
        o = self._offsets[0]
        return self.state[o:o + 4 * self.B * self.P].view(torch.int32).view(self.B, self.P)

    # ---- autotuning -------------------------------------------------------
    def autotune(self, params, dL_dimages=None, reps=3):
        """Time forward + backward for every blend variant and keep the
        fastest.  Returns {variant: ms}."""
        self._require_plan()
        if dL_dimages is None:
            dL_dimages = torch.full((3 * self.total_pixels,), 1e-3, device=self.device)
        images = self.new_images()
        grads = torch.empty(CHANNELS, self.P, dtype=torch.float32, device=self.device)
        times = {}
        for v in range(len(RENDER_VARIANTS)):
            self.variant = v
            self.forward(params, images)
            self.backward(params, dL_dimages, grads)
            torch.cuda.synchronize(self.device)
            t0 = time.perf_counter()
            for _ in range(reps):
                self.forward(params, images)
                self.backward(params, dL_dimages, grads)
            torch.cuda.synchronize(self.device)
            times[v] = 1000.0 * (time.perf_counter() - t0) / reps
        self.variant = min(times, key=times.get)
        return times


class BatchedLoss:
    """Summed per-view (1 - lambda) L1 + lambda (1 - SSIM) against fixed
    targets (train_geo_stage3.py:122-124 applied to every view of the
    batch), on the fused loss kernels; no host synchronisation."""

    def __init__(self, renderer, targets, lambda_dssim=0.2):
        if len(targets) != renderer.B:
            raise ValueError("one target per view")
        self.r, self.lam = renderer, float(lambda_dssim)
        self.lib = N.load()
        dev = renderer.device
        self.targets = renderer.new_images()
        for v, t in enumerate(targets):
            renderer.image(self.targets, v).copy_(t.to(device=dev, dtype=torch.float32))
        # consecutive views of one size are back to back in the packed buffers:
        # each such run is one grouped launch pair (one loss per view)
        self.runs = []
        for v, hw in enumerate(renderer.sizes):
            if self.runs and self.runs[-1][2] == hw:
                self.runs[-1][1] = v + 1
            else:
                self.runs.append([v, v + 1, hw])
        self.ws = [torch.empty(self.lib.dmgs_dssim_l1_workspace_size(3 * (v1 - v0), h, w), dtype=torch.uint8,
                               device=dev) for v0, v1, (h, w) in self.runs]
        # the targets' blurred maps (G*y, G*y^2) are fixed: computed once here
        self.stats = []
        s = N.stream_handle(dev)
        for v0, v1, (h, w) in self.runs:
            st = torch.empty(2 * 3 * (v1 - v0) * h * w, dtype=torch.float32, device=dev)
            N.check(self.lib.dmgs_dssim_target_stats(N.ptr(renderer.image(self.targets, v0)), 3 * (v1 - v0), h, w,
                                                     N.ptr(st), s), "dmgs_dssim_target_stats")
            self.stats.append(st)
        self.losses = torch.zeros(renderer.B, 3, dtype=torch.float32, device=dev)
        self.ones = torch.ones(renderer.B, dtype=torch.float32, device=dev)

    def forward(self, images):
        """(B, 3) device tensor: (loss, l1, ssim) per view."""
        r, s = self.r, N.stream_handle(self.r.device)
        for ws, st, (v0, v1, (h, w)) in zip(self.ws, self.stats, self.runs):
            N.check(self.lib.dmgs_dssim_l1_forward_groups_stats(
                N.ptr(r.image(images, v0)), N.ptr(r.image(self.targets, v0)), N.ptr(st), v1 - v0, 3, h, w, self.lam,
                N.ptr(ws), N.ptr(self.losses[v0]), s), "dmgs_dssim_l1_forward_groups_stats")
        return self.losses

    def backward(self, images, dimages):
        """dL/dimages of the summed per-view loss (every view's dloss = 1)."""
        r, s = self.r, N.stream_handle(self.r.device)
        for ws, (v0, v1, (h, w)) in zip(self.ws, self.runs):
            N.check(self.lib.dmgs_dssim_l1_backward_groups(
                N.ptr(r.image(images, v0)), N.ptr(r.image(self.targets, v0)), v1 - v0, 3, h, w, self.lam, N.ptr(ws),
                N.ptr(self.ones[v0:v1]), N.ptr(r.image(dimages, v0)), s), "dmgs_dssim_l1_backward_groups")
        return dimages


class BatchedStep:
    """One fwd+bwd training step over all views: batched render, summed loss,
    loss backward, batched render backward into raw-parameter gradients.

    ``run()`` launches eagerly; ``capture()`` records the same launches in a
    HIP graph and ``replay()`` re-issues them (the scene tensor, targets and
    views are fixed at capture; update ``params`` in place between replays).

    Instance capacity: the lists are sized once by ``plan`` (with headroom).
    As the parameters change (Gaussians grow, move closer) a step can need
    more instances than planned; its lists are then truncated and its images
    and gradients are wrong.  Every step therefore copies the instance count
    and the status word into pinned host memory behind its kernels, and
    ``check()`` — called by ``run`` / ``replay`` unless ``check=False`` —
    waits for them and, if the step overflowed (or, with view-major lists,
    met a depth those keys cannot order), re-plans the capacity from the
    step's own count (or switches to global cell lists), re-captures the
    graph if there is one and runs the step again, so the step's outputs are
    correct when ``check`` returns.  With ``check=False`` (timed loops) call
    ``check()`` before using a step's outputs."""

    def __init__(self, renderer, loss, params):
        self.r, self.loss, self.params = renderer, loss, params
        self.images = renderer.new_images()
        self.dimages = renderer.new_images()
        self.grads = torch.empty(CHANNELS, renderer.P, dtype=torch.float32, device=renderer.device)
        self.graph = None
        self.replans = 0
        self._status = torch.zeros(2, dtype=torch.int32, pin_memory=True)  # instances, status word
        self._status_event = torch.cuda.Event()
        self._status_pending = False

    def _launch(self):
        self.r.forward(self.params, self.images)
        self.loss.forward(self.images)
        self.loss.backward(self.images, self.dimages)
        self.r.backward(self.params, self.dimages, self.grads)

    def _record_status(self):
        # behind the step on the same stream, outside any graph
        self._status[0:1].copy_(self.r.instance_count(), non_blocking=True)
        o = self.r._offsets[8]
        self._status[1:2].copy_(self.r.state[o:o + 4].view(torch.int32), non_blocking=True)
        self._status_event.record()
        self._status_pending = True

    def run(self, check=True):
        self._launch()
        self._record_status()
        if check:
            self.check()
        return self.loss.losses

    def capture(self):
        self._launch()  # warm the kernels and the allocator outside capture
        self._record_status()
        torch.cuda.synchronize(self.r.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._launch()
        self.graph = g
        return g

    def replay(self, check=True):
        if self.graph is None:
            raise RuntimeError("capture() first")
        self.graph.replay()
        self._record_status()
        if check:
            self.check()
        return self.loss.losses

    def overflowed(self):
        """Whether the last step must be redone (waits for its status)."""
        if not self._status_pending:
            return False
        self._status_event.synchronize()
        K = int(self._status[0].item()) & 0xFFFFFFFF
        far = self.r.cell_mode == 1 and (int(self._status[1].item()) & 1) != 0
        return far or K > self.r.K_cap

    def check(self):
        """Wait for the last step; if it overflowed, re-plan and redo it (see
        the class docstring).  Returns True when the step was redone."""
        redone = False
        while self.overflowed():
            K = int(self._status[0].item()) & 0xFFFFFFFF
            if self.r.cell_mode == 1 and (int(self._status[1].item()) & 1) != 0:
                self.r.cell_mode = 0
            if K > self.r.K_cap:
                self.r.plan(self.params, K_cap=max(1024, int(math.ceil(K * self.r.headroom))))
            self.replans += 1
            redone = True
            if self.graph is not None:
                self.graph = None
                self.capture()  # its eager launch is the redone step
            else:
                self._launch()
                self._record_status()
        self._status_pending = False
        return redone
