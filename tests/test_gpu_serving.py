"""GPU tests of the render server (continuous batching over resident scenes).

* A request's image is the same alone and in a mix of concurrent batches of
  other scenes, cameras and resolutions (bit-identical; the stated bar is
  1e-4 abs), and equals the drop-in single-view ``render`` within 1e-4 on
  every pixel the FP64 oracle does not flag as a blend threshold tie.
* Admission respects the HBM budget: requests that can never fit are
  rejected; under a budget that holds one batch at a time everything still
  completes and the reservation never exceeds the budget.
* Batches overlap: with several streams more than one batch is in flight.
"""
import pytest
import torch

from dmgs_amd import batched as BT
from dmgs_amd import cameras as C
from dmgs_amd import scenes as S
from dmgs_amd import serving as SV
from dmgs_amd.gaussian_renderer import PipelineParams, render
from oracle import raster as OR

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-4


def _scenes():
    a = S.free_scene(20000, seed=21, layout="mixture", shell_radius=3.0)
    b = S.free_scene(40000, seed=22, layout="uniform")
    return {"a": a, "b": b}


def _packed(models):
    return {k: BT.pack_model(m.to("cuda")) for k, m in models.items()}


def _cam(pos, w, h):
    return C.camera_at(pos, w, h, fov_deg=50.0, device="cuda")


def _cpu(cam):
    class Cam:
        pass
    c = Cam()
    for k in ("image_width", "image_height", "FoVx", "FoVy"):
        setattr(c, k, getattr(cam, k))
    for k in ("world_view_transform", "full_proj_transform", "camera_center"):
        setattr(c, k, getattr(cam, k).detach().cpu())
    return c


def test_same_result_alone_and_in_the_mix():
    print("case: serving alone vs mix", flush=True)
    models = _scenes()
    scenes = _packed(models)
    probe_pos, probe_size = (0.5, -1.0, 3.6), (200, 150)
    # alone
    srv = SV.RenderServer(scenes, max_views=4, num_streams=3)
    alone = srv.run([SV.Request("a", _cam(probe_pos, *probe_size))])
    assert len(alone) == 1 and alone[0].status == "done", alone[0].error
    img_alone = alone[0].image.clone()
    # in a mix: same camera among other views of both scenes, all at once
    pos = C.sphere_positions(11, (3.0, 6.0), seed=5)
    sizes = [(128, 128), (256, 144), (96, 200), (320, 180)]
    mix = [SV.Request("a" if k % 2 else "b", _cam(pos[k], *sizes[k % 4])) for k in range(11)]
    probe = SV.Request("a", _cam(probe_pos, *probe_size))
    mix.insert(5, probe)
    srv2 = SV.RenderServer(scenes, max_views=4, num_streams=3)
    done = srv2.run(mix)
    assert all(q.status == "done" for q in done), [q.error for q in done if q.status != "done"]
    assert probe.batch_size > 1, "the probe should share a batch with other views"
    assert srv2.max_active > 1, "batches should overlap"
    err = (probe.image - img_alone).abs().max().item()
    print(f"alone vs mix max diff {err:.3g}; probe batch size {probe.batch_size}; max in flight {srv2.max_active}",
          flush=True)
    assert torch.equal(probe.image, img_alone)
    # and the drop-in single-view render of that camera
    m = models["a"]
    ref = render(probe.camera, m.to("cuda"), PipelineParams(), torch.ones(3, device="cuda"))["render"]
    st = OR.settings_for(_cpu(probe.camera), torch.ones(3, dtype=torch.float64), 3)
    _, _, state = OR.rasterize(m.get_xyz, m.get_opacity, st, shs=m.get_features, scales=m.get_scaling,
                               rotations=m.get_rotation)
    ok = ~state["ambiguous"].cuda()
    e2 = (probe.image - ref).abs()[:, ok].max().item()
    assert e2 < IMG_TOL, e2
    # every mixed request equals its own alone render
    for q in done[:4]:
        solo = SV.RenderServer(scenes, max_views=1, num_streams=1).run([SV.Request(q.scene, q.camera)])[0]
        assert torch.equal(solo.image, q.image), f"request {q.id} differs alone"


def test_admission_by_hbm_budget():
    print("case: serving admission", flush=True)
    scenes = _packed(_scenes())
    pos = C.sphere_positions(8, 4.0, seed=7)
    reqs = [SV.Request("a" if k < 4 else "b", _cam(pos[k], 160, 120)) for k in range(8)]
    # a budget far below one request: everything is rejected before running
    tiny = SV.RenderServer(scenes, hbm_budget=1 << 16)
    out = tiny.run([SV.Request(q.scene, q.camera) for q in reqs])
    assert all(q.status == "rejected" for q in out) and tiny.batches_run == 0
    # a budget that holds about one single-view batch: all complete, budget respected
    probe = SV.RenderServer(scenes)
    one = max(probe.footprint("a", [reqs[0].camera]), probe.footprint("b", [reqs[4].camera]))
    tight = SV.RenderServer(scenes, hbm_budget=int(1.5 * one), max_views=8, num_streams=4)
    out = tight.run(reqs)
    assert all(q.status == "done" for q in out), [q.error for q in out]
    assert tight.peak_reserved <= tight.budget
    assert tight.batches_run >= 2
    # the same requests with a generous budget give the same images
    free = SV.RenderServer(scenes, max_views=8, num_streams=4)
    again = free.run([SV.Request(q.scene, q.camera) for q in reqs])
    for q1, q2 in zip(sorted(out, key=lambda q: q.id), sorted(again, key=lambda q: q.id)):
        assert torch.equal(q1.image, q2.image)


def test_poisson_stream_statistics():
    print("case: serving Poisson arrivals", flush=True)
    scenes = _packed(_scenes())
    srv = SV.RenderServer(scenes, max_views=8, num_streams=4)
    reqs = SV.poisson_requests(["a", "b"], 60, rate=400.0, seed=9, sizes=[(256, 256), (512, 512), (640, 360)])
    stats = SV.serve_once(srv, reqs)
    print(stats, flush=True)
    assert stats["completed"] == 60 and stats["rejected"] == 0
    assert 0 < stats["p50_ms"] <= stats["p99_ms"]
    assert stats["renders_per_s"] > 0


def test_cancel_and_timeout_leave_other_requests_intact():
    """One stream, batches of two: a running and a queued request are
    cancelled and a request with a zero timeout expires in the queue; every
    other request completes with the image it gets when served alone."""
    print("case: serving cancellation and timeouts", flush=True)
    scenes = _packed(_scenes())
    srv = SV.RenderServer(scenes, max_views=2, num_streams=1)
    pos = C.sphere_positions(6, 3.8, 31)
    reqs = [SV.Request("a" if i % 2 else "b", _cam(p, 128, 96)) for i, p in enumerate(pos)]
    for q in reqs:
        srv.submit(q)
    srv.poll()
    running = [q for q in reqs if q.status == "running"]
    pending = [q for q in reqs if q.status == "pending"]
    assert running and pending
    assert srv.cancel(running[0]) and srv.cancel(pending[0])
    late = srv.submit(SV.Request("b", _cam(pos[0], 64, 64), timeout_s=0.0))
    done = srv.drain()
    torch.cuda.synchronize()
    status = {q.id: q.status for q in done}
    assert status[running[0].id] == "cancelled" and status[pending[0].id] == "cancelled"
    assert status[late.id] == "timed_out" and late.image is None
    survivors = [q for q in reqs if q not in (running[0], pending[0])]
    assert all(status[q.id] == "done" for q in survivors)
    for q in survivors:
        solo = SV.RenderServer(scenes, max_views=1, num_streams=1).run([SV.Request(q.scene, q.camera)])
        assert torch.equal(solo[0].image, q.image), f"request {q.id} differs from its solo render"
