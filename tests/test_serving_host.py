"""Host logic of the render server (CPU only).

* Blend-shape choice: with no fixed variant the server times each shape of
  AUTO_VARIANTS twice per (scene, size class) and then keeps the one with the
  lower median of its last eight batches; a fixed variant is used for every
  batch and nothing is timed.
* Cancellation and timeouts of queued and running requests.
"""
import collections
import time
import types

from dmgs_amd import serving as SV


def _server(variant=None):
    srv = types.SimpleNamespace(variant=variant, variant_ms={})
    srv.choose = lambda scene, tiles: SV.RenderServer._choose_variant(srv, scene, tiles)
    return srv


def test_fixed_variant_is_not_timed():
    srv = _server(variant=5)
    assert srv.choose("a", 1000) == (5, None)
    assert srv.variant_ms == {}


def test_each_shape_is_timed_twice_then_the_faster_kept():
    srv = _server()
    shapes = SV.AUTO_VARIANTS
    seen = []
    for _ in range(2 * len(shapes)):
        v, key = srv.choose("a", 3000)
        assert key == ("a", (3000).bit_length())
        seen.append(v)
        srv.variant_ms[key][v].append(1.0 if v == shapes[-1] else 2.0)
    assert sorted(seen) == sorted(list(shapes) * 2)
    assert srv.choose("a", 3000)[0] == shapes[-1]
    # another size class or scene starts its own trials
    assert srv.choose("a", 100)[0] == shapes[0]
    assert srv.choose("b", 3000)[0] == shapes[0]


def test_choice_follows_recent_times():
    srv = _server()
    key = ("a", (5000).bit_length())
    srv.variant_ms[key] = {v: [1.0, 1.0] for v in SV.AUTO_VARIANTS}
    fast, slow = SV.AUTO_VARIANTS[0], SV.AUTO_VARIANTS[1]
    srv.variant_ms[key][slow] = [0.5, 0.5]
    assert srv.choose("a", 5000)[0] == slow
    srv.variant_ms[key][slow] += [3.0] * 8  # only the last eight count
    assert srv.choose("a", 5000)[0] == fast


# ---- cancellation and timeouts ----
def _queue_server(reqs):
    srv = types.SimpleNamespace(pending=collections.deque(reqs), completed=[])
    srv.cancel = lambda q: SV.RenderServer.cancel(srv, q)
    srv.expire = lambda now: SV.RenderServer._expire(srv, now)
    return srv


def _req(timeout=None):
    q = SV.Request("a", camera=None, timeout_s=timeout)
    q.status = "pending"
    q.t_submit = 100.0
    return q


def test_cancel_pending_and_running():
    a, b, c = _req(), _req(), _req()
    srv = _queue_server([a, b])
    assert srv.cancel(a) and a.status == "cancelled" and a.image is None
    assert list(srv.pending) == [b] and srv.completed == [a]
    c.status = "running"  # on the GPU: flagged, judged when its batch finishes
    assert srv.cancel(c) and c.cancel_requested and c.status == "running"
    done = _req()
    done.status = "done"
    assert not srv.cancel(done) and done.status == "done"


def test_queued_requests_time_out_at_their_deadline():
    fast, slow, free = _req(0.001), _req(10.0), _req(None)
    srv = _queue_server([fast, slow, free])
    srv.expire(100.0005)
    assert list(srv.pending) == [fast, slow, free]
    srv.expire(100.002)
    assert fast.status == "timed_out" and list(srv.pending) == [slow, free]
    srv.expire(1e9)
    assert slow.status == "timed_out" and list(srv.pending) == [free] and free.status == "pending"
    assert all(q.t_done is not None for q in srv.completed)
    stats = SV.latency_stats(srv.completed)
    assert stats["timed_out"] == 2 and stats["completed"] == 0
    assert SV.Request("a", None).deadline is None and time.perf_counter() > 0
