"""Host logic of the render server's blend-shape choice (CPU only).

With no fixed variant the server times each shape of AUTO_VARIANTS twice per
(scene, size class) and then keeps the one with the lower median of its last
eight batches; a fixed variant is used for every batch and nothing is timed.
"""
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
