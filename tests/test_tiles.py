"""Tile shapes of the fused two-factor separable kernel: the rule table and
its ordering (CPU), and on the GPU that 64 x 64 tiles give bit-identical
results to 64 x 32 tiles (single-image and batched), within 1e-5 of the FP64
oracle, and that the library's lookup follows the installed rules."""
import ctypes as C

import numpy as np
import pytest

import imagefiltering as IF
from imagefiltering import KernelFactors, _lib, serving, tiles

TOL = 1e-5
ALL = 1 << 62


# ---- CPU ---------------------------------------------------------------------------------------
def test_rule_layout_matches_c_abi():
    assert C.sizeof(_lib.IfmTileRule) == 24
    assert _lib.IfmTileRule.max_pixels.offset == 16


def test_rules_are_ordered_by_dtype_kernel_then_size():
    ch = {("BFloat16", 17, 100): 64, ("Float32", 64, 100): 32, ("Float32", 9, ALL): 64, ("Float32", 9, 100): 32}
    r = tiles.rules_from_choices(ch)
    assert [(x["dtype"], x["kmax"], x["max_pixels"], x["tile_h"]) for x in r] == [
        (2, 9, 100, 32), (2, 9, ALL, 64), (2, 64, 100, 32), (4, 17, 100, 64)]
    assert all(x["op"] == _lib.TILE_SEP2D for x in r)
    with pytest.raises(ValueError):
        tiles.rules_from_choices({("Float32", 9, 100): 48})


def test_buckets_cover_every_call():
    assert tiles.KBUCKETS[-1][0] >= 64          # the longest factor the fused kernel takes
    assert tiles.SIZEBUCKETS[-1][0] >= 1 << 62
    for (kmax, k) in tiles.KBUCKETS:
        assert k <= kmax
    for (maxpx, side) in tiles.SIZEBUCKETS:
        assert side * side <= maxpx


# ---- GPU ---------------------------------------------------------------------------------------
@pytest.fixture
def restore_rules():
    yield
    if tiles.load() is None:  # back to this machine's cached rules, else the default tiles
        tiles.install([])


def _force(th):
    tiles.install(tiles.rules_from_choices({(dt, 64, ALL): th for dt in ("Float32", "BFloat16", "Float64",
                                                                         "Float16", "N0f8", "UInt8")}))


@pytest.mark.gpu
def test_tile_height_lookup_follows_rules(restore_rules):
    tiles.install([])
    assert tiles.tile_height("Float32", 13, 1024 * 1024) == 32
    tiles.install(tiles.rules_from_choices({("Float32", 17, 1536 ** 2): 64, ("Float32", 64, ALL): 32}))
    assert tiles.tile_height("Float32", 13, 1024 * 1024) == 64
    assert tiles.tile_height("Float32", 17, 1536 ** 2) == 64       # bounds are inclusive
    assert tiles.tile_height("Float32", 18, 1024 * 1024) == 32
    assert tiles.tile_height("Float32", 13, 1536 ** 2 + 1) == 32
    assert tiles.tile_height("BFloat16", 13, 1024 * 1024) == 32    # no rule for BF16
    L, ctx = _lib.load(), _lib.ctx()
    bad = (_lib.IfmTileRule * 1)()
    bad[0].op, bad[0].dtype, bad[0].kmax, bad[0].tile_h, bad[0].max_pixels = 1, 2, 9, 48, 100
    with pytest.raises(IF.ArgumentError):
        _lib.check(L.ifm_set_tile_rules(ctx, bad, 1))
    with pytest.raises(IF.ArgumentError):
        _lib.check(L.ifm_set_tile_rules(ctx, bad, _lib.MAX_TILE_RULES + 1))
    assert tiles.tile_height("Float32", 13, 1024 * 1024) == 64       # a rejected call keeps the rules


@pytest.mark.gpu
@pytest.mark.parametrize("border", ["replicate", "circular", "symmetric", "reflect", IF.Fill(0.3)])
def test_tall_tiles_bit_identical_and_oracle(border, restore_rules):
    from oracle import oracle as O
    rng = np.random.default_rng(31)
    for shape, sig in [((70, 97), (1.0, 2.0)), ((300, 65), (3.0, 0.8)), ((64, 128), (7.5, 7.5)),
                       ((130, 33), (0.6, 5.0)), ((257, 200), (15.0, 2.0))]:
        for et in ("f32", "bf16"):
            x = rng.random(shape).astype(np.float32)
            img = IF.bf16(x) if et == "bf16" else x
            kern = KernelFactors.gaussian(sig)
            _force(32)
            short = np.asarray(IF.imfilter(img, kern, border))
            _force(64)
            tall = np.asarray(IF.imfilter(img, kern, border))
            np.testing.assert_array_equal(tall, short, err_msg="%r %s %r" % (shape, et, sig))
            x64 = (x if et == "f32" else IF.bf16_to_float32(img)).astype(np.float64)
            # a Fill value is converted to the image's eltype first (src/borderarray.jl:11-24)
            fv = float(np.float32(0.3)) if et == "f32" else float(IF.bf16_to_float32(IF.bf16(0.3)))
            ob = ("fill", fv) if isinstance(border, IF.Fill) else border
            want = O.imfilter_factors(x64, [(f.Npre, f.data.data.astype(np.float64), f.data.first[0]) for f in kern],
                                      ob)
            y = np.asarray(tall, dtype=np.float64)
            assert float(np.max(np.abs(y - want)) / np.max(np.abs(want))) <= TOL, (shape, et, sig)


@pytest.mark.gpu
def test_tall_tiles_batched_bit_identical(restore_rules):
    from test_serving import _buffers, _outputs, _small_mix
    mix = _small_mix(24, seed=17)
    mix["op"][:] = np.where(np.arange(24) % 2 == 0, serving.OPS.index("gaussian"), serving.OPS.index("sobel"))
    bufs, keep = _buffers(mix)
    arr, k2 = serving.requests(mix, bufs)
    L, ctx = _lib.load(), _lib.ctx()
    res = {}
    for th in (32, 64):
        _force(th)
        _lib.check(L.ifm_batch_run(ctx, arr, mix["n"]))
        _lib.check(L.ifm_ctx_synchronize(ctx))
        res[th] = _outputs(keep)
    for i in range(mix["n"]):
        np.testing.assert_array_equal(res[64][i][0], res[32][i][0], err_msg="request %d" % i)
        if res[32][i][1] is not None:
            np.testing.assert_array_equal(res[64][i][1], res[32][i][1])


def test_damaged_cache_is_ignored(tmp_path, monkeypatch):
    p = tmp_path / "tiles.json"
    p.write_text("{not json")
    installed = []
    monkeypatch.setattr(tiles, "install", lambda rules: installed.append(rules))
    with pytest.warns(UserWarning):
        assert tiles.load(str(p)) is None
    p.write_text('{"rules": [{"op": 1, "dtype": 2, "kmax": 9, "tile_h": 64, "max_pixels": 100}]}')
    assert tiles.load(str(p))["rules"][0]["tile_h"] == 64
    assert installed and installed[-1][0]["kmax"] == 9
    assert tiles.load(str(tmp_path / "absent.json")) is None
