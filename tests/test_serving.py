"""Batched serving: the request stream and the C request layout (CPU), and on
the GPU the batched results against the single-image kernels (bit-exact, alone
and in the mix), against the oracle (1e-5 normwise / bit-exact medians), and
the admission loop under a small HBM budget."""
import ctypes as C

import numpy as np
import pytest

import imagefiltering as IF
from imagefiltering import _lib, serving
from imagefiltering.eltypes import BFloat16, Float32, Float64

TOL = 1e-5


# ---- CPU ---------------------------------------------------------------------------------------
def test_request_layout_matches_c_abi():
    assert C.sizeof(_lib.IfmRequest) == 544
    assert C.sizeof(_lib.IfmServeConfig) == 24
    assert _lib.IfmRequest.in_.offset == 8
    assert _lib.IfmRequest.out.offset == 96
    assert _lib.IfmRequest.border.offset == 184
    assert _lib.IfmRequest.factors.offset == 264
    assert _lib.IfmRequest.kernel.offset == 312
    assert _lib.IfmRequest.window.offset == 392


def test_request_mix_is_seeded_and_in_range():
    a = serving.request_mix(500, rate=1000.0, seed=7)
    b = serving.request_mix(500, rate=1000.0, seed=7)
    for k in ("w", "h", "dtype", "op", "border", "sigma", "arrival"):
        np.testing.assert_array_equal(a[k], b[k])
    assert a["w"].min() >= 64 and a["w"].max() <= 4096
    assert np.all(np.diff(a["arrival"]) >= 0)
    assert set(np.unique(a["op"])) == {0, 1, 2, 3}
    assert set(np.unique(a["border"])) == {0, 1, 2, 3, 4}
    assert 0.5 <= a["sigma"].min() and a["sigma"].max() <= 8.0
    # log-uniform sides: the mean side of [64, 4096] is (4096 - 64) / ln 64 = 969.5
    big = serving.request_mix(20000, seed=1)
    assert abs(big["w"].mean() - 969.5) < 30


def test_requests_struct_fields():
    mix = serving.request_mix(40, seed=3)
    arr, keep = serving.requests(mix)
    for i in range(40):
        q = arr[i]
        op = serving.OPS[mix["op"][i]]
        assert q.in_.size[0] == mix["w"][i] and q.in_.size[1] == mix["h"][i]
        assert q.in_.dtype == (_lib.BF16 if mix["dtype"][i] else _lib.F32)
        if op == "median7":
            assert q.op == _lib.OP_MEDIAN and q.window == 7 and q.out.dtype == q.in_.dtype
        elif op == "dense15":
            assert q.op == _lib.OP_DENSE and q.kernel.size[0] == 15 and q.out.dtype == _lib.F64
            taps = np.ctypeslib.as_array(q.kernel.taps, (225,))
            assert abs(taps.sum() - 1.0) < 1e-12
        elif op == "gaussian":
            s = mix["sigma"][i]
            assert q.op == _lib.OP_SEPARABLE and q.nfactors == 2
            assert q.factors[0].len == 4 * int(np.ceil(s)) + 1 and q.factors[1].axis == 1
        else:
            assert q.nfactors == 2 and q.nfactors2 == 2 and q.out2.dtype == _lib.F64
    assert serving.request_bytes(mix).sum() > 0


# ---- GPU ---------------------------------------------------------------------------------------
def _small_mix(n, seed):
    mix = serving.request_mix(n, seed=seed, side=(20, 300))
    mix["op"][:4] = [0, 1, 2, 3]                         # every op at least once
    nb = min(5, n)
    mix["border"][:nb] = np.arange(nb)
    return mix


def _buffers(mix):
    """Device buffers per request, inputs filled with the request's seed."""
    L, ctx = _lib.load(), _lib.ctx()
    bufs, keep = [], []
    for i in range(mix["n"]):
        shape = (int(mix["w"][i]), int(mix["h"][i]))
        et = BFloat16 if mix["dtype"][i] else Float32
        op = serving.OPS[mix["op"][i]]
        din = IF.DeviceArray(shape, et)
        a = serving._array(shape, et, din.ptr)
        _lib.check(L.ifm_fill_uniform(ctx, C.byref(a), 1000 + i))
        oet = et if op == "median7" else Float64
        dout = IF.DeviceArray(shape, oet)
        dout2 = IF.DeviceArray(shape, Float64) if op == "sobel" else None
        keep += [din, dout, dout2]
        bufs.append((din.ptr, dout.ptr, dout2.ptr if dout2 is not None else None))
    return bufs, keep


def _outputs(keep):
    return [(keep[3 * i + 1].to_host(), keep[3 * i + 2].to_host() if keep[3 * i + 2] is not None else None)
            for i in range(len(keep) // 3)]


@pytest.mark.gpu
def test_batched_equals_alone_and_percall_and_oracle():
    from oracle import oracle as O
    mix = _small_mix(28, seed=11)
    mix["arrival"] = np.linspace(0.0, 0.02, mix["n"])
    bufs, keep = _buffers(mix)
    serving.serve(mix, "batched", budget_bytes=1 << 30, max_batch=8, buffers=bufs)
    mixed = _outputs(keep)
    inputs = [keep[3 * i].to_host() for i in range(mix["n"])]
    # each request alone through ifm_batch_run
    L, ctx = _lib.load(), _lib.ctx()
    arr, k2 = serving.requests(mix, bufs)
    for i in range(mix["n"]):
        one = (_lib.IfmRequest * 1)(arr[i])
        _lib.check(L.ifm_batch_run(ctx, one, 1))
        _lib.check(L.ifm_ctx_synchronize(ctx))
        alone = (keep[3 * i + 1].to_host(), keep[3 * i + 2].to_host() if keep[3 * i + 2] is not None else None)
        np.testing.assert_array_equal(np.asarray(alone[0]), np.asarray(mixed[i][0]), err_msg="request %d" % i)
        if alone[1] is not None:
            np.testing.assert_array_equal(alone[1], mixed[i][1])
    # per-call dispatch through the single-image entry points
    serving.serve(mix, "percall", budget_bytes=1 << 30, buffers=bufs)
    percall = _outputs(keep)
    for i in range(mix["n"]):
        np.testing.assert_array_equal(np.asarray(percall[i][0]), np.asarray(mixed[i][0]), err_msg="request %d" % i)
    # the oracle
    for i in range(mix["n"]):
        op = serving.OPS[mix["op"][i]]
        b = serving.BORDERS[mix["border"][i]]
        ob = ("fill", 0.0) if b == "fill0" else b
        x = inputs[i]
        x64 = (IF.bf16_to_float32(x) if mix["dtype"][i] else np.asarray(x)).astype(np.float64)
        if op == "median7":
            want = O.median(x64, [-3, -3], [7, 7], ob)
            got = (IF.bf16_to_float32(mixed[i][0]) if mix["dtype"][i] else np.asarray(mixed[i][0])).astype(np.float64)
            np.testing.assert_array_equal(got, want, err_msg="median request %d" % i)
            continue
        if op == "dense15":
            k = serving.dense15(int(mix["kseed"][i]))
            want = O.imfilter_dense(x64, k, [-7, -7], ob)
            e = float(np.max(np.abs(mixed[i][0] - want)) / np.max(np.abs(want)))
            assert e <= TOL, (i, op, b, e)
            continue
        if op == "gaussian":
            s = float(mix["sigma"][i])
            parts = [(IF.KernelFactors.gaussian((s, s)), mixed[i][0])]
        else:
            parts = [(IF.KernelFactors.sobel((True, True), 0), mixed[i][0]),
                     (IF.KernelFactors.sobel((True, True), 1), mixed[i][1])]
        for kern, got in parts:
            fs = [(f.Npre, np.asarray(f.data.data, np.float64).ravel(), int(f.data.first[0])) for f in kern]
            want = O.imfilter_factors(x64, fs, ob)
            e = float(np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1e-300))
            assert e <= TOL, (i, op, b, e)


@pytest.mark.gpu
def test_server_submit_poll_alone_equals_mix():
    """ifm_server_submit / poll: every request's result is the same submitted
    alone (submit, wait) and submitted in a stream with the others (submits
    interleaved with polls, requests joining batches while earlier ones run),
    bit for bit, and equal to ifm_batch_run's; a small budget and batch
    limit force several batches in flight."""
    mix = _small_mix(24, seed=21)
    bufs, keep = _buffers(mix)
    arr, k2 = serving.requests(mix, bufs)
    n = mix["n"]
    alone = []
    with serving.Server(budget_bytes=0, max_batch=8) as srv:
        for i in range(n):
            t = srv.submit(arr[i])
            srv.wait(t)
            assert srv.poll(t)
            alone.append(_outputs(keep)[i])
    for i in range(n):  # scramble the outputs so the stream has to rewrite them
        _lib.check(_lib.load().ifm_memset_d(_lib.ctx(), keep[3 * i + 1].ptr, 0xFF, keep[3 * i + 1].nbytes))
    need = int(serving.request_bytes(mix).max())
    with serving.Server(budget_bytes=3 * need, max_batch=5, max_inflight=3) as srv:
        tickets = []
        for i in range(n):
            tickets.append(srv.submit(arr[i]))
            srv.poll(tickets[0])
        assert tickets == sorted(tickets) and len(set(tickets)) == n
        srv.wait(tickets[n // 2])
        assert all(srv.poll(t) for t in tickets[:n // 2 + 1])
        srv.wait()
        assert all(srv.poll(t) for t in tickets)
    mixed = _outputs(keep)
    for i in range(n):
        np.testing.assert_array_equal(np.asarray(mixed[i][0]), np.asarray(alone[i][0]), err_msg="request %d" % i)
        if alone[i][1] is not None:
            np.testing.assert_array_equal(mixed[i][1], alone[i][1])
    L, ctx = _lib.load(), _lib.ctx()
    _lib.check(L.ifm_batch_run(ctx, arr, n))
    _lib.check(L.ifm_ctx_synchronize(ctx))
    batch = _outputs(keep)
    for i in range(n):
        np.testing.assert_array_equal(np.asarray(batch[i][0]), np.asarray(mixed[i][0]), err_msg="request %d" % i)


@pytest.mark.gpu
def test_server_rejects_bad_requests_at_submit():
    mix = _small_mix(4, seed=2)
    bufs, keep = _buffers(mix)
    arr, k2 = serving.requests(mix, bufs)
    with serving.Server() as srv:
        bad = (_lib.IfmRequest * 1)(arr[2])
        bad[0].window = 9
        with pytest.raises(NotImplementedError):
            srv.submit(bad[0])
        t = srv.submit(arr[0])
        srv.wait(t)
    with serving.Server(budget_bytes=16) as srv:
        with pytest.raises(RuntimeError, match="budget"):
            srv.submit(arr[0])


@pytest.mark.gpu
def test_serve_generated_inputs_under_a_small_budget():
    mix = serving.request_mix(60, rate=5000.0, seed=5, side=(64, 700))
    budget = 24 << 20
    need = serving.arena_bytes(mix)
    keep = need <= budget
    for k in ("w", "h", "dtype", "op", "border", "sigma", "kseed", "arrival"):
        mix[k] = mix[k][keep]
    mix["n"] = int(keep.sum())
    for mode in ("batched", "percall"):
        admit, done = serving.serve(mix, mode, budget_bytes=budget, max_batch=16)
        assert np.all(admit >= mix["arrival"] - 1e-9)
        assert np.all(done >= admit)
        st = serving.stats(mix, done)
        assert st["images_per_s"] > 0


@pytest.mark.gpu
def test_serve_request_filling_the_whole_budget():
    """A two-output request (sobel) whose 256-byte-rounded arrays fill the
    budget exactly is served in both modes; one granule less is refused with
    a budget error instead of waiting forever."""
    mix = serving.request_mix(3, rate=None, seed=9, side=(97, 131))
    mix["op"][:] = serving.OPS.index("sobel")
    need = serving.arena_bytes(mix)
    budget = int(need.max())
    assert budget > serving.request_bytes(mix).max()   # the rounding matters here
    for mode in ("batched", "percall"):
        admit, done = serving.serve(mix, mode, budget_bytes=budget, max_batch=4)
        assert np.all(done >= admit) and np.all(done > 0)
    for mode in ("batched", "percall"):
        with pytest.raises(RuntimeError, match="budget"):
            serving.serve(mix, mode, budget_bytes=budget - 256, max_batch=4)


@pytest.mark.gpu
def test_batch_rejects_bad_requests():
    mix = _small_mix(4, seed=2)
    bufs, keep = _buffers(mix)
    arr, k2 = serving.requests(mix, bufs)
    L, ctx = _lib.load(), _lib.ctx()
    bad = (_lib.IfmRequest * 1)(arr[0])
    bad[0].out.size[0] += 1                      # rejected before any launch
    bad[0].out.stride[1] = bad[0].out.size[0]
    with pytest.raises(IF.DimensionMismatch):
        _lib.check(L.ifm_batch_run(ctx, bad, 1))
    bad = (_lib.IfmRequest * 1)(arr[2])
    bad[0].window = 9
    with pytest.raises(NotImplementedError):
        _lib.check(L.ifm_batch_run(ctx, bad, 1))


@pytest.mark.gpu
def test_schedule_autotuner_picks_a_tried_schedule():
    mix = serving.request_mix(80, seed=11, side=(32, 400))
    sched, table = serving.autotune(mix, budget_bytes=1 << 30, sample=40, batches=(8, 32), inflight=(1, 3))
    assert len(table) == 4
    assert {(r["max_batch"], r["max_inflight"]) for r in table} == {(8, 1), (8, 3), (32, 1), (32, 3)}
    best = max(r["images_per_s"] for r in table)
    chosen = [r for r in table if (r["max_batch"], r["max_inflight"]) == (sched["max_batch"], sched["max_inflight"])][0]
    assert chosen["images_per_s"] >= 0.98 * best
    # no pair within the slack has fewer batches in flight, or as few and smaller batches
    for r in table:
        if r["images_per_s"] >= 0.98 * best:
            assert (r["max_inflight"], r["max_batch"]) >= (chosen["max_inflight"], chosen["max_batch"])
    admit, done = serving.serve(mix, "batched", budget_bytes=1 << 30, **sched)
    assert np.all(done >= admit)
