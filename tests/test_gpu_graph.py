"""GPU tests of CC iterations captured as graphs.

One iteration: a packed FP64 batch, a table-kernel batch (result blocks cut
into pair groups), a split-precision batch, a single contraction, a permute, a seeded fill and an energy sum, all
on one stream.  It is run eagerly once, then captured and replayed three
times; before every replay the blocks are reset from host memory and other
eager work runs on the same stream, and every replay must reproduce the
eager results bit for bit (the same kernels, the same fixed summation
orders).  Planning, allocation and host copies inside a capture fail with
SIPGPU_ECAPTURE and leave the capture usable.
"""
import numpy as np
import pytest

from aces4_amd import sipgpu as sg
from aces4_amd import workloads as wl

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager(monkeypatch):
    s = sg.Stream()
    w1 = wl.ring(o_segs=(9, 6), v_segs=(14, 11, 8))
    w2 = wl.ladder()
    w3 = wl.ring(o_segs=(5, 4), v_segs=(10, 7), seed=7)
    rng = np.random.default_rng(3)
    for w in (w1, w2, w3):
        w["D"].host[:] = rng.standard_normal(w["D"].host.size)
    monkeypatch.setenv("SIPGPU_FP64_PATH", "packed")
    b_packed = wl.batch_for(w1)
    monkeypatch.delenv("SIPGPU_FP64_PATH")
    # the table kernel with the ladder's result blocks cut into pair groups
    # (a GEMM launch, then the groups' sums)
    b_tab = sg.plan_choice(wl.contractions_for(w2), "tables", tail=1)
    b_split = wl.batch_for(w3, moduli=13)
    assert (b_packed.kind, b_tab.kind, b_split.kind) == ("packed", "tables", "split")
    assert b_tab.ntiles > 64
    # a single contraction, a permute, a fill and an energy sum on their own buffers
    n = 24
    A = sg.DeviceArray.from_host(rng.standard_normal(n * n))
    B = sg.DeviceArray.from_host(rng.standard_normal(n * n))
    C = sg.DeviceArray(n * n)
    db0 = next(iter(w1["D"].block_ids()))
    dims0 = [int(x) for x in w1["D"].dims(db0)]
    P = sg.DeviceArray(int(np.prod(dims0)))
    F = sg.DeviceArray(1000)
    eps = [sg.DeviceArray.from_host(rng.uniform(0.5, 2.0, d)) for d in dims0]
    out = sg.DeviceArray(1)
    ptrn = [1, -1, -2, 2]  # D(i,j) = L(i,k) R(k,j)
    # block operations: a slice of the permuted block, its squared norm, and a
    # two-part sum with its energy
    sdims = [max(1, d // 2) for d in dims0]
    Sl = sg.DeviceArray(int(np.prod(sdims)))
    N2 = sg.DeviceArray(1)
    Q = sg.DeviceArray.from_host(rng.standard_normal(int(np.prod(dims0))))
    T2 = sg.DeviceArray(int(np.prod(dims0)))
    out2 = sg.DeviceArray(1)

    def iteration(st):
        b_packed.execute(st)
        b_tab.execute(st)
        b_split.execute(st)
        sg.contract(ptrn, A.ptr, [n, n], B.ptr, [n, n], C.ptr, [n, n], stream=st)
        sg.permute(dims0, [1, 4, 3, 2, 1], w1["D"].addr(db0), P.ptr, stream=st)
        sg.fill_random(F.ptr, 1000, 11, 3, "normal", 0.5, s=st)
        sg.energy_sum(dims0, w1["D"].addr(db0), P.ptr, [e.ptr for e in eps], [1, -1, 1, -1],
                      0.5, out.ptr, s=st)
        sg.block_slice(P.ptr, dims0, Sl.ptr, sdims, [1] * len(dims0), stream=st)
        sg.block_norm2(P.ptr, dims0, N2.ptr, stream=st)
        sg.block_init(T2.ptr, dims0, 0.0, stream=st)
        sg.block_add(T2.ptr, P.ptr, dims0, 0.5, stream=st)
        sg.energy_sum_parts(dims0, T2.ptr, [Q.ptr, P.ptr], [e.ptr for e in eps], [1, -1, 1, -1],
                            0.25, out2.ptr, s=st)

    def reset():
        for w in (w1, w2, w3):
            w["D"].dev.upload(w["D"].host)
        C.zero()
        out.zero()
        out2.zero()

    def results():
        return [w1["D"].dev.download(), w2["D"].dev.download(), w3["D"].dev.download(),
                C.download(), P.download(), F.download(), out.download(), Sl.download(),
                N2.download(), T2.download(), out2.download()]

    reset()
    iteration(s)
    s.synchronize()
    want = results()
    g = sg.Graph.capture(iteration, s)
    assert g.nodes >= 9
    for rep in range(3):
        reset()
        b_tab.execute(s)  # eager work on the stream between replays
        s.synchronize()
        reset()
        g.launch(s)
        s.synchronize()
        got = results()
        for k, (x, y) in enumerate(zip(got, want)):
            assert np.array_equal(x, y), (rep, k)
    g.destroy()
    for b in (b_packed, b_tab, b_split):
        b.destroy()
    s.destroy()


def test_capture_refuses_planning_and_host_copies():
    s = sg.Stream()
    w = wl.ring(o_segs=(5,), v_segs=(6, 4))
    b = wl.batch_for(w)
    x = sg.DeviceArray(16)
    sg.Graph.begin(s)
    b.execute(s)
    for call in (lambda: wl.batch_for(w), lambda: x.zero(), lambda: x.download(),
                 lambda: sg.contract([1, -1, -2, 2], x.ptr, [3, 5], x.ptr, [5, 3], x.ptr, [3, 3],
                                     stream=s)):
        with pytest.raises(sg.SipGpuError) as ei:
            call()
        assert ei.value.code == -111
    g = sg.Graph.end(s)
    assert g.nodes >= 1
    with pytest.raises(sg.SipGpuError):
        sg.Graph.end(s)  # not capturing any more
    w["D"].dev.zero()
    g.launch(s)
    s.synchronize()
    ref = w["D"].dev.download()
    w["D"].dev.zero()
    b.execute(s)
    s.synchronize()
    assert np.array_equal(w["D"].dev.download(), ref)
    g.destroy()
    b.destroy()
    s.destroy()


def test_graph_of_two_stream_triples_tuple():
    """A triples tuple whose batches alternate over two streams (fork / join
    with events) captured from the caller's stream: every replay reproduces
    the eager W and energy bit for bit."""
    from aces4_amd import triples as tp
    w = tp.triples(o_segs=(2, 3), v_segs=(4, 3, 5), seed=5)
    tp.to_device(w)
    tup = (0, 1, 2, 0, 1, 0)
    n = int(np.prod(tp.tuple_dims(w, tup)))
    wbuf = sg.DeviceArray(n)
    parts = [sg.DeviceArray(n) for _ in range(5)]
    out = sg.DeviceArray(1)
    te = tp.TupleEnergy(w, tup, wbuf.ptr, out.ptr, parts=[p.ptr for p in parts], two_streams=True)
    s = sg.Stream()
    out.zero()
    te.execute(s)
    s.synchronize()
    want_w, want_e = wbuf.download(), out.download()[0]
    g = sg.Graph.capture(lambda st: te.execute(st), s)
    for rep in range(3):
        wbuf.zero()
        out.zero()
        s.synchronize()
        g.launch(s)
        s.synchronize()
        assert np.array_equal(wbuf.download().view(np.uint64), want_w.view(np.uint64)), rep
        assert out.download()[0] == want_e, rep
    g.destroy()
    te.destroy()
    s.destroy()


def test_graph_of_split_single_contraction():
    """A single contraction that runs split along K (one 64 x 64 tile, K =
    1000: four segments, three in the stream's workspace, then summed in
    order) captured and replayed three times, with other single contractions
    on the same stream between the replays: every replay bitwise equal to the
    eager result, and the graph holds the split GEMM and the summing pass."""
    s = sg.Stream()
    rng = np.random.default_rng(41)
    ptrn = [-1, 1, -1, 2]
    ld, rd, dd = [1000, 60], [1000, 50], [60, 50]
    L, R = rng.standard_normal(60000), rng.standard_normal(50000)
    D0 = rng.standard_normal(3000)
    dl, dr, d = (sg.DeviceArray.from_host(x) for x in (L, R, D0))
    other = sg.DeviceArray(3000)
    sg.contract(ptrn, dl.ptr, ld, dr.ptr, rd, d.ptr, dd, accumulate=True, stream=s)
    s.synchronize()
    eager = d.download()
    sg.Graph.begin(s)
    sg.contract(ptrn, dl.ptr, ld, dr.ptr, rd, d.ptr, dd, accumulate=True, stream=s)
    g = sg.Graph.end(s)
    assert g.nodes >= 2
    for rep in range(3):
        print("split graph replay", rep, flush=True)
        d.upload(D0)
        # eager split work on the capture stream between replays (it shares
        # the workspace; the stream orders it)
        sg.contract(ptrn, dl.ptr, ld, dr.ptr, rd, other.ptr, dd, stream=s)
        g.launch(s)
        s.synchronize()
        assert np.array_equal(d.download().view(np.uint64), eager.view(np.uint64)), rep
    g.destroy()
    for x in (dl, dr, d, other):
        x.free()
    s.destroy()


def test_split_workspace_growth_refused_during_capture():
    """While this thread captures stream A, a K-split contraction on a fresh
    stream B (whose workspace would have to grow) is refused with
    SIPGPU_ECAPTURE instead of allocating or running unsplit; A's capture
    stays usable, and after it the same contraction on B runs split and
    matches the oracle.  A shape warmed on A and captured on a fresh stream
    is refused the same way (a replay could not reproduce the eager bits)."""
    from checker import Oracle
    from cases import rel_fro
    a, b = sg.Stream(), sg.Stream()
    rng = np.random.default_rng(5)
    ptrn = [-1, 1, -1, 2]
    ld, rd, dd = [1000, 60], [1000, 50], [60, 50]
    L, R = rng.standard_normal(60000), rng.standard_normal(50000)
    dl, dr = sg.DeviceArray.from_host(L), sg.DeviceArray.from_host(R)
    d = sg.DeviceArray.from_host(np.zeros(3000))
    w = wl.ring(o_segs=(5,), v_segs=(6, 4))
    bt = wl.batch_for(w)
    sg.contract(ptrn, dl.ptr, ld, dr.ptr, rd, d.ptr, dd, stream=a)  # warm the shape on A
    a.synchronize()
    sg.Graph.begin(a)
    bt.execute(a)
    with pytest.raises(sg.SipGpuError) as ei:
        sg.contract(ptrn, dl.ptr, ld, dr.ptr, rd, d.ptr, dd, stream=b)
    assert ei.value.code == -111
    g = sg.Graph.end(a)
    assert g.nodes >= 1
    w["D"].dev.zero()
    g.launch(a)
    a.synchronize()
    ref = w["D"].dev.download()
    w["D"].dev.zero()
    bt.execute(a)
    a.synchronize()
    assert np.array_equal(w["D"].dev.download(), ref)
    # capture on a fresh stream C of a shape warmed on A only
    c = sg.Stream()
    sg.Graph.begin(c)
    bt.execute(c)
    with pytest.raises(sg.SipGpuError) as ei:
        sg.contract(ptrn, dl.ptr, ld, dr.ptr, rd, d.ptr, dd, stream=c)
    assert ei.value.code == -111
    sg.Graph.end(c).destroy()
    # eager on B: grows B's workspace, runs split, matches the oracle
    sg.contract(ptrn, dl.ptr, ld, dr.ptr, rd, d.ptr, dd, stream=b)
    b.synchronize()
    want, e = Oracle().contract(ptrn, L, ld, R, rd, dd)
    assert e == 0 and rel_fro(d.download(), want) <= 1e-13
    g.destroy()
    bt.destroy()
    for x in (dl, dr, d):
        x.free()
    for s in (a, b, c):
        s.destroy()


def test_cache_calls_refused_during_capture():
    """Acquire, release, read and flush of the block cache inside a capture
    fail with SIPGPU_ECAPTURE (an upload, memset or fence recorded into a
    graph would bake in an arena offset a later eviction reuses); the
    capture stays usable and the cache works after it.  A ZERO block is
    cleared by a kernel, not a memset."""
    s = sg.Stream()
    cache = sg.Cache(1 << 20)
    cache.insert(1, 1000, init="zero")
    w = wl.ring(o_segs=(5,), v_segs=(6, 4))
    bt = wl.batch_for(w)
    sg.Graph.begin(s)
    bt.execute(s)
    for call in (lambda: cache.acquire(1, "rw", s), lambda: cache.acquire(1, "r"),
                 lambda: cache.read(1, 1000), lambda: cache.flush()):
        with pytest.raises(sg.SipGpuError) as ei:
            call()
        assert ei.value.code == -111
    g = sg.Graph.end(s)
    assert g.nodes >= 1
    g.launch(s)
    s.synchronize()
    dev = cache.acquire(1, "r", s)
    assert dev
    cache.release(1, s)
    s.synchronize()
    assert np.array_equal(cache.read(1, 1000), np.zeros(1000))
    g.destroy()
    bt.destroy()
    cache.destroy()
    s.destroy()
