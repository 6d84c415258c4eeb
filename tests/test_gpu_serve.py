"""GPU tests of the device block cache and the continuous-batching server.

Cache: values survive eviction and refill bit for bit (clean blocks are
re-read from their homes, dirty ones are spilled first), LRU order decides
the victims, pinned blocks are never evicted.  Server: the results of a
stream of single block contractions equal the SIP's sequential oracle
(tensor_block_contract into a temporary, then block add), with blocks
evicted and spilled while the stream runs; dependent requests keep
submission order.
"""
import ctypes

import numpy as np
import pytest

from cases import rel_fro
from checker import Oracle

from aces4_amd import sipgpu as sg
from aces4_amd import workloads as wl

pytestmark = pytest.mark.gpu
o = Oracle()
MB = 1 << 20


def d2h(ptr, n):
    out = np.empty(n)
    sg._check(sg.lib().sipgpu_memcpy_d2h(out.ctypes.data, ptr, n * 8))
    return out


def h2d(ptr, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    sg._check(sg.lib().sipgpu_memcpy_h2d(ptr, x.ctypes.data, x.nbytes))


def test_cache_hit_miss_evict_refill():
    n = MB // 8  # 1 MiB blocks
    c = sg.Cache(int(2.5 * MB))
    homes = {k: wl.uniform(k, n) for k in (1, 2, 3)}
    for k, h in homes.items():
        c.insert(k, n, host=h)
    for k in (1, 2, 3, 1):
        p = c.acquire(k, "r")
        assert np.array_equal(d2h(p, n), homes[k])
        c.release(k)
    st = c.stats()
    assert st["misses"] == 4 and st["evictions"] == 2 and st["spills"] == 0
    assert st["resident_bytes"] <= st["budget_bytes"]
    c.destroy()


def test_cache_dirty_block_spills_and_returns():
    n = MB // 8
    c = sg.Cache(int(2.5 * MB))
    c.insert(1, n, init="zero")
    c.insert(2, n, init="uniform", seed=5, stream=1)
    c.insert(3, n, init="normal", seed=5, stream=2, a=0.5)
    p = c.acquire(1, "w")
    new = wl.uniform(99, n)
    h2d(p, new)
    c.release(1)
    for k in (2, 3):  # forces 1 out: it must be written back first
        c.acquire(k, "r")
        c.release(k)
    st = c.stats()
    assert st["spills"] == 1
    p = c.acquire(1, "r")
    assert np.array_equal(d2h(p, n), new)
    c.release(1)
    assert np.array_equal(c.read(2, n), wl.uniform(5, n, stream=1))
    c.destroy()


def test_cache_pinned_blocks_are_not_evicted():
    n = MB // 8
    c = sg.Cache(int(2.5 * MB))
    for k in (1, 2, 3):
        c.insert(k, n, init="zero")
    c.acquire(1, "r")
    c.acquire(2, "r")
    with pytest.raises(sg.SipGpuError) as ei:
        c.acquire(3, "r")
    assert ei.value.code == -109  # SIPGPU_EBUDGET
    c.release(1)
    c.acquire(3, "r")  # now 1 can go
    c.release(3)
    c.release(2)
    c.destroy()


def test_cache_scopes_remove_temps_without_write_back():
    """leave_scope semantics: the scope's temps are forgotten (no spill, no
    write to their homes), their HBM is reusable without evictions, other
    blocks and outer scopes are untouched."""
    n = MB // 8
    c = sg.Cache(int(4.5 * MB))         # all four blocks fit: nothing is evicted
    home = wl.uniform(7, n)
    keep = home.copy()
    h3 = home.copy()
    c.insert(1, n, host=home)           # an ordinary block
    c.scope_enter()
    c.insert(2, n, init="zero")
    c.insert(3, n, host=h3)
    c.scope_add(2)
    c.scope_add(3)
    c.scope_enter()                     # nested scope
    c.insert(4, n, init="zero")
    c.scope_add(4)
    for k in (1, 2, 3, 4):
        p = c.acquire(k, "w")
        h2d(p, np.full(n, float(k)))
        c.release(k)
    c.scope_leave()                     # removes 4 only
    with pytest.raises(sg.SipGpuError) as ei:
        c.read(4, n)
    assert ei.value.code == -110        # SIPGPU_ENOTFOUND
    p = c.acquire(2, "r")               # a pinned temp blocks leave_scope...
    with pytest.raises(sg.SipGpuError) as ei:
        c.scope_leave()
    assert ei.value.code == -107
    assert np.array_equal(d2h(p, n), np.full(n, 2.0))   # ...and nothing was removed
    c.release(2)
    st0 = c.stats()
    c.scope_leave()
    st = c.stats()
    assert st["blocks"] == 1 and st["spills"] == st0["spills"] and st["bytes_d2h"] == st0["bytes_d2h"]
    assert st["resident_bytes"] == MB
    assert np.array_equal(h3, keep)     # the dirty temp was not written back
    # the freed space takes three new blocks without evicting block 1
    for k in (5, 6, 7):
        c.insert(k, n, init="zero")
        c.acquire(k, "r")
        c.release(k)
    assert c.stats()["evictions"] == st["evictions"] == 0
    assert np.array_equal(c.read(1, n), np.full(n, 1.0))
    assert np.array_equal(home, keep)   # block 1 stayed resident: its home is untouched
    with pytest.raises(sg.SipGpuError):
        c.scope_leave()                 # no scope open
    c.destroy()


def test_cache_remove_and_persistent_blocks():
    """delete_block and set_persistent / end of program semantics."""
    n = MB // 8
    c = sg.Cache(int(2.5 * MB))
    c.insert(1, n, init="zero")
    c.insert(2, n, init="zero")
    c.insert(3, n, init="uniform", seed=3, stream=0)
    p = c.acquire(1, "w")
    h2d(p, np.full(n, 1.5))
    with pytest.raises(sg.SipGpuError) as ei:
        c.remove(1)                     # pinned
    assert ei.value.code == -107
    c.release(1)
    p = c.acquire(2, "w")
    h2d(p, np.full(n, 2.5))
    c.release(2)
    c.acquire(3, "r")                   # evicts and spills block 1
    c.release(3)
    assert c.stats()["spills"] == 1
    c.remove(3)
    with pytest.raises(sg.SipGpuError) as ei:
        c.remove(3)
    assert ei.value.code == -110
    c.set_persistent(1)                 # spilled, persistent
    c.set_persistent(2)                 # resident, persistent
    c.insert(4, n, init="zero")         # dies with the program
    assert c.end_program() == 2
    assert c.stats()["blocks"] == 2
    assert np.array_equal(c.read(1, n), np.full(n, 1.5))
    assert np.array_equal(c.read(2, n), np.full(n, 2.5))
    # the next program keeps only what it marks again
    c.set_persistent(2)
    assert c.end_program() == 1
    assert c.stats()["blocks"] == 1
    assert np.array_equal(c.read(2, n), np.full(n, 2.5))
    c.scope_enter()
    with pytest.raises(sg.SipGpuError):
        c.end_program()                 # a scope is open
    c.scope_leave()
    assert c.end_program() == 0 and c.stats()["blocks"] == 0
    assert c.stats()["resident_bytes"] == 0
    c.destroy()


def ring_keys(w):
    keys = {}
    for tag, arr in ((1, w["L"]), (2, w["R"]), (3, w["D"])):
        for j, blk in enumerate(arr.block_ids()):
            keys[(tag, blk)] = (tag << 40) + j
    return keys


@pytest.mark.parametrize("precision", ["fp64", "split", "mixed"])
@pytest.mark.parametrize("budget_frac", [2.0, 0.4], ids=["fits", "evicts"])
def test_server_matches_sequential_oracle(budget_frac, precision):
    """Plain FP64 requests (1e-13), split-precision ones over 14 moduli
    (1e-12), and a stream mixing both per request (batches break at each
    change of precision), against the sequential oracle."""
    w = wl.ring(o_segs=(5, 4), v_segs=(10, 10, 7))
    D, L, R, ptrn = w["D"], w["L"], w["R"], w["ptrn"]
    total = 8 * (D.host.size + L.host.size + R.host.size)
    c = sg.Cache(int(total * budget_frac))
    keys = ring_keys(w)
    homes = []
    for tag, arr in ((1, L), (2, R), (3, D)):
        for blk in arr.block_ids():
            h = arr.block(blk).copy()
            homes.append(h)
            c.insert(keys[(tag, blk)], h.size, host=h)
    srv = sg.Server(c, max_batch=64)
    rng = np.random.default_rng(2)
    order = rng.permutation(len(w["entries"]))
    mods = {"fp64": [0], "split": [14], "mixed": [0, 14]}[precision]
    tickets = []
    for idx in order:
        db, lb, rb = w["entries"][idx]
        req = sg.make_request(ptrn, keys[(1, lb)], L.dims(lb), keys[(2, rb)], R.dims(rb),
                              keys[(3, db)], D.dims(db), accumulate=True,
                              moduli=mods[int(rng.integers(len(mods)))])
        tickets.append(srv.submit(req))
    lat = [srv.wait(t) for t in tickets]
    assert all(x >= 0.0 for x in lat)
    st = srv.stats()
    assert st["requests"] == len(w["entries"]) and st["batches"] >= 1
    srv.destroy()
    worst = 0.0
    for db in D.block_ids():
        want = D.block(db).copy()
        for e in w["entries"]:
            if e[0] == db:
                tmp, _ = o.contract(ptrn, L.block(e[1]), L.dims(e[1]), R.block(e[2]),
                                    R.dims(e[2]), D.dims(db))
                want += tmp
        got = c.read(keys[(3, db)], want.size)
        worst = max(worst, rel_fro(got, want))
    if budget_frac < 1.0:
        assert c.stats()["evictions"] > 0
    c.destroy()
    assert worst <= (1e-13 if precision == "fp64" else 1e-12)


@pytest.mark.parametrize("moduli", [0, 14])
def test_server_result_same_alone_and_in_mix(moduli):
    """Each request's result computed alone (one request in flight) equals
    its result inside a stream of other requests batched with it, within the
    path's tolerance."""
    rng = np.random.default_rng(12)
    ptrn = [1, -1, -2, 2]  # D(i,j) = L(i,k) R(k,j)
    shapes = [(int(rng.integers(8, 64)), int(rng.integers(8, 64)), int(rng.integers(8, 64)))
              for _ in range(24)]
    c = sg.Cache(256 * MB)
    key = 1
    reqs = []
    for m, k, n in shapes:
        lk, rk, dk = key, key + 1, key + 2
        key += 3
        c.insert(lk, m * k, host=rng.standard_normal(m * k))
        c.insert(rk, k * n, host=rng.standard_normal(k * n))
        c.insert(dk, m * n, init="zero")
        reqs.append(sg.make_request(ptrn, lk, [m, k], rk, [k, n], dk, [m, n], moduli=moduli))
    srv = sg.Server(c, max_batch=64)
    alone = []
    for r in reqs:
        srv.wait(srv.submit(r))
        alone.append(c.read(r.dkey, r.ddims[0] * r.ddims[1]))
    ts = [srv.submit(r) for r in reqs]
    for t in ts:
        srv.wait(t)
    st = srv.stats()
    assert st["max_batch"] > 1
    tol = 1e-13 if moduli == 0 else 1e-12
    for r, a in zip(reqs, alone):
        assert rel_fro(c.read(r.dkey, r.ddims[0] * r.ddims[1]), a) <= tol
    srv.destroy()
    c.destroy()


def test_server_rejects_bad_moduli():
    c = sg.Cache(16 * MB)
    for k in (1, 2, 3):
        c.insert(k, 16, init="zero")
    srv = sg.Server(c)
    with pytest.raises(sg.SipGpuError):
        srv.submit(sg.make_request([1, -1, -2, 2], 1, [4, 4], 2, [4, 4], 3, [4, 4], moduli=5))
    srv.destroy()
    c.destroy()


def test_server_dependent_requests_keep_order():
    """X = A * B, then Y = X * C, then X = A * B' (overwrite): Y must see the
    first X, and the final X is the second product."""
    n = 24
    rng = np.random.default_rng(8)
    A, B, B2, C = (rng.standard_normal(n * n) for _ in range(4))
    c = sg.Cache(64 * MB)
    for k, h in ((1, A), (2, B), (3, B2), (4, C)):
        c.insert(k, n * n, host=h)
    c.insert(10, n * n, init="zero")  # X
    c.insert(11, n * n, init="zero")  # Y
    ptrn = [1, -1, -2, 2]  # D(i,j) = L(i,k) R(k,j)
    dims = [n, n]
    srv = sg.Server(c, max_batch=64)
    t = [srv.submit(sg.make_request(ptrn, 1, dims, 2, dims, 10, dims)),
         srv.submit(sg.make_request(ptrn, 10, dims, 4, dims, 11, dims)),
         srv.submit(sg.make_request(ptrn, 1, dims, 3, dims, 10, dims))]
    for x in t:
        srv.wait(x)
    srv.destroy()
    Am, Bm, B2m, Cm = (v.reshape(n, n, order="F") for v in (A, B, B2, C))
    X1 = Am @ Bm
    assert rel_fro(c.read(11, n * n), (X1 @ Cm).ravel(order="F")) <= 1e-13
    assert rel_fro(c.read(10, n * n), (Am @ B2m).ravel(order="F")) <= 1e-13
    c.destroy()


def test_server_rejects_self_hazard():
    c = sg.Cache(16 * MB)
    c.insert(1, 16, init="zero")
    c.insert(2, 16, init="zero")
    srv = sg.Server(c)
    with pytest.raises(sg.SipGpuError) as ei:
        srv.submit(sg.make_request([1, -1, -2, 2], 1, [4, 4], 2, [4, 4], 1, [4, 4]))
    assert ei.value.code == -108
    srv.destroy()
    c.destroy()


def test_server_replay_latencies():
    w = wl.ring(o_segs=(5,), v_segs=(6, 4))
    D, L, R, ptrn = w["D"], w["L"], w["R"], w["ptrn"]
    c = sg.Cache(64 * MB)
    keys = ring_keys(w)
    for tag, arr in ((1, L), (2, R), (3, D)):
        for blk in arr.block_ids():
            c.insert(keys[(tag, blk)], int(np.prod(arr.dims(blk))), init="zero")
    srv = sg.Server(c, max_batch=32)
    reqs = [sg.make_request(ptrn, keys[(1, lb)], L.dims(lb), keys[(2, rb)], R.dims(rb),
                            keys[(3, db)], D.dims(db), accumulate=True)
            for db, lb, rb in w["entries"]]
    arrivals = np.cumsum(np.random.default_rng(4).exponential(20e-6, len(reqs)))
    lat, wall = srv.replay(reqs, arrivals)
    assert np.all(lat > 0) and wall >= arrivals[-1]
    srv.destroy()
    c.destroy()


def test_server_large_requests_on_panels():
    """Requests of over 1e9 flop run as packed FP64 panel batches bound into
    the server's scratch buffer; each result matches the product to 1e-13,
    including one that accumulates."""
    rng = np.random.default_rng(14)
    shapes = [(1000, 1100, 900), (1030, 970, 1010)]  # (M, K, N): ~2e9 flop each
    c = sg.Cache(256 * MB)
    ptrn = [1, -1, -2, 2]  # D(i,j) = L(i,k) R(k,j)
    reqs, want, key = [], [], 1
    for idx, (m, k, n) in enumerate(shapes):
        L, R, D0 = rng.standard_normal(m * k), rng.standard_normal(k * n), rng.standard_normal(m * n)
        c.insert(key, m * k, host=L)
        c.insert(key + 1, k * n, host=R)
        c.insert(key + 2, m * n, host=D0)
        acc = idx == 1
        reqs.append(sg.make_request(ptrn, key, [m, k], key + 1, [k, n], key + 2, [m, n],
                                    accumulate=acc))
        prod = (L.reshape(m, k, order="F") @ R.reshape(k, n, order="F")).ravel(order="F")
        want.append(prod + (D0 if acc else 0.0))
        key += 3
    srv = sg.Server(c, max_batch=8)
    for t in [srv.submit(r) for r in reqs]:
        srv.wait(t)
    srv.destroy()
    for r, wv in zip(reqs, want):
        assert rel_fro(c.read(r.dkey, wv.size), wv) <= 1e-13
    c.destroy()
