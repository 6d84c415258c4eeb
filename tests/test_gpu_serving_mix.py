"""GPU parity of the serving stream: mixed ladder / ring / hole-hole requests
with segments 8..64 (the serving configuration's generator, seed 4, Poisson
arrivals) through the device block cache and the two-stream server.

Every touched result block is checked against the SIP's sequential oracle
(tensor_block_contract into a temporary, then R += TMP, in submission
order): plain FP64 within 1e-13, split precision (14 moduli) within 1e-12
relative Frobenius error, with the HBM budget fitting the working set and
with a budget that forces LRU eviction and spills.  The oracle's CPU time
bounds the request count: the stream's first requests of at most 2e9 flop
each.  Operands are generated in HBM by the seeded device generator and read
back for the oracle; results start at zero.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from cases import rel_fro
from checker import Oracle

from aces4_amd import sipgpu as sg
from aces4_amd import workloads as wl

pytestmark = pytest.mark.gpu
o = Oracle()
MAX_FLOP = 2e9
COUNT = 24


def _subset(moduli):
    """The stream's first requests of at most MAX_FLOP each, COUNT / 3 per
    term (a term is told by its operand arrays: the ladder and hole-hole
    terms share a pattern), in arrival order."""
    mix = wl.serving_mix(n_requests=400, seed=4, moduli=moduli)
    per, pick = {}, []
    for i, r in enumerate(mix["requests"]):
        term = (r.lkey >> 48, r.rkey >> 48)
        if mix["flops"][i] <= MAX_FLOP and per.get(term, 0) < COUNT // 3:
            per[term] = per.get(term, 0) + 1
            pick.append(i)
    reqs = [mix["requests"][i] for i in pick]
    arrivals = mix["arrivals"][pick] - mix["arrivals"][pick[0]]
    keys = set()
    for r in reqs:
        keys.update((r.lkey, r.rkey, r.dkey))
    blocks = {k: mix["blocks"][k] for k in keys}
    return mix, reqs, arrivals, blocks, len(per)


def _cache(mix, blocks, frac):
    total = 8 * sum(n for _, _, n in blocks.values())
    c = sg.Cache(max(int(total * frac), 64 << 20) if frac >= 1.0 else int(total * frac))
    sub = dict(mix, blocks=blocks)
    wl.populate_cache(c, sub)
    return c, total


def _dims(r, which):
    rank = getattr(r, which[0] + "rank")
    return [int(x) for x in getattr(r, which + "dims")[:rank]]


def _oracle(reqs, operands):
    """R += L*R per request in submission order (TMP then add), results from zero."""
    def one(r):
        ld, rd, dd = _dims(r, "l"), _dims(r, "r"), _dims(r, "d")
        tmp, e = o.contract(list(r.ptrn[:len(ld) + len(rd)]), operands[r.lkey], ld,
                            operands[r.rkey], rd, dd)
        assert e == 0
        return tmp
    with ThreadPoolExecutor(8) as ex:  # the oracle's C calls release the GIL
        tmps = list(ex.map(one, reqs))
    out = {}
    for r, t in zip(reqs, tmps):
        out[r.dkey] = out.get(r.dkey, 0.0) + t
    return out


_EXPECTED = {}


def _expected(mix, reqs, blocks):
    """Oracle results of the subset (operands read back from a cache filled
    by the device generator, so they are the values the server used)."""
    if "want" not in _EXPECTED:
        c, _ = _cache(mix, blocks, 2.0)
        dkeys = {r.dkey for r in reqs}
        operands = {}
        for k, (_, _, n) in blocks.items():
            if k in dkeys:
                continue
            c.acquire(k, "r")  # generated in HBM on first use
            c.release(k)
            operands[k] = c.read(k, n)
        c.destroy()
        _EXPECTED["want"] = _oracle(reqs, operands)
    return _EXPECTED["want"]


@pytest.mark.parametrize("moduli", [0, 14], ids=["fp64", "split14"])
@pytest.mark.parametrize("budget", [2.0, 0.45], ids=["fits", "spills"])
def test_serving_mix_matches_oracle(moduli, budget):
    mix, reqs, arrivals, blocks, nterms = _subset(moduli)
    assert nterms == 3 and len(reqs) == COUNT
    segs = mix["segs"]
    assert min(segs["v"] + segs["o"]) >= 8 and max(segs["v"] + segs["o"]) <= 64
    want = _expected(mix, reqs, blocks)
    c, total = _cache(mix, blocks, budget)
    srv = sg.Server(c, max_batch=64)
    print("serving mix", moduli, budget, "requests", len(reqs), "bytes", total, flush=True)
    # arrivals compressed 20x so requests queue behind running batches
    lat, wall = srv.replay(reqs, arrivals / 20.0)
    assert np.all(lat > 0)
    st = srv.stats()
    assert st["requests"] == len(reqs)
    lanes, _ = srv.lanes()
    assert lanes == 2
    srv.destroy()
    tol = 1e-13 if moduli == 0 else 1e-12
    worst = 0.0
    for dk, wv in want.items():
        got = c.read(dk, wv.size)
        worst = max(worst, rel_fro(got, wv))
    cs = c.stats()
    c.destroy()
    if budget < 1.0:
        assert cs["evictions"] > 0 and cs["spills"] > 0
    assert worst <= tol, worst


@pytest.mark.parametrize("moduli", [0, 14], ids=["fp64", "split14"])
def test_serving_mix_same_alone_and_in_mix(moduli):
    """Each result block computed with one request in flight at a time equals
    the one computed from the batched, two-stream replay of the stream, within
    the path's tolerance."""
    mix, reqs, arrivals, blocks, _ = _subset(moduli)
    dkeys = sorted({r.dkey for r in reqs})
    results = []
    for mode in ("alone", "mix"):
        c, _ = _cache(mix, blocks, 2.0)
        srv = sg.Server(c, max_batch=64)
        if mode == "alone":
            for r in reqs:
                srv.wait(srv.submit(r))
        else:
            srv.replay(reqs, np.zeros(len(reqs)))
            assert srv.stats()["max_batch"] > 1
        srv.destroy()
        results.append({k: c.read(k, blocks[k][2]) for k in dkeys})
        c.destroy()
    tol = 1e-13 if moduli == 0 else 1e-12
    for k in dkeys:
        assert rel_fro(results[1][k], results[0][k]) <= tol, k


def test_server_permutes_in_order_with_contractions():
    """X = A * B, then P = permute(X), then Y = P * C, then X = A * B2
    (overwrite), submitted back to back: the permute reads the first X, Y reads
    the permuted block, and the last X is the second product, each within
    1e-13 of the oracle's sequence."""
    n, m = 20, 12
    rng = np.random.default_rng(21)
    A = rng.standard_normal(n * m)
    B, B2 = rng.standard_normal(m * n), rng.standard_normal(m * n)
    C = rng.standard_normal(n * n)
    c = sg.Cache(64 << 20)
    for k, h in ((1, A), (2, B), (3, B2), (4, C)):
        c.insert(k, h.size, host=h)
    for k in (10, 11, 12):
        c.insert(k, n * n, init="zero")
    ptrn = [1, -1, -2, 2]  # D(i,j) = L(i,k) R(k,j)
    srv = sg.Server(c, max_batch=64)
    t = [srv.submit(sg.make_request(ptrn, 1, [n, m], 2, [m, n], 10, [n, n])),
         srv.submit_permute([n, n], [1, 2, 1], 10, 11),
         srv.submit(sg.make_request(ptrn, 11, [n, n], 4, [n, n], 12, [n, n])),
         srv.submit(sg.make_request(ptrn, 1, [n, m], 3, [m, n], 10, [n, n]))]
    for x in t:
        srv.wait(x)
    srv.destroy()
    X1, _ = o.contract(ptrn, A, [n, m], B, [m, n], [n, n])
    P, _ = o.copy([n, n], [1, 2, 1], X1)
    Y, _ = o.contract(ptrn, P, [n, n], C, [n, n], [n, n])
    X2, _ = o.contract(ptrn, A, [n, m], B2, [m, n], [n, n])
    assert rel_fro(c.read(11, n * n), P) <= 1e-13
    assert rel_fro(c.read(12, n * n), Y) <= 1e-13
    assert rel_fro(c.read(10, n * n), X2) <= 1e-13
    c.destroy()


def test_server_permute_is_bit_exact_and_checks_sizes():
    """A rank-4 permute through the server equals the oracle's bit for bit;
    extents that do not match the cached blocks, a non-permutation and an
    unknown key are refused at submission."""
    dims = [7, 5, 9, 4]
    x = np.random.default_rng(3).standard_normal(int(np.prod(dims)))
    c = sg.Cache(16 << 20)
    c.insert(1, x.size, host=x)
    c.insert(2, x.size, init="zero")
    srv = sg.Server(c)
    srv.wait(srv.submit_permute(dims, [1, 3, 1, 4, 2], 1, 2))
    want, e = o.copy(dims, [1, 3, 1, 4, 2], x)
    assert e == 0 and np.array_equal(c.read(2, x.size), want)
    for args, code in (((dims[:3], [1, 1, 2, 3], 1, 2), -107),
                       ((dims, [1, 1, 1, 3, 4], 1, 2), -103),
                       ((dims, [1, 2, 1, 3, 4], 1, 99), -110),
                       ((dims, [1, 2, 1, 3, 4], 1, 1), -108)):
        with pytest.raises(sg.SipGpuError) as ei:
            srv.submit_permute(*args)
        assert ei.value.code == code
    srv.destroy()
    c.destroy()


def test_server_dependency_chain_across_streams():
    """A chain X_{t+1} = X_t * C of eight requests, each depending on the one
    before, submitted at once: consecutive batches alternate streams, so each
    must wait for the other stream's batch on the device.  The result equals
    the oracle's chain within 1e-13."""
    n = 96
    rng = np.random.default_rng(5)
    X0 = rng.standard_normal(n * n)
    C = rng.standard_normal(n * n) / np.sqrt(n)
    c = sg.Cache(64 << 20)
    c.insert(1, n * n, host=X0)
    c.insert(2, n * n, host=C)
    for k in range(10, 18):
        c.insert(k, n * n, init="zero")
    ptrn = [1, -1, -2, 2]
    srv = sg.Server(c, max_batch=64)
    src = 1
    reqs = []
    for k in range(10, 18):
        reqs.append(sg.make_request(ptrn, src, [n, n], 2, [n, n], k, [n, n]))
        src = k
    tk = [srv.submit(r) for r in reqs]
    for t in tk:
        srv.wait(t)
    st = srv.stats()
    srv.destroy()
    assert st["batches"] == 8  # each request reads the block the one before writes
    want = X0
    for _ in range(8):
        want, _ = o.contract(ptrn, want, [n, n], C, [n, n], [n, n])
    assert rel_fro(c.read(17, n * n), want) <= 1e-13
    c.destroy()


def test_server_permute_batches_bit_exact():
    """Forty permutes of ranks 2..5 submitted back to back (consecutive
    permutes run as one batched launch), a source read by two of them, and a
    chain X -> Y -> Z whose second permute reads the first one's destination
    (it must wait for it): every destination equals the oracle's copy bit for
    bit."""
    rng = np.random.default_rng(17)
    c = sg.Cache(256 << 20)
    cases = []
    for q in range(40):
        rank = 2 + q % 4
        dims = [int(x) for x in rng.integers(2, 13, size=rank)]
        perm = [int(x) + 1 for x in rng.permutation(rank)]
        x = rng.standard_normal(int(np.prod(dims)))
        c.insert(1000 + q, x.size, host=x)
        c.insert(2000 + q, x.size, init="zero")
        cases.append((dims, [1] + perm, x, 2000 + q, 1000 + q))
    # fan-out: block 1000 also into 3000
    d0, t0, x0 = cases[0][0], cases[0][1], cases[0][2]
    c.insert(3000, x0.size, init="zero")
    cases.append((d0, t0, x0, 3000, 1000))
    # chain: X -> Y (reverse), Y -> Z (reverse back)
    dims = [6, 7, 8]
    X = rng.standard_normal(int(np.prod(dims)))
    c.insert(4000, X.size, host=X)
    c.insert(4001, X.size, init="zero")
    c.insert(4002, X.size, init="zero")
    srv = sg.Server(c, max_batch=64)
    tickets = [srv.submit_permute(d, t, src, dst) for d, t, _, dst, src in cases]
    tickets.append(srv.submit_permute(dims, [1, 3, 2, 1], 4000, 4001))
    tickets.append(srv.submit_permute(dims[::-1], [1, 3, 2, 1], 4001, 4002))
    for tk in tickets:
        srv.wait(tk)
    print("server stats", srv.stats(), flush=True)
    srv.destroy()
    for d, t, x, dst, _ in cases:
        want, e = o.copy(d, t, x)
        assert e == 0 and np.array_equal(c.read(dst, x.size), want), (d, t)
    Y, _ = o.copy(dims, [1, 3, 2, 1], X)
    assert np.array_equal(c.read(4001, X.size), Y)
    assert np.array_equal(c.read(4002, X.size), X)
    c.destroy()
