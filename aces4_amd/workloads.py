"""Synthetic coupled-cluster contraction workloads on segmented blocks.

Each workload is a SIAL-style contraction over segmented index ranges, e.g.
the ring term R(a,i,b,j) += sum_{k,c} W(k,b,c,j) * T(a,i,c,k): the SIP runs it
as a `pardo a,i,b,j` over result blocks with an inner `do k; do c` over
operand blocks, one `contract` super-instruction per block pair.  Here the
arrays are generated whole (seeded, portable counter-based RNG), cut into
dense column-major blocks exactly as the SIP's block shapes
(segment extents per index), and packed into one device arena per array.

Input data are bit-identical on every machine for the uniform draws
(SplitMix64 is pure integer arithmetic); normal draws use numpy's log/cos
and are reproducible within one machine.
"""
import itertools

import numpy as np

from . import sipgpu as sg

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, n, stream=0):
    """n 64-bit outputs of SplitMix64 for (seed, stream), counter-based."""
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + np.uint64(stream) * np.uint64(
            0xD1B54A32D192ED03)
        z = base + (np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n, stream=0, lo=0.0, hi=1.0):
    u = (splitmix64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return lo + (hi - lo) * u


def normal(seed, n, stream=0, sigma=1.0):
    m = (n + 1) // 2
    u1 = 1.0 - uniform(seed, m, 2 * stream + 1)  # (0, 1]
    u2 = uniform(seed, m, 2 * stream + 2)
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])[:n]
    return sigma * z


def seg_bounds(segs):
    b = [0]
    for s in segs:
        b.append(b[-1] + s)
    return b


class BlockedArray:
    """A dense array cut into blocks by per-index segment lists.

    `data` is the full array (Fortran order).  Blocks are packed contiguously
    (each one column-major) in `host`, in lexicographic block order with the
    first index fastest; `offset[blk]` is the element offset of a block.
    """

    def __init__(self, name, segs, data):
        self.name = name
        self.segs = [list(s) for s in segs]
        self.bounds = [seg_bounds(s) for s in self.segs]
        self.data = data
        self.offset = {}
        parts = []
        off = 0
        for blk in self.block_ids():
            sl = tuple(slice(self.bounds[d][blk[d]], self.bounds[d][blk[d] + 1])
                       for d in range(len(blk)))
            piece = np.asfortranarray(data[sl]).ravel(order="F")
            self.offset[blk] = off
            parts.append(piece)
            off += piece.size
        self.host = np.concatenate(parts) if parts else np.zeros(0)
        self.dev = None

    def block_ids(self):
        ranges = [range(len(s)) for s in self.segs]
        # first index fastest, as SIP block ids are enumerated
        for t in itertools.product(*reversed(ranges)):
            yield tuple(reversed(t))

    def dims(self, blk):
        return [self.segs[d][blk[d]] for d in range(len(blk))]

    def block(self, blk, host=None):
        h = self.host if host is None else host
        n = int(np.prod(self.dims(blk)))
        return h[self.offset[blk]:self.offset[blk] + n]

    def to_device(self):
        self.dev = sg.DeviceArray.from_host(self.host)
        return self.dev

    def addr(self, blk):
        return self.dev.at(self.offset[blk])


def contraction_entries(dest, left, right, dlab, llab, rlab, accumulate=True):
    """All block pairs of D(dlab) += sum L(llab) R(rlab), SIP pardo order.

    Returns a list of (dblk, lblk, rblk) tuples plus the contraction pattern
    (computed by the library's get_contraction_ptrn).
    """
    labels = list(dict.fromkeys(list(dlab) + list(llab) + list(rlab)))
    code = {lab: 1 + i for i, lab in enumerate(labels)}
    aces = [code[x] for x in list(dlab) + list(llab) + list(rlab)]
    ptrn, ierr = sg.get_contraction_ptrn(len(dlab), len(llab), len(rlab), aces)
    if ierr:
        raise ValueError("bad contraction labels (ierr %d)" % ierr)
    nseg = {}
    for arr, labs in ((dest, dlab), (left, llab), (right, rlab)):
        for d, lab in enumerate(labs):
            nseg[lab] = len(arr.segs[d])
    summed = [lab for lab in labels if lab not in dlab]
    out = []
    for dblk in dest.block_ids():
        idx = dict(zip(dlab, dblk))
        for combo in itertools.product(*[range(nseg[s]) for s in reversed(summed)]):
            idx.update(zip(reversed(summed), combo))
            lblk = tuple(idx[x] for x in llab)
            rblk = tuple(idx[x] for x in rlab)
            out.append((dblk, lblk, rblk))
    return out, [int(x) for x in ptrn]


def ring(o_segs=(15, 15), v_segs=(30, 30, 30, 30, 15, 15), seed=2):
    """R(a,i,b,j) += sum_{k,c} W(k,b,c,j) T(a,i,c,k); W ~ U(-.5,.5), T ~ N(0,1e-2^2)."""
    o, v = sum(o_segs), sum(v_segs)
    W = uniform(seed, o * v * v * o, stream=1, lo=-0.5, hi=0.5).reshape((o, v, v, o), order="F")
    T = normal(seed, v * o * v * o, stream=2, sigma=1e-2).reshape((v, o, v, o), order="F")
    R0 = np.zeros((v, o, v, o), order="F")
    Wb = BlockedArray("W", [o_segs, v_segs, v_segs, o_segs], W)
    Tb = BlockedArray("T", [v_segs, o_segs, v_segs, o_segs], T)
    Rb = BlockedArray("R", [v_segs, o_segs, v_segs, o_segs], R0)
    entries, ptrn = contraction_entries(Rb, Wb, Tb, "aibj", "kbcj", "aick")
    return dict(name="ring", D=Rb, L=Wb, R=Tb, entries=entries, ptrn=ptrn,
                labels=("aibj", "kbcj", "aick"))


def ladder(o_segs=(10,), v_segs=(10, 10, 10, 10), seed=1):
    """R(a,b,i,j) += sum_{c,d} V(a,b,c,d) T(c,d,i,j); V symmetrised, V(ab,cd) = V(cd,ab)."""
    o, v = sum(o_segs), sum(v_segs)
    X = normal(seed, v ** 4, stream=1, sigma=0.1).reshape((v, v, v, v), order="F")
    V = 0.5 * (X + X.transpose(2, 3, 0, 1))
    T = normal(seed, v * v * o * o, stream=2, sigma=1e-2).reshape((v, v, o, o), order="F")
    R0 = np.zeros((v, v, o, o), order="F")
    Vb = BlockedArray("V", [v_segs] * 4, np.asfortranarray(V))
    Tb = BlockedArray("T", [v_segs, v_segs, o_segs, o_segs], T)
    Rb = BlockedArray("R", [v_segs, v_segs, o_segs, o_segs], R0)
    entries, ptrn = contraction_entries(Rb, Vb, Tb, "abij", "abcd", "cdij")
    return dict(name="ladder", D=Rb, L=Vb, R=Tb, entries=entries, ptrn=ptrn,
                labels=("abij", "abcd", "cdij"))


class ContractionList(list):
    """Contractions on a workload's device arrays; `owner` (the workload)
    stays alive with the list, so the arrays are not freed while a batch
    planned from the list can still read them."""
    owner = None


def contractions_for(w, accumulate=True):
    """The workload's block contractions on its device arrays (uploaded on
    demand).  The list holds the workload: a batch planned from it (which
    keeps the list) keeps the arrays."""
    for k in ("D", "L", "R"):
        if w[k].dev is None:
            w[k].to_device()
    D, L, R, ptrn = w["D"], w["L"], w["R"], w["ptrn"]
    out = ContractionList(
        sg.make_contraction(ptrn, L.addr(lb), L.dims(lb), R.addr(rb), R.dims(rb), D.addr(db),
                            D.dims(db), accumulate=accumulate)
        for db, lb, rb in w["entries"])
    out.owner = w
    return out


def batch_for(w, accumulate=True, moduli=None):
    """A planned device batch for workload `w` (arrays uploaded on demand);
    `moduli` selects the split-precision path."""
    return sg.Batch(contractions_for(w, accumulate), moduli=moduli)


def flops(w):
    """Sum of 2*M*N*K over the workload's block pairs."""
    return sum(2.0 * m * n * k for m, n, k in mnk(w))


def mnk(w):
    """(M, N, K) of every block pair."""
    dl, ll, rl = w["labels"]
    out = []
    for db, lb, rb in w["entries"]:
        ld = dict(zip(ll, w["L"].dims(lb)))
        rd = dict(zip(rl, w["R"].dims(rb)))
        M = int(np.prod([ld[x] for x in ll if x in dl] or [1]))
        N = int(np.prod([rd[x] for x in rl if x in dl] or [1]))
        K = int(np.prod([rd[x] for x in rl if x not in dl] or [1]))
        out.append((M, N, K))
    return out


def random_segments(total, lo, hi, rng):
    """Segment extents drawn i.i.d. uniform in lo..hi until `total` is
    covered; the last one is truncated to fit."""
    segs = []
    while sum(segs) < total:
        segs.append(int(rng.integers(lo, hi + 1)))
    segs[-1] -= sum(segs) - total
    return segs


class DeviceBlocks:
    """A block-sparse array whose blocks live only in HBM.

    Kept blocks (those whose segments' irreps XOR to zero) are packed
    contiguously, column-major each, in block-id order; block `j` (ordinal
    among kept blocks) holds the seeded sequence (seed, stream = tag * 2^32 +
    j) drawn on the device, so any block can be regenerated or read back for
    checking without a host copy of the array.
    """

    def __init__(self, name, segs, irreps, seed, tag, dist, a=0.0, b=1.0, zero=False):
        self.name = name
        self.segs = [list(s) for s in segs]
        self.irreps = [list(r) for r in irreps]
        self.offset, self.ordinal = {}, {}
        off = 0
        for blk in self.all_block_ids():
            if self.kept(blk):
                self.ordinal[blk] = len(self.offset)
                self.offset[blk] = off
                off += int(np.prod(self.dims(blk)))
        self.size = off
        self.seed, self.tag, self.dist, self.a, self.b, self.zero = seed, tag, dist, a, b, zero
        self.dev = None

    def all_block_ids(self):
        ranges = [range(len(s)) for s in self.segs]
        for t in itertools.product(*reversed(ranges)):
            yield tuple(reversed(t))

    def block_ids(self):
        return iter(self.offset)

    def kept(self, blk):
        x = 0
        for d, s in enumerate(blk):
            x ^= self.irreps[d][s]
        return x == 0

    def dims(self, blk):
        return [self.segs[d][blk[d]] for d in range(len(blk))]

    def to_device(self):
        self.dev = sg.DeviceArray(max(self.size, 1))
        if self.zero:
            self.dev.zero()
        else:
            for blk, j in self.ordinal.items():
                sg.fill_random(self.addr(blk), int(np.prod(self.dims(blk))), self.seed,
                               (self.tag << 32) + j, self.dist, self.a, self.b)
        sg.synchronize()
        return self.dev

    def addr(self, blk):
        return self.dev.at(self.offset[blk])

    def download_block(self, blk, full=None):
        n = int(np.prod(self.dims(blk)))
        if full is not None:
            return full[self.offset[blk]:self.offset[blk] + n]
        return self.dev.download_range(self.offset[blk], n)


def sparse_ladder(o=40, v=200, seg_lo=12, seg_hi=40, probs=(0.75, 1 / 12, 1 / 12, 1 / 12),
                  seed=3):
    """Block-sparse ladder R(a,b,i,j) += sum_{c,d} V(a,b,c,d) T(c,d,i,j) with
    random uneven segments (extents uniform in seg_lo..seg_hi) and one of 4
    irreps per segment (probabilities `probs`); a block exists iff the XOR of
    its segments' irreps is 0.  V ~ N(0, 0.1^2), T ~ N(0, 1e-2^2) per block,
    generated in HBM; R starts at zero."""
    rng = np.random.default_rng(seed)
    vs = random_segments(v, seg_lo, seg_hi, rng)
    os_ = random_segments(o, seg_lo, seg_hi, rng)
    irr = np.random.default_rng([seed, 1])  # an independent stream for the irreps
    vi = [int(x) for x in irr.choice(4, size=len(vs), p=probs)]
    oi = [int(x) for x in irr.choice(4, size=len(os_), p=probs)]
    V = DeviceBlocks("V", [vs] * 4, [vi] * 4, seed, 1, "normal", 0.1)
    T = DeviceBlocks("T", [vs, vs, os_, os_], [vi, vi, oi, oi], seed, 2, "normal", 1e-2)
    R = DeviceBlocks("R", [vs, vs, os_, os_], [vi, vi, oi, oi], seed, 3, "normal", zero=True)
    labels = ("abij", "abcd", "cdij")
    code = {"a": 1, "b": 2, "i": 3, "j": 4, "c": 5, "d": 6}
    aces = [code[x] for x in "".join(labels)]
    ptrn, ierr = sg.get_contraction_ptrn(4, 4, 4, aces)
    assert ierr == 0
    entries = []
    nv = len(vs)
    for rb in R.block_ids():
        a, b, i, j = rb
        for d in range(nv):
            for c in range(nv):
                vb, tb = (a, b, c, d), (c, d, i, j)
                if vb in V.offset and tb in T.offset:
                    entries.append((rb, vb, tb))
    dense = 2.0 * float(v) ** 4 * float(o) ** 2
    return dict(name="sparse-ladder", D=R, L=V, R=T, entries=entries,
                ptrn=[int(x) for x in ptrn], labels=labels, v_segs=vs, o_segs=os_,
                v_irreps=vi, o_irreps=oi, dense_flops=dense)


SERVE_TERMS = {
    # name: (dest labels, left labels, right labels, left array, right array, dest array)
    "ladder": ("abij", "abcd", "cdij", "V", "T", "R"),
    "ring": ("aibj", "kbcj", "aick", "W", "U", "Q"),
    "hole": ("abij", "abkl", "klij", "T", "H", "R"),
}
SERVE_ARRAYS = {
    # name: (index kinds, initial value, a, b)
    "V": ("vvvv", "normal", 0.1, 0.0),
    "T": ("vvoo", "normal", 1e-2, 0.0),
    "W": ("ovvo", "uniform", -0.5, 0.5),
    "U": ("vovo", "normal", 1e-2, 0.0),
    "H": ("oooo", "normal", 0.1, 0.0),
    "R": ("vvoo", "zero", 0.0, 0.0),
    "Q": ("vovo", "zero", 0.0, 0.0),
}


def serving_mix(n_requests=3000, rate=3000.0, seed=4, v_nseg=8, o_nseg=4, seg_lo=8, seg_hi=64,
                mix=(("ladder", 0.4), ("ring", 0.4), ("hole", 0.2)), moduli=0):
    """A stream of single block contractions for the serving path.

    Segment extents are uniform in seg_lo..seg_hi (v_nseg virtual and o_nseg
    occupied segments); each request picks a term (ladder / ring / hole-hole)
    with the given probabilities and uniform segment indices, and accumulates
    one block pair into its result block.  Arrivals are Poisson with `rate`
    requests per second.  Blocks are named by 64-bit keys (array id << 48 |
    block index) and generated in HBM on first use.  `moduli` > 0 runs every
    request on the split-precision path.
    """
    rng = np.random.default_rng(seed)
    segs = {"v": [int(x) for x in rng.integers(seg_lo, seg_hi + 1, v_nseg)],
            "o": [int(x) for x in rng.integers(seg_lo, seg_hi + 1, o_nseg)]}
    names = sorted(SERVE_ARRAYS)
    arr_id = {n: i + 1 for i, n in enumerate(names)}

    def key(name, blk):
        kinds = SERVE_ARRAYS[name][0]
        lin = 0
        for d in reversed(range(len(blk))):
            lin = lin * len(segs[kinds[d]]) + blk[d]
        return (arr_id[name] << 48) | lin

    def dims(name, blk):
        kinds = SERVE_ARRAYS[name][0]
        return [segs[kinds[d]][blk[d]] for d in range(len(blk))]

    ptrns = {}
    for t, (dl, ll, rl, *_rest) in SERVE_TERMS.items():
        labels = list(dict.fromkeys(dl + ll + rl))
        code = {lab: 1 + i for i, lab in enumerate(labels)}
        p, e = sg.get_contraction_ptrn(4, 4, 4, [code[x] for x in dl + ll + rl])
        assert e == 0
        ptrns[t] = [int(x) for x in p]
    terms = [t for t, _ in mix]
    probs = np.array([p for _, p in mix], dtype=np.float64)
    probs /= probs.sum()
    kind_of = {"a": "v", "b": "v", "c": "v", "d": "v", "i": "o", "j": "o", "k": "o", "l": "o"}
    reqs, flops, blocks = [], [], {}
    for _ in range(n_requests):
        t = terms[int(rng.choice(len(terms), p=probs))]
        dl, ll, rl, ln, rn, dn = SERVE_TERMS[t]
        idx = {lab: int(rng.integers(0, len(segs[kind_of[lab]]))) for lab in set(dl + ll + rl)}
        lb, rb, db = (tuple(idx[x] for x in ll), tuple(idx[x] for x in rl),
                      tuple(idx[x] for x in dl))
        for name, blk in ((ln, lb), (rn, rb), (dn, db)):
            blocks[key(name, blk)] = (name, blk, int(np.prod(dims(name, blk))))
        reqs.append(sg.make_request(ptrns[t], key(ln, lb), dims(ln, lb), key(rn, rb),
                                    dims(rn, rb), key(dn, db), dims(dn, db), accumulate=True,
                                    moduli=moduli))
        ext = {lab: segs[kind_of[lab]][idx[lab]] for lab in idx}
        mn = float(np.prod([ext[x] for x in dl]))
        kk = float(np.prod([ext[x] for x in ll if x not in dl]))
        flops.append(2.0 * mn * kk)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    return dict(requests=reqs, arrivals=arrivals, flops=np.array(flops), blocks=blocks,
                segs=segs, seed=seed, rate=rate, terms=dict(mix))


def populate_cache(cache, mix):
    """Registers every block the stream touches (generated in HBM on first
    use; seeded per block)."""
    for k, (name, blk, n) in mix["blocks"].items():
        _, init, a, b = SERVE_ARRAYS[name]
        cache.insert(k, n, init=init, seed=mix["seed"], stream=k, a=a, b=b)
    return sum(n for _, _, n in mix["blocks"].values()) * 8
