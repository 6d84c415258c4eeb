"""(T)-style triples: W blocks from twelve permuted contractions, and their
energy through the orbital-energy denominator.

For a block tuple (A, B, C, I, J, K) of segments, the W block (indices
a, b, c, i, j, k, column-major in that order) is

    W = sum over the 6 permutations P of the pairs (a,i), (b,j), (c,k) of
        sum_d V(b,c,d,k) T(a,d,i,j)  +  sum_l V'(l,c,j,k) T(a,b,i,l)

with the labels permuted by P (V' holds -V(l,c,j,k)), and the tuple's
energy is  E = factor * sum W^2 / D,  D = e_i + e_j + e_k - e_a - e_b - e_c.
This is the structure of the reference's rccsdpt_aaa.sialx: X blocks
accumulated from contractions, energy_denominator_rhf applied to the block
(energy_denominator_rhf.F:399-444), then contracted with itself to a scalar
(rccsdpt_aaa.sialx:2164-2174).

On the device the twelve terms run as six batches: term 1 under a pair
permutation and term 2 under the reversed one split W's indices into the
same row and column sets, so each such pair is one GEMM task over the d and
l segments together.  The six batches accumulate into one W buffer in
stream order (each has its own row / column grouping of W's six indices, so
they cannot share a launch), then one fused pass divides by D, squares and
reduces, without writing the divided block.
"""
import itertools

import numpy as np

from . import sipgpu as sg
from .workloads import DeviceBlocks, uniform

PAIRS = (("a", "i"), ("b", "j"), ("c", "k"))
W_LABELS = "abcijk"
W_SIGNS = (-1, -1, -1, 1, 1, 1)
# box slots per kernel launch of sipgpu_triples_run (SIPGPU_TRIPLES_LAUNCH_BOXES)
LAUNCH_BOXES = 1 << 21


def term_labels():
    """The twelve (left labels, right labels, left array, right array,
    summed label) of W, in a fixed order."""
    out = []
    for perm in itertools.permutations(range(3)):
        (x1, y1), (x2, y2), (x3, y3) = (PAIRS[p] for p in perm)
        out.append((x2 + x3 + "d" + y3, x1 + "d" + y1 + y2, "V1", "T", "d"))
        out.append(("l" + x3 + y2 + y3, x1 + x2 + y1 + "l", "V2", "T", "l"))
    return out


def term_groups():
    """The twelve terms as six pairs that split W's indices into the same row
    and column sets: term 1 under a pair permutation P (V1 (x2 x3 d y3) times
    T (x1 d y1 y2)) and term 2 under P reversed, written with its operands
    swapped (T (x3 x2 y3 l) times V2 (l x1 y2 y1)) so that in both the left
    operand holds {x2, x3, y3}.  Each pair runs as one GEMM task over the d
    and l segments together (one pass over W instead of two)."""
    out = []
    for perm in itertools.permutations(range(3)):
        (x1, y1), (x2, y2), (x3, y3) = (PAIRS[p] for p in perm)
        out.append([(x2 + x3 + "d" + y3, x1 + "d" + y1 + y2, "V1", "T", "d"),
                    (x3 + x2 + y3 + "l", "l" + x1 + y2 + y1, "T", "V2", "l")])
    return out


def triples(o_segs=(20, 10), v_segs=(20,) * 12, seed=5):
    """Arrays of the triples term, generated in HBM per block: V1(b,c,d,k)
    ~ N(0, 0.1^2), V2(l,c,j,k) ~ N(0, 0.1^2) (held as -V), T(a,d,i,j) ~
    N(0, 1e-2^2); orbital energies e_occ ~ U(-2, -0.3), e_vir ~ U(0.1, 3)."""
    o, v = sum(o_segs), sum(v_segs)
    kinds = {"v": list(v_segs), "o": list(o_segs)}
    zero = {"v": [0] * len(v_segs), "o": [0] * len(o_segs)}

    def arr(name, kind, tag, sigma):
        return DeviceBlocks(name, [kinds[x] for x in kind], [zero[x] for x in kind], seed, tag,
                            "normal", sigma)

    w = dict(V1=arr("V1", "vvvo", 11, 0.1), V2=arr("V2", "ovoo", 12, 0.1),
             T=arr("T", "vvoo", 13, 1e-2), o_segs=list(o_segs), v_segs=list(v_segs), seed=seed)
    w["e_occ"] = uniform(seed, o, stream=21, lo=-2.0, hi=-0.3)
    w["e_vir"] = uniform(seed, v, stream=22, lo=0.1, hi=3.0)
    w["kind"] = {"a": "v", "b": "v", "c": "v", "d": "v", "i": "o", "j": "o", "k": "o",
                 "l": "o"}
    return w


def to_device(w):
    for k in ("V1", "V2", "T"):
        if w[k].dev is None:
            w[k].to_device()
    if "e_occ_dev" not in w:
        w["e_occ_dev"] = sg.DeviceArray.from_host(w["e_occ"])
        w["e_vir_dev"] = sg.DeviceArray.from_host(w["e_vir"])


def seg_offset(segs, s):
    return int(sum(segs[:s]))


def tuple_dims(w, tup):
    segs = {"v": w["v_segs"], "o": w["o_segs"]}
    return [segs[w["kind"][lab]][s] for lab, s in zip(W_LABELS, tup)]


def tuple_entries(w, tup, specs=None):
    """For W block `tup` = (A, B, C, I, J, K): per term (of `specs`, default
    the twelve of term_labels), the contraction pattern and the (left block,
    right block) pairs over the summed segments."""
    idx = dict(zip(W_LABELS, tup))
    nseg = {"d": len(w["v_segs"]), "l": len(w["o_segs"])}
    terms = []
    for ll, rl, ln, rn, summed in (term_labels() if specs is None else specs):
        labels = list(dict.fromkeys(W_LABELS + ll + rl))
        code = {lab: 1 + i for i, lab in enumerate(labels)}
        ptrn, e = sg.get_contraction_ptrn(6, 4, 4, [code[x] for x in W_LABELS + ll + rl])
        assert e == 0
        pairs = []
        for s in range(nseg[summed]):
            idx[summed] = s
            pairs.append((ln, tuple(idx[x] for x in ll), rn, tuple(idx[x] for x in rl)))
        terms.append(([int(x) for x in ptrn], pairs))
    return terms


class TupleEnergy:
    """Planned device work for one W block: six batches (the term pairs of
    term_groups, each one GEMM task) into `wbuf` and the fused denominator /
    energy reduction into `out` (device double)."""

    def __init__(self, w, tup, wbuf, out, moduli=None, factor=1.0, parts=None, tune=False,
                 two_streams=False):
        """`parts` (optional): five device buffers of at least W's size.  The
        first batch then assigns W, batch g > 0 assigns its own products into
        parts[g - 1], and one pass adds the parts to W in batch order and
        reduces the energy: the same roundings as accumulating every batch
        into W, but the reads of the partial sums stream through HBM in one
        kernel instead of waiting in five GEMM epilogues.  `tune`: plan each
        batch with the autotuner (its timing runs write scratch blocks only;
        batches of equal shape share one tuning).  `two_streams` (with
        `parts`): the batches alternate between the caller's stream and a
        second one, joined before the summing pass, so one batch's pack and
        last partial round of tiles overlap the next batch's GEMM; the
        batches write distinct buffers and own their panels, so the result
        is the same."""
        to_device(w)
        self.dims = tuple_dims(w, tup)
        self.batches = []
        groups = term_groups()
        if parts is not None and len(parts) < len(groups) - 1:
            raise ValueError("need %d partial buffers" % (len(groups) - 1))
        if parts is not None and len(set(parts[:len(groups) - 1]) | {wbuf}) != len(groups):
            raise ValueError("the partial buffers must be distinct from each other and from W")
        self.parts = list(parts[:len(groups) - 1]) if parts is not None else None
        first = True
        for g, group in enumerate(groups):
            dest = wbuf if (self.parts is None or g == 0) else self.parts[g - 1]
            if self.parts is not None:
                first = True  # every batch assigns its own buffer
            cs = []
            for ptrn, pairs in tuple_entries(w, tup, group):
                for ln, lb, rn, rb in pairs:
                    L, R = w[ln], w[rn]
                    cs.append(sg.make_contraction(ptrn, L.addr(lb), L.dims(lb), R.addr(rb),
                                                  R.dims(rb), dest, self.dims,
                                                  accumulate=not first))
                    first = False
            # entries of two patterns into one block plan on panels (the table
            # kernel takes one pattern per block)
            self.batches.append(sg.tune(cs, moduli=moduli)[0] if tune else sg.Batch(cs, moduli=moduli))
        segs = {"v": w["v_segs"], "o": w["o_segs"]}
        self.eps = []
        for lab, s in zip(W_LABELS, tup):
            kind = w["kind"][lab]
            base = w["e_vir_dev"] if kind == "v" else w["e_occ_dev"]
            self.eps.append(base.at(seg_offset(segs[kind], s)))
        self.wbuf, self.out, self.factor = wbuf, out, factor
        self.flops = sum(b.flops for b in self.batches)
        self.aux = None
        if two_streams and self.parts is not None:
            self.aux = sg.Stream()
            self.fork, self.join = sg.Event(), sg.Event()

    def execute(self, s=None):
        if self.aux is not None:
            # fork: the second stream starts after the caller's earlier work
            self.fork.record(s)
            self.aux.wait(self.fork)
            for g, b in enumerate(self.batches):
                b.execute(s if g % 2 == 0 else self.aux)
            self.join.record(self.aux)
            sg.stream_wait(self.join, s)
        else:
            for b in self.batches:
                b.execute(s)
        if self.parts is not None:
            # the parts' sum and the energy in one pass
            sg.energy_sum_parts(self.dims, self.wbuf, self.parts, self.eps, W_SIGNS, self.factor,
                                self.out, s=s)
            return
        sg.energy_sum(self.dims, self.wbuf, self.wbuf, self.eps, W_SIGNS, self.factor, self.out,
                      s=s)

    def destroy(self):
        for b in self.batches:
            b.destroy()
        self.batches = []
        if self.aux is not None:
            self.aux.synchronize()
            self.aux.destroy()
            self.aux = None


class FusedTriples:
    """The whole unrestricted (T) sum with W kept on chip
    (sipgpu_triples_*): one workgroup per 5^6 box of W runs the six GEMMs of
    the box on FP64 MFMA, adds them in LDS, and reduces the box's energy;
    W is never written.  The operands are packed once into two panel arrays.

        ft = FusedTriples(w)        # plan + pack (~4.3 GB at o = 30, v = 240)
        ft.run()                    # all box slots, or ft.run(box0, nbox)
        ft.energy(out_ptr)          # out += factor * sum of the partials
    """

    def __init__(self, w, stream=None):
        import ctypes
        to_device(w)
        self.w = w
        nv, no = len(w["v_segs"]), len(w["o_segs"])

        def table(name, counts):
            arr = w[name]
            ptrs = [arr.addr(t) for t in itertools.product(*[range(c) for c in counts])]
            return (ctypes.c_void_p * len(ptrs))(*ptrs)

        self._T = table("T", (nv, nv, no, no))
        self._V1 = table("V1", (nv, nv, nv, no))
        self._V2 = table("V2", (no, nv, no, no))
        vs, os_ = sg._ints(w["v_segs"]), sg._ints(w["o_segs"])
        p = ctypes.c_void_p()
        sg._check(sg.lib().sipgpu_triples_create(
            nv, sg._iptr(vs), no, sg._iptr(os_), ctypes.cast(self._T, ctypes.c_void_p),
            ctypes.cast(self._V1, ctypes.c_void_p), ctypes.cast(self._V2, ctypes.c_void_p),
            w["e_occ_dev"].ptr, w["e_vir_dev"].ptr, ctypes.byref(p)))
        self.handle = p.value
        slots, fpb, pbytes = ctypes.c_longlong(0), ctypes.c_double(0), ctypes.c_size_t(0)
        sg._check(sg.lib().sipgpu_triples_info(self.handle, ctypes.byref(slots), ctypes.byref(fpb),
                                               ctypes.byref(pbytes)))
        self.slots, self.mfma_flop_per_box, self.panel_bytes = slots.value, fpb.value, pbytes.value
        v, o = sum(w["v_segs"]), sum(w["o_segs"])
        # the algorithmic work of the whole sum: 12 terms of 2 x v^3 o^3 (v or o)
        self.flops = 12.0 * (v ** 3) * (o ** 3) * (v + o)
        self.pack(stream)

    def pack(self, stream=None):
        sg._check(sg.lib().sipgpu_triples_pack(self.handle, stream.handle if stream else None))

    def run(self, box0=0, nbox=None, stream=None):
        if nbox is None:
            nbox = self.slots - box0
        sg._check(sg.lib().sipgpu_triples_run(self.handle, int(box0), int(nbox),
                                              stream.handle if stream else None))

    def energy(self, out, factor=1.0, stream=None):
        sg._check(sg.lib().sipgpu_triples_energy(self.handle, float(factor), out,
                                                 stream.handle if stream else None))

    def tuple(self, tup, out, w_out=None, factor=1.0, stream=None):
        t = sg._ints(tup)
        sg._check(sg.lib().sipgpu_triples_tuple(self.handle, sg._iptr(t), float(factor), w_out,
                                                out, stream.handle if stream else None))

    def destroy(self):
        if self.handle:
            sg._check(sg.lib().sipgpu_triples_destroy(self.handle))
            self.handle = None
