"""ctypes bindings to the CPU checkers used by the tests.

`Oracle` wraps oracle/liboracle.so, the C restatement of the reference
kernels.  `RefLib` wraps oracle/_ref/libtensordil_ref.so, the reference's own
Fortran kernel module compiled from its sources (present only where the
reference tree was available at build time).  Both are checkers, never the
code under test.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libtensordil_ref.so")

_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)


def _ints(seq):
    a = np.ascontiguousarray(np.asarray(list(seq) if not isinstance(seq, np.ndarray) else seq,
                                        dtype=np.int32))
    if a.size == 0:
        a = np.zeros(1, dtype=np.int32)
    return a


def _iptr(a):
    return a.ctypes.data_as(_ip)


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """The C restatement (oracle/oracle.c)."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise RuntimeError("oracle library missing: %s (run `make -C oracle`)" % path)
        self.lib = ctypes.CDLL(path)
        self.lib.orc_tensor_size_by_shape.restype = ctypes.c_longlong

    def size_by_shape(self, dims):
        d = _ints(dims)
        ierr = ctypes.c_int(0)
        s = self.lib.orc_tensor_size_by_shape(len(dims), _iptr(d), ctypes.byref(ierr))
        return s, ierr.value

    def get_contraction_ptrn(self, drank, lrank, rrank, aces):
        a = _ints(aces)
        out = np.zeros(max(lrank + rrank, 1), dtype=np.int32)
        ierr = ctypes.c_int(0)
        self.lib.orc_get_contraction_ptrn(drank, lrank, rrank, _iptr(a), _iptr(out),
                                          ctypes.byref(ierr))
        return out[:lrank + rrank].copy(), ierr.value

    def copy(self, dims, dim_transp, x):
        d = _ints(dims)
        t = _ints(dim_transp)
        x = _f64(x).ravel()
        out = np.empty_like(x)
        ierr = ctypes.c_int(0)
        self.lib.orc_tensor_block_copy(len(dims), _iptr(d), _iptr(t), _dptr(x), _dptr(out),
                                       ctypes.byref(ierr))
        return out, ierr.value

    def contract(self, ptrn, l, ldims, r, rdims, ddims, d_init=None):
        p = _ints(ptrn)
        ld, rd, dd = _ints(ldims), _ints(rdims), _ints(ddims)
        l = _f64(l).ravel()
        r = _f64(r).ravel()
        dsize = int(np.prod(ddims)) if len(ddims) else 1
        d = np.full(dsize, np.nan) if d_init is None else _f64(d_init).ravel().copy()
        ierr = ctypes.c_int(0)
        self.lib.orc_tensor_block_contract(_iptr(p), _dptr(l), len(ldims), _iptr(ld),
                                           _dptr(r), len(rdims), _iptr(rd), _dptr(d),
                                           len(ddims), _iptr(dd), ctypes.byref(ierr))
        return d, ierr.value

    def energy_denominator(self, dims, x, eps, signs):
        """x / eps (energy_denominator_rhf), eps per index from `eps[d]`."""
        x = _f64(x).ravel().copy()
        ev = [np.ascontiguousarray(e, dtype=np.float64) for e in eps]
        ptrs = (ctypes.c_void_p * len(dims))(*[e.ctypes.data for e in ev])
        d, sg_ = _ints(dims), _ints(signs)
        self.lib.orc_energy_denominator(len(dims), _iptr(d), _dptr(x), ptrs, _iptr(sg_))
        return x

    def fcontract(self, l, r, d=0.0):
        l, r = _f64(l).ravel(), _f64(r).ravel()
        self.lib.orc_fcontract.restype = ctypes.c_double
        self.lib.orc_fcontract.argtypes = [ctypes.c_longlong, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_double]
        return self.lib.orc_fcontract(l.size, l.ctypes.data, r.ctypes.data, float(d))

    def block_init(self, dims, t, sval):
        d = _ints(dims)
        t = _f64(t).ravel().copy()
        ierr = ctypes.c_int(0)
        self.lib.orc_block_init(len(dims), _iptr(d), _dptr(t), ctypes.c_double(sval),
                                ctypes.byref(ierr))
        return t, ierr.value

    def block_scale(self, dims, t, s):
        d = _ints(dims)
        t = _f64(t).ravel().copy()
        ierr = ctypes.c_int(0)
        self.lib.orc_block_scale(len(dims), _iptr(d), _dptr(t), ctypes.c_double(s),
                                 ctypes.byref(ierr))
        return t, ierr.value

    def block_norm2(self, dims, t):
        d = _ints(dims)
        t = _f64(t).ravel()
        ierr = ctypes.c_int(0)
        self.lib.orc_block_norm2.restype = ctypes.c_double
        v = self.lib.orc_block_norm2(len(dims), _iptr(d), _dptr(t), ctypes.byref(ierr))
        return v, ierr.value

    def block_add(self, dims, t0, t1, s):
        d = _ints(dims)
        t0 = _f64(t0).ravel().copy()
        ierr = ctypes.c_int(0)
        self.lib.orc_tensor_block_add(len(dims), _iptr(d), _dptr(t0), _dptr(_f64(t1).ravel()),
                                      ctypes.c_double(s), ctypes.byref(ierr))
        return t0, ierr.value

    def block_slice(self, tens, tens_ext, slice_ext, ext_beg, out=None):
        te, se, eb = _ints(tens_ext), _ints(slice_ext), _ints(ext_beg)
        n = max(int(np.prod(slice_ext)) if len(slice_ext) else 1, 1)
        out = np.full(n, np.nan) if out is None else _f64(out).ravel().copy()
        ierr = ctypes.c_int(0)
        self.lib.orc_block_slice(len(tens_ext), _dptr(_f64(tens).ravel()), _iptr(te),
                                 _dptr(out), _iptr(se), _iptr(eb), ctypes.byref(ierr))
        return out, ierr.value

    def block_insert(self, tens, tens_ext, slc, slice_ext, ext_beg):
        te, se, eb = _ints(tens_ext), _ints(slice_ext), _ints(ext_beg)
        tens = _f64(tens).ravel().copy()
        ierr = ctypes.c_int(0)
        self.lib.orc_block_insert(len(tens_ext), _dptr(tens), _iptr(te),
                                  _dptr(_f64(slc).ravel()), _iptr(se), _iptr(eb),
                                  ctypes.byref(ierr))
        return tens, ierr.value

    def perms(self, ptrn, ldims, rdims, ddims):
        """determine_index_permutations: (lo2n, ro2n, do2n, flags, M, N, K)."""

        class P(ctypes.Structure):
            _fields_ = [("lo2n", ctypes.c_int * 33), ("ro2n", ctypes.c_int * 33),
                        ("do2n", ctypes.c_int * 33), ("ltransp", ctypes.c_int),
                        ("rtransp", ctypes.c_int), ("dtransp", ctypes.c_int),
                        ("nlu", ctypes.c_int), ("nru", ctypes.c_int), ("ncd", ctypes.c_int),
                        ("lld", ctypes.c_longlong), ("lrd", ctypes.c_longlong),
                        ("lcd", ctypes.c_longlong), ("lsize", ctypes.c_longlong),
                        ("rsize", ctypes.c_longlong), ("dsize", ctypes.c_longlong)]

        p = P()
        pt = _ints(ptrn)
        ld, rd, dd = _ints(ldims), _ints(rdims), _ints(ddims)
        self.lib.orc_determine_index_permutations(_iptr(pt), len(ldims), _iptr(ld), len(rdims),
                                                  _iptr(rd), len(ddims), _iptr(dd),
                                                  ctypes.byref(p))
        lr, rr, dr = len(ldims), len(rdims), len(ddims)
        return dict(lo2n=list(p.lo2n[1:lr + 1]), ro2n=list(p.ro2n[1:rr + 1]),
                    do2n=list(p.do2n[1:dr + 1]), ltransp=p.ltransp, rtransp=p.rtransp,
                    dtransp=p.dtransp, M=p.lld, N=p.lrd, K=p.lcd)


class RefLib:
    """The reference's own Fortran kernels, built by oracle/Makefile."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise RuntimeError("reference build missing: %s" % path)
        self.lib = ctypes.CDLL(path)
        self.lib.tensor_size_by_shape_.restype = ctypes.c_longlong

    @staticmethod
    def available(path=REF_SO):
        return os.path.exists(path)

    def get_contraction_ptrn(self, drank, lrank, rrank, aces):
        a = _ints(aces)
        out = np.zeros(max(lrank + rrank, 1), dtype=np.int32)
        ierr = ctypes.c_int(0)
        self.lib.get_contraction_ptrn_(ctypes.byref(ctypes.c_int(drank)),
                                       ctypes.byref(ctypes.c_int(lrank)),
                                       ctypes.byref(ctypes.c_int(rrank)), _iptr(a), _iptr(out),
                                       ctypes.byref(ierr))
        return out[:lrank + rrank].copy(), ierr.value

    def copy(self, dims, dim_transp, x, nthreads=8):
        d = _ints(dims)
        t = _ints(dim_transp)
        x = _f64(x).ravel()
        out = np.empty_like(x)
        ierr = ctypes.c_int(0)
        self.lib.tensor_block_copy__(ctypes.byref(ctypes.c_int(nthreads)),
                                     ctypes.byref(ctypes.c_int(len(dims))), _iptr(d), _iptr(t),
                                     _dptr(x), _dptr(out), ctypes.byref(ierr))
        return out, ierr.value

    def contract(self, ptrn, l, ldims, r, rdims, ddims, d_init=None, nthreads=8):
        p = _ints(ptrn)
        ld, rd, dd = _ints(ldims), _ints(rdims), _ints(ddims)
        l = _f64(l).ravel()
        r = _f64(r).ravel()
        dsize = int(np.prod(ddims)) if len(ddims) else 1
        d = np.full(dsize, np.nan) if d_init is None else _f64(d_init).ravel().copy()
        ierr = ctypes.c_int(0)
        self.lib.tensor_block_contract__(ctypes.byref(ctypes.c_int(nthreads)), _iptr(p),
                                         _dptr(l), ctypes.byref(ctypes.c_int(len(ldims))),
                                         _iptr(ld), _dptr(r),
                                         ctypes.byref(ctypes.c_int(len(rdims))), _iptr(rd),
                                         _dptr(d), ctypes.byref(ctypes.c_int(len(ddims))),
                                         _iptr(dd), ctypes.byref(ierr))
        return d, ierr.value

    def add(self, dims, t0, t1, scale=1.0, nthreads=8):
        d = _ints(dims)
        ierr = ctypes.c_int(0)
        self.lib.tensor_block_add__(ctypes.byref(ctypes.c_int(nthreads)),
                                    ctypes.byref(ctypes.c_int(len(dims))), _iptr(d), _dptr(t0),
                                    _dptr(_f64(t1)), ctypes.byref(ctypes.c_double(scale)),
                                    ctypes.byref(ierr))
        return ierr.value

    def _c(self, v, ctype=ctypes.c_int):
        return ctypes.byref(ctype(v))

    def block_init(self, dims, t, sval, nthreads=1):
        d = _ints(dims)
        t = _f64(t).ravel().copy()
        ierr = ctypes.c_int(0)
        self.lib.tensor_block_init__(self._c(nthreads), _dptr(t), self._c(len(dims)), _iptr(d),
                                     self._c(sval, ctypes.c_double), ctypes.byref(ierr))
        return t, ierr.value

    def block_scale(self, dims, t, s, nthreads=1):
        d = _ints(dims)
        t = _f64(t).ravel().copy()
        ierr = ctypes.c_int(0)
        self.lib.tensor_block_scale__(self._c(nthreads), _dptr(t), self._c(len(dims)), _iptr(d),
                                      self._c(s, ctypes.c_double), ctypes.byref(ierr))
        return t, ierr.value

    def block_norm2(self, dims, t, nthreads=1):
        d = _ints(dims)
        t = _f64(t).ravel()
        ierr = ctypes.c_int(0)
        self.lib.tensor_block_norm2__.restype = ctypes.c_double
        v = self.lib.tensor_block_norm2__(self._c(nthreads), _dptr(t), self._c(len(dims)),
                                          _iptr(d), ctypes.byref(ierr))
        return v, ierr.value

    def block_slice(self, tens, tens_ext, slice_ext, ext_beg, nthreads=1):
        te, se, eb = _ints(tens_ext), _ints(slice_ext), _ints(ext_beg)
        n = max(int(np.prod(slice_ext)) if len(slice_ext) else 1, 1)
        out = np.full(n, np.nan)
        ierr = ctypes.c_int(0)
        self.lib.tensor_block_slice__(self._c(nthreads), self._c(len(tens_ext)),
                                      _dptr(_f64(tens).ravel()), _iptr(te), _dptr(out),
                                      _iptr(se), _iptr(eb), ctypes.byref(ierr))
        return out, ierr.value

    def block_insert(self, tens, tens_ext, slc, slice_ext, ext_beg, nthreads=1):
        te, se, eb = _ints(tens_ext), _ints(slice_ext), _ints(ext_beg)
        tens = _f64(tens).ravel().copy()
        ierr = ctypes.c_int(0)
        self.lib.tensor_block_insert__(self._c(nthreads), self._c(len(tens_ext)), _dptr(tens),
                                       _iptr(te), _dptr(_f64(slc).ravel()), _iptr(se),
                                       _iptr(eb), ctypes.byref(ierr))
        return tens, ierr.value
