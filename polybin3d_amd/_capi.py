"""ctypes binding of libpb3d.so (include/pb3d.h).

The library is loaded from this package directory (built in-tree by ``python -m
polybin3d_amd.build``).  There is no fallback: if the library or a GPU is missing, calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PB3D_LIB overrides the library path (an alternative in-tree build of the same sources)
LIB_PATH = os.environ.get("PB3D_LIB", os.path.join(_HERE, "libpb3d.so"))

c_int32, c_int64, c_float, c_double = ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_double
c_void_p, c_size_t = ctypes.c_void_p, ctypes.c_size_t

PB3D_NO_BIN = 255

# symbol -> (restype, argtypes); the full list of exported entry points in include/pb3d.h
EXPORTS = {
    "pb3d_version": (ctypes.c_int, []),
    "pb3d_last_error": (ctypes.c_char_p, []),
    "pb3d_fft_counter": (c_int64, [ctypes.c_int]),
    "pb3d_prof_enable": (ctypes.c_int, [ctypes.c_int]),
    "pb3d_prof_read": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(c_int64), ctypes.POINTER(c_double),
                                      ctypes.POINTER(c_double)]),
    "pb3d_plan_create": (ctypes.c_int, [ctypes.POINTER(c_void_p), c_void_p]),
    "pb3d_plan_destroy": (ctypes.c_int, [c_void_p]),
    "pb3d_fft_r2c": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, c_void_p, c_void_p, c_void_p]),
    "pb3d_fft_c2r": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, c_void_p, c_void_p, c_void_p]),
    "pb3d_fft_pointing": (ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.c_int, c_void_p, c_void_p, c_float, c_void_p,
                                         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pb3d_fft_pointing_work_size": (c_int64, [c_void_p, ctypes.c_int, ctypes.c_int, c_void_p, c_void_p]),
    "pb3d_icsum_partial_size": (c_int64, [c_void_p, ctypes.c_int, ctypes.c_int]),
    "pb3d_pointing_rows_mode": (ctypes.c_int, [ctypes.c_int]),
    "pb3d_kmap": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_int, c_void_p, c_void_p]),
    "pb3d_kmap_packed": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, ctypes.c_int, c_void_p,
                                        c_void_p]),
    "pb3d_gather_rows": (ctypes.c_int, [c_void_p, c_int64, ctypes.c_int, c_void_p, ctypes.c_int, c_int64, c_int64,
                                        c_void_p, c_void_p, c_void_p]),
    "pb3d_grf_device": (ctypes.c_int, [c_void_p, ctypes.c_uint64, c_void_p, c_float, c_void_p, c_void_p]),
    "pb3d_grf_device_at": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_uint64, c_void_p, c_float, c_void_p,
                                          c_void_p]),
    "pb3d_grf_hermitian": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_void_p]),
    "pb3d_bin_reduce_work_size": (c_size_t, [c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "pb3d_bin_reduce": (ctypes.c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_int64, ctypes.c_int,
                                       ctypes.c_int, c_void_p, c_void_p, ctypes.c_int, c_void_p, ctypes.c_int,
                                       c_void_p, c_void_p, c_void_p]),
    "pb3d_bin_reduce_modes": (ctypes.c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_int64, ctypes.c_int,
                                             ctypes.c_int, c_void_p, c_void_p, ctypes.c_int, c_void_p, ctypes.c_int,
                                             c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "pb3d_bin_reduce_modes_noalloc": (ctypes.c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_int64,
                                                     ctypes.c_int, ctypes.c_int, c_void_p, c_void_p, ctypes.c_int,
                                                     c_void_p, ctypes.c_int, c_void_p, c_int64, c_void_p, c_void_p]),
    "pb3d_pair_gemm_work_size": (c_size_t, [c_int64, ctypes.c_int, ctypes.c_int]),
    "pb3d_pair_gemm": (ctypes.c_int, [c_void_p, ctypes.c_int, c_int64, c_void_p, ctypes.c_int, c_void_p, c_void_p,
                                      ctypes.c_int, ctypes.c_int, c_void_p, c_void_p, c_void_p]),
    "pb3d_pair_gemm_cfg": (ctypes.c_int, [c_void_p, ctypes.c_int, c_int64, c_void_p, ctypes.c_int, c_void_p, c_void_p,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_void_p, c_void_p,
                                          c_void_p]),
    "pb3d_pair_adjoint_work_size": (c_size_t, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "pb3d_pair_adjoint": (ctypes.c_int, [c_void_p, ctypes.c_int, c_int64, c_void_p, c_void_p, ctypes.c_int, c_void_p,
                                         ctypes.c_int, ctypes.c_int, c_void_p, c_void_p, c_void_p]),
    "pb3d_fisher_tiles": (ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.c_int, c_void_p, ctypes.c_int]),
    "pb3d_fisher_contract": (ctypes.c_int, [c_void_p, c_void_p, c_int64, ctypes.c_int, c_void_p, ctypes.c_int,
                                            c_void_p, ctypes.c_int, c_void_p, c_void_p, c_void_p, c_double,
                                            ctypes.c_int, c_void_p, c_int64, c_void_p]),
    "pb3d_triplet_work_size": (c_size_t, [c_int64, ctypes.c_int, ctypes.c_int]),
    "pb3d_triplet_sum": (ctypes.c_int, [c_void_p, ctypes.c_int, c_int64, c_void_p, c_void_p, c_void_p, ctypes.c_int,
                                        ctypes.c_int, c_void_p, c_void_p, c_void_p]),
}


class Grid(ctypes.Structure):
    """pb3d_grid"""
    _fields_ = [("n", c_int32 * 3), ("local", c_int32), ("boxsize", c_double * 3),
                ("boxcenter", c_double * 3), ("sight", c_double * 3)]


class ROp(ctypes.Structure):
    """pb3d_rop: real-space voxel op (R2C prologue / C2R epilogue)."""
    _fields_ = [("kind", c_int32), ("ell", c_int32), ("m", c_int32), ("accumulate", c_int32),
                ("a", c_void_p), ("b", c_void_p), ("c", c_void_p), ("d", c_void_p),
                ("batch_stride_a", c_int64), ("batch_stride_b", c_int64), ("batch_stride_c", c_int64),
                ("batch_stride_d", c_int64), ("scale", c_float), ("pad_", c_int32)]


class KOp(ctypes.Structure):
    """pb3d_kop: Fourier-space mode op (C2R prologue / R2C epilogue)."""
    _fields_ = [("kind", c_int32), ("ell", c_int32), ("ell2", c_int32), ("m", c_int32),
                ("accumulate", c_int32), ("pad_", c_int32), ("bin_table", c_void_p), ("bins", c_void_p),
                ("w", c_void_p), ("x_batch_stride", c_int64), ("scale", c_float), ("pad2_", c_int32),
                ("pack_index", c_void_p), ("packed_len", c_int64), ("trip", c_void_p), ("pair_index", c_void_p),
                ("ft", c_void_p), ("ft_lstride", c_int64), ("nk", c_int32), ("terms", c_int32),
                ("trip0", c_int32), ("leg3_legendre", c_int32), ("support", c_int32 * 3), ("x_batch_group", c_int32)]


class SubMap(ctypes.Structure):
    """pb3d_submap: one weighted sub-map of pb3d_fft_pointing."""
    _fields_ = [("kop", KOp), ("inp", c_void_p), ("ell", c_int32), ("m", c_int32), ("coef", c_float),
                ("pad_", c_int32)]


class FisherSeg(ctypes.Structure):
    """pb3d_fisher_seg: one bin segment of the Fisher contraction."""
    _fields_ = [("a_off", c_int64), ("q_off", c_int64), ("k", c_int32), ("rows", c_int32), ("row0", c_int32),
                ("pad_", c_int32)]


class IcSum(ctypes.Structure):
    """pb3d_icsum: per-bin weighted sums of the pointing chain's combined real map."""
    _fields_ = [("weight", c_void_p), ("bin", c_void_p), ("nbins", c_int32), ("pad_", c_int32), ("sums", c_void_p),
                ("partial", c_void_p)]


ROP_PLAIN, ROP_MUL, ROP_DIFF, ROP_YLM, ROP_MULYLM, ROP_SQUARE = range(6)
KOP_PLAIN, KOP_FILTER, KOP_YLM, KOP_WEIGHT, KOP_BKQ, KOP_PACK, KOP_BINPOLY = range(7)

_lib = None


def lib():
    """Load libpb3d.so once; raise with a clear message when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libpb3d.so not found at %s; build it with `python -m polybin3d_amd.build`"
                               % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class PB3DError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        msg = lib().pb3d_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError("pb3d: " + msg)
        raise PB3DError("pb3d error %d: %s" % (rc, msg))


def rop(kind=ROP_PLAIN, a=None, b=None, c=None, d=None, ell=0, m=0, accumulate=0, scale=1.0,
        sa=0, sb=0, sc=0, sd=0):
    o = ROp()
    o.kind, o.ell, o.m, o.accumulate = kind, ell, m, int(accumulate)
    o.a, o.b, o.c, o.d = (_ptr(a), _ptr(b), _ptr(c), _ptr(d))
    o.batch_stride_a, o.batch_stride_b, o.batch_stride_c, o.batch_stride_d = sa, sb, sc, sd
    o.scale = float(scale)
    o._t = {k: v for k, v in (('a', a), ('b', b), ('c', c), ('d', d)) if v is not None and not isinstance(v, int)}
    return o


def kop(kind=KOP_PLAIN, bin_table=None, bins=None, w=None, ell=0, ell2=0, m=0, accumulate=0, scale=1.0,
        x_batch_stride=0, pack_index=None, packed_len=0, trip=None, pair_index=None, ft=None, ft_lstride=0,
        nk=0, terms=0, trip0=0, leg3_legendre=0, support=None, x_batch_group=0):
    o = KOp()
    o.kind, o.ell, o.ell2, o.m, o.accumulate = kind, ell, ell2, m, int(accumulate)
    o.bin_table, o.bins, o.w = _ptr(bin_table), _ptr(bins), _ptr(w)
    o.x_batch_stride = int(x_batch_stride)
    o.x_batch_group = int(x_batch_group)
    o.scale = float(scale)
    o.pack_index, o.packed_len = _ptr(pack_index), int(packed_len)
    o.trip, o.pair_index, o.ft, o.ft_lstride = _ptr(trip), _ptr(pair_index), _ptr(ft), int(ft_lstride)
    o.nk, o.terms, o.trip0, o.leg3_legendre = int(nk), int(terms), int(trip0), int(leg3_legendre)
    sup = (-1, -1, -1) if support is None else support
    for i in range(3):
        o.support[i] = int(sup[i])
    o._t = {k: v for k, v in (('bin_table', bin_table), ('bins', bins), ('w', w), ('pack_index', pack_index),
                              ('trip', trip), ('pair_index', pair_index), ('ft', ft))
            if v is not None and not isinstance(v, int)}
    return o


def _ptr(t):
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def byref_or_none(s):
    return None if s is None else ctypes.byref(s)


def make_grid(gridsize, boxsize, boxcenter, sight, local):
    g = Grid()
    for i in range(3):
        g.n[i] = int(gridsize[i])
        g.boxsize[i] = float(boxsize[i])
        g.boxcenter[i] = float(boxcenter[i])
        g.sight[i] = float(sight[i])
    g.local = int(bool(local))
    return g


def int_array(vals):
    arr = (c_int32 * len(vals))(*[int(v) for v in vals])
    return arr


__all__ = ["lib", "check", "rop", "kop", "Grid", "make_grid", "EXPORTS", "PB3D_NO_BIN", "np"]
