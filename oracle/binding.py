"""ctypes bindings for the oracle restatement and the compiled reference.

TEST INFRASTRUCTURE ONLY (see package docstring).

Both libraries expose the same set of functions under different prefixes
(``oracle_*`` for the restatement in icbc_oracle.c, ``ref_*`` for the
reference header compiled by ref_entry.cpp), so one wrapper class serves both.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "libicbc_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_VARIANTS = {"scalar": "libicbc_ref_scalar.so", "avx2": "libicbc_ref_avx2.so",
                "avx512": "libicbc_ref_avx512.so"}

_f = C.POINTER(C.c_float)
_u8 = C.POINTER(C.c_uint8)
_i = C.POINTER(C.c_int)


def build(with_reference=None):
    """Compile the oracle (and the reference build when the checkout exists)."""
    target = "all"
    if with_reference is False:
        target = "oracle"
    subprocess.check_call(["make", "-s", "-C", HERE, target])


def reference_available(variant="scalar"):
    return os.path.exists(os.path.join(REF_DIR, REF_VARIANTS[variant]))


def _fp(a):
    return a.ctypes.data_as(_f)


def _u8p(a):
    return a.ctypes.data_as(_u8)


class _Encoder:
    prefix = ""

    def __init__(self, path):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run make -C oracle)")
        self.lib = C.CDLL(path)
        p = self.prefix
        L = self.lib
        self._init = getattr(L, p + "init")
        self._init.argtypes = [C.c_int]
        self._compress = getattr(L, p + "compress_bc1")
        self._compress.restype = C.c_float
        self._decode = getattr(L, p + "decode_bc1")
        self._decode.argtypes = [C.c_void_p, _u8, C.c_int]
        self._eval = getattr(L, p + "evaluate_bc1_error")
        self._eval.restype = C.c_float
        self._eval.argtypes = [_u8, C.c_void_p, C.c_int]
        self._decode3 = getattr(L, p + "decode_bc3")
        self._decode3.argtypes = [C.c_void_p, _u8, C.c_int]
        self._eval3 = getattr(L, p + "evaluate_bc3_error")
        self._eval3.restype = C.c_float
        self._eval3.argtypes = [_u8, C.c_void_p, C.c_int, C.c_int]
        for name, res in [("reduce_colors", C.c_int), ("compute_sat", C.c_int),
                          ("principal_component", None), ("cluster_fit_four", None),
                          ("cluster_fit_three", None), ("output_block4", C.c_float),
                          ("output_block3", C.c_float), ("optimize_end_points4", C.c_int),
                          ("box_fit", None), ("refine_endpoints", C.c_float),
                          ("single_color", None), ("vector3_to_color16", C.c_uint32),
                          ("evaluate_palette", None), ("compute_indices4", C.c_uint32),
                          ("compute_indices3", C.c_uint32), ("compute_indices", C.c_uint32),
                          ("evaluate_mse_block", C.c_float), ("tables", None),
                          ("midpoints", None)]:
            fn = getattr(L, p + name)
            fn.restype = res
            setattr(self, "_" + name, fn)
        self.decoder = None

    def init(self, decoder=0):
        self._init(decoder)
        self.decoder = decoder

    # ---- whole-block API -------------------------------------------------
    def compress_block(self, level, colors, weights, cw=(1, 1, 1), tcm=False, tcb=False,
                       prior=None):
        colors = np.ascontiguousarray(colors, dtype=np.float32).reshape(64)
        weights = np.ascontiguousarray(weights, dtype=np.float32).reshape(16)
        cw = np.ascontiguousarray(cw, dtype=np.float32)
        out = np.zeros(8, np.uint8) if prior is None else np.array(prior, np.uint8).copy()
        e = self._call_compress(level, colors, weights, cw, tcm, tcb, out)
        return out, e

    def decode(self, block8, decoder=0):
        rgba = np.zeros(64, np.uint8)
        b = np.ascontiguousarray(block8, np.uint8)
        self._decode(b.ctypes.data, _u8p(rgba), decoder)
        return rgba

    def evaluate_error(self, rgba64, block8, decoder=0):
        r = np.ascontiguousarray(rgba64, np.uint8)
        b = np.ascontiguousarray(block8, np.uint8)
        return float(self._eval(_u8p(r), b.ctypes.data, decoder))

    def decode_bc3(self, block16, decoder=0):
        rgba = np.zeros(64, np.uint8)
        b = np.ascontiguousarray(block16, np.uint8)
        self._decode3(b.ctypes.data, _u8p(rgba), decoder)
        return rgba

    def evaluate_bc3_error(self, rgba64, block16, alpha_blend, decoder=0):
        r = np.ascontiguousarray(rgba64, np.uint8)
        b = np.ascontiguousarray(block16, np.uint8)
        return float(self._eval3(_u8p(r), b.ctypes.data, int(bool(alpha_blend)), decoder))

    # ---- stages ------------------------------------------------------------
    def reduce_colors(self, colors64, weights16, threshold):
        c = np.ascontiguousarray(colors64, np.float32)
        w = np.ascontiguousarray(weights16, np.float32)
        oc = np.zeros(48, np.float32)
        ow = np.zeros(16, np.float32)
        ab = C.c_int(0)
        n = self._reduce_colors(_fp(c), _fp(w), C.c_float(threshold), _fp(oc), _fp(ow),
                                C.byref(ab))
        return n, oc[:3 * n].reshape(n, 3), ow[:n], bool(ab.value)

    def compute_sat(self, rgb, w):
        n = len(w)
        c = np.zeros(48, np.float32)
        c[:3 * n] = np.asarray(rgb, np.float32).reshape(-1)
        ww = np.zeros(16, np.float32)
        ww[:n] = w
        out = np.zeros(64, np.float32)
        self._compute_sat(_fp(c), _fp(ww), n, _fp(out))
        return out

    def principal_component(self, rgb, w):
        n = len(w)
        c = np.zeros(48, np.float32)
        c[:3 * n] = np.asarray(rgb, np.float32).reshape(-1)
        ww = np.zeros(16, np.float32)
        ww[:n] = w
        out = np.zeros(3, np.float32)
        self._principal_component(_fp(c), _fp(ww), n, _fp(out))
        return out

    def cluster_fit(self, sat64, n, metric, four=True):
        s = np.ascontiguousarray(sat64, np.float32)
        m = np.ascontiguousarray(metric, np.float32)
        a = np.zeros(3, np.float32)
        b = np.zeros(3, np.float32)
        fn = self._cluster_fit_four if four else self._cluster_fit_three
        fn(_fp(s), n, _fp(m), _fp(a), _fp(b))
        return a, b

    def output_block(self, colors64, weights16, cw, v0, v1, four=True, allow_black=False):
        c = np.ascontiguousarray(colors64, np.float32)
        w = np.ascontiguousarray(weights16, np.float32)
        cw = np.ascontiguousarray(cw, np.float32)
        a = np.ascontiguousarray(v0, np.float32)
        b = np.ascontiguousarray(v1, np.float32)
        out = np.zeros(8, np.uint8)
        if four:
            e = self._output_block4(_fp(c), _fp(w), _fp(cw), _fp(a), _fp(b), out.ctypes.data_as(C.c_void_p))
        else:
            e = self._output_block3(_fp(c), _fp(w), _fp(cw), C.c_int(int(allow_black)), _fp(a),
                                    _fp(b), out.ctypes.data_as(C.c_void_p))
        return out, e

    def optimize_end_points4(self, indices, colors64):
        c = np.ascontiguousarray(colors64, np.float32)
        a = np.zeros(3, np.float32)
        b = np.zeros(3, np.float32)
        ok = self._optimize_end_points4(C.c_uint32(indices), _fp(c), _fp(a), _fp(b))
        return bool(ok), a, b

    def box_fit(self, rgb):
        n = len(rgb)
        c = np.zeros(48, np.float32)
        c[:3 * n] = np.asarray(rgb, np.float32).reshape(-1)
        a = np.zeros(3, np.float32)
        b = np.zeros(3, np.float32)
        self._box_fit(_fp(c), n, _fp(a), _fp(b))
        return a, b

    def refine(self, colors64, weights16, cw, tcm, err, block8):
        c = np.ascontiguousarray(colors64, np.float32)
        w = np.ascontiguousarray(weights16, np.float32)
        cw = np.ascontiguousarray(cw, np.float32)
        blk = np.array(block8, np.uint8).copy()
        args = [_fp(c), _fp(w), _fp(cw), C.c_int(int(tcm)), C.c_float(err),
                blk.ctypes.data_as(C.c_void_p)]
        if self.prefix == "oracle_":
            steps = C.c_int(0)
            e = self._refine_endpoints(*args, C.byref(steps))
        else:
            e = self._refine_endpoints(*args)
        return blk, e

    def single_color(self, rgb):
        r = np.ascontiguousarray(rgb, np.uint8)
        out = np.zeros(8, np.uint8)
        self._single_color(_u8p(r), out.ctypes.data_as(C.c_void_p))
        return out

    def vector3_to_color16(self, v):
        a = np.ascontiguousarray(v, np.float32)
        return int(self._vector3_to_color16(_fp(a)))

    def evaluate_palette(self, c0, c1):
        out = np.zeros(16, np.uint8)
        self._evaluate_palette(C.c_uint32(c0), C.c_uint32(c1), _u8p(out))
        return out

    def compute_indices(self, kind, colors64, cw, pal12, allow_black=False):
        c = np.ascontiguousarray(colors64, np.float32)
        cw = np.ascontiguousarray(cw, np.float32)
        p = np.ascontiguousarray(pal12, np.float32)
        if kind == 4:
            return int(self._compute_indices4(_fp(c), _fp(cw), _fp(p)))
        if kind == 3:
            return int(self._compute_indices3(_fp(c), _fp(cw), C.c_int(int(allow_black)), _fp(p)))
        return int(self._compute_indices(_fp(c), _fp(cw), _fp(p)))

    def evaluate_mse_block(self, colors64, weights16, cw, block8):
        c = np.ascontiguousarray(colors64, np.float32)
        w = np.ascontiguousarray(weights16, np.float32)
        cw = np.ascontiguousarray(cw, np.float32)
        b = np.ascontiguousarray(block8, np.uint8)
        return float(self._evaluate_mse_block(_fp(c), _fp(w), _fp(cw), b.ctypes.data_as(C.c_void_p)))

    def tables(self):
        four = np.zeros((976, 4), np.uint8)
        three = np.zeros((160, 4), np.uint8)
        ft = np.zeros(16, np.int32)
        tt = np.zeros(16, np.int32)
        m5 = np.zeros((256, 2), np.uint8)
        m6 = np.zeros((256, 2), np.uint8)
        w = np.zeros(5, np.float32)
        self._tables(_u8p(four), _u8p(three), ft.ctypes.data_as(_i), tt.ctypes.data_as(_i),
                     _u8p(m5), _u8p(m6), _fp(w))
        return dict(four=four, three=three, four_total=ft, three_total=tt, match5=m5,
                    match6=m6, weights=w)

    def midpoints(self):
        m5 = np.zeros(32, np.float32)
        m6 = np.zeros(64, np.float32)
        self._midpoints(_fp(m5), _fp(m6))
        return m5, m6


class Oracle(_Encoder):
    """The CPU restatement (icbc_oracle.c)."""
    prefix = "oracle_"

    def __init__(self, path=ORACLE_SO):
        super().__init__(path)
        L = self.lib
        self._compress.argtypes = [C.c_int, _f, _f, _f, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        self._many = L.oracle_compress_bc1_many
        self._many.restype = C.c_double
        self._many.argtypes = [C.c_int, _f, _f, _f, C.c_int, C.c_int, _u8, _f, C.c_void_p,
                               C.c_size_t, C.c_int]

    def _call_compress(self, level, colors, weights, cw, tcm, tcb, out):
        return float(self._compress(level, _fp(colors), _fp(weights), _fp(cw), int(tcm), int(tcb),
                                    out.ctypes.data_as(C.c_void_p), None))

    def compress_many(self, level, colors, weights=None, cw=(1, 1, 1), tcm=False, tcb=False,
                      nthreads=1, prior=None, with_stats=False):
        """Encode block-major colours (nblocks, 16, 4); returns (codes u8[n,8], errors, stats|None, seconds)."""
        colors = np.ascontiguousarray(colors, dtype=np.float32)
        n = colors.shape[0]
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        cw = np.ascontiguousarray(cw, dtype=np.float32)
        out = np.zeros((n, 8), np.uint8) if prior is None else np.array(prior, np.uint8).copy()
        err = np.zeros(n, np.float32)
        stats = np.zeros((n, 4), np.int32) if with_stats else None
        secs = self._many(level, _fp(colors), None if w is None else _fp(w), _fp(cw), int(tcm),
                          int(tcb), _u8p(out), _fp(err),
                          None if stats is None else stats.ctypes.data, n, nthreads)
        return out, err, stats, secs


class Reference(_Encoder):
    """The reference encoder compiled from its own header (oracle/_ref)."""
    prefix = "ref_"

    def __init__(self, variant="scalar"):
        super().__init__(os.path.join(REF_DIR, REF_VARIANTS[variant]))
        L = self.lib
        self.variant = variant
        self._compress.argtypes = [C.c_int, _f, _f, _f, C.c_int, C.c_int, C.c_void_p]
        self._many = L.ref_compress_bc1_many
        self._many.argtypes = [C.c_int, _f, _f, _f, C.c_int, C.c_int, _u8, _f, C.c_size_t]
        self._par = L.ref_compress_bc1_parallel
        self._par.restype = C.c_double
        self._par.argtypes = [C.c_int, _f, _f, _f, C.c_int, C.c_int, _u8, _f, C.c_size_t, C.c_int]
        self.simd_level = L.ref_simd_level()

    def _call_compress(self, level, colors, weights, cw, tcm, tcb, out):
        return float(self._compress(level, _fp(colors), _fp(weights), _fp(cw), int(tcm), int(tcb),
                                    out.ctypes.data_as(C.c_void_p)))

    def compress_many(self, level, colors, weights=None, cw=(1, 1, 1), tcm=False, tcb=False,
                      nthreads=0, prior=None):
        """Encode block-major colours; nthreads=0 runs the serial loop,
        otherwise the chunked multithreaded loop.  Returns (codes, errors, seconds|None)."""
        colors = np.ascontiguousarray(colors, dtype=np.float32)
        n = colors.shape[0]
        w = (np.ones((n, 16), np.float32) if weights is None
             else np.ascontiguousarray(weights, dtype=np.float32))
        cw = np.ascontiguousarray(cw, dtype=np.float32)
        out = np.zeros((n, 8), np.uint8) if prior is None else np.array(prior, np.uint8).copy()
        err = np.zeros(n, np.float32)
        if nthreads <= 0:
            self._many(level, _fp(colors), _fp(w), _fp(cw), int(tcm), int(tcb), _u8p(out), _fp(err), n)
            secs = None
        else:
            secs = self._par(level, _fp(colors), _fp(w), _fp(cw), int(tcm), int(tcb), _u8p(out),
                             _fp(err), n, nthreads)
        return out, err, secs
