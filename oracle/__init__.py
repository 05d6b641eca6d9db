"""ctypes bindings of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py, as the checker / the timed CPU baseline.  The product never imports it.
Arrays are sample-major (B, N); nets are per-row ELL (L, N, K).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_I32P = ctypes.POINTER(ctypes.c_int32)
_I64P = ctypes.POINTER(ctypes.c_int64)
_F32P = ctypes.POINTER(ctypes.c_float)
_F64P = ctypes.POINTER(ctypes.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing; build the oracle (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_plain_f32.restype = ctypes.c_int
        L.oracle_plain_f32.argtypes = [ctypes.c_int] * 3 + [_I32P, _F32P, ctypes.c_int, _F32P,
                                                            ctypes.c_float, ctypes.c_int64, _F32P,
                                                            _F32P, _I32P, _I64P, ctypes.c_int]
        L.oracle_plain_f32_ex.restype = ctypes.c_int
        L.oracle_plain_f32_ex.argtypes = L.oracle_plain_f32.argtypes + [_F64P]
        L.oracle_plain_f64.restype = ctypes.c_int
        L.oracle_plain_f64.argtypes = [ctypes.c_int] * 3 + [_I32P, _F32P, ctypes.c_int, _F32P,
                                                            ctypes.c_float, ctypes.c_int64, _F32P,
                                                            _F64P, _I32P, ctypes.c_int]
        L.oracle_layer_f64.restype = ctypes.c_int
        L.oracle_layer_f64.argtypes = [ctypes.c_int, ctypes.c_int, _I32P, _F32P, ctypes.c_int,
                                       _F32P, ctypes.c_float, ctypes.c_int64, _F32P, _F64P]
        L.oracle_snicit_exact.restype = ctypes.c_int
        L.oracle_snicit_exact.argtypes = [ctypes.c_int] * 3 + [
            _I32P, _F32P, ctypes.c_int, _F32P, ctypes.c_float, ctypes.c_int64, _F32P,
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, _I32P, _F32P, _I32P, _F32P,
            _I32P, _I32P, _I64P, _I64P, _I64P, ctypes.c_int]
        ex = list(L.oracle_snicit_exact.argtypes)
        L.oracle_snicit_exact_ex.restype = ctypes.c_int
        L.oracle_snicit_exact_ex.argtypes = ex[:14] + [ctypes.c_int] + ex[14:]
        model = [ctypes.c_int, _F32P, _F32P, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                 ctypes.c_int, _I32P, _F32P, ctypes.c_int, _F32P, ctypes.c_float, ctypes.c_int,
                 _F32P, _F32P, ctypes.c_int64, _F32P]
        L.oracle_model_plain.restype = ctypes.c_int
        L.oracle_model_plain.argtypes = model + [_F32P, _F32P, _I32P, ctypes.c_int]
        L.oracle_model_snicit_lossy.restype = ctypes.c_int
        L.oracle_model_snicit_lossy.argtypes = model + [
            ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
            ctypes.c_int, _I32P, _F32P, _F32P, _I32P, _I32P, _I32P, _I64P, _I64P, ctypes.c_int]
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _bias(bias, L, N):
    b = np.atleast_1d(np.ascontiguousarray(bias, dtype=np.float32)).ravel()
    kind = 0 if b.size == 1 else 1 if b.size == L else 2
    if kind == 2:
        assert b.size == L * N
    return kind, b


def plain_f32(cols, vals, bias, clip, y0, nthreads=0, want_y=True, timing=False):
    """Plain FP32 forward.  Returns (yL, cat, live_per_layer) and, with timing=True, also the
    wall seconds of the layer loop alone."""
    cols = np.ascontiguousarray(cols, np.int32)
    vals = np.ascontiguousarray(vals, np.float32)
    y0 = np.ascontiguousarray(y0, np.float32)
    L, N, K = cols.shape
    B = y0.shape[0]
    kind, b = _bias(bias, L, N)
    yL = np.empty((B, N), np.float32) if want_y else None
    cat = np.empty(B, np.int32)
    live = np.empty(L, np.int64)
    secs = np.zeros(1, np.float64)
    rc = lib().oracle_plain_f32_ex(N, L, K, _p(cols, _I32P), _p(vals, _F32P), kind,
                                   _p(b, _F32P), clip, B, _p(y0, _F32P), _p(yL, _F32P),
                                   _p(cat, _I32P), _p(live, _I64P), nthreads, _p(secs, _F64P))
    assert rc == 0
    if timing:
        return yL, cat, live, float(secs[0])
    return yL, cat, live


def plain_f64(cols, vals, bias, clip, y0, nthreads=0):
    cols = np.ascontiguousarray(cols, np.int32)
    vals = np.ascontiguousarray(vals, np.float32)
    y0 = np.ascontiguousarray(y0, np.float32)
    L, N, K = cols.shape
    B = y0.shape[0]
    kind, b = _bias(bias, L, N)
    yL = np.empty((B, N), np.float64)
    cat = np.empty(B, np.int32)
    rc = lib().oracle_plain_f64(N, L, K, _p(cols, _I32P), _p(vals, _F32P), kind, _p(b, _F32P),
                                clip, B, _p(y0, _F32P), _p(yL, _F64P), _p(cat, _I32P), nthreads)
    assert rc == 0
    return yL, cat


def layer_f64(cols_l, vals_l, bias_l, clip, yin):
    cols_l = np.ascontiguousarray(cols_l, np.int32)
    vals_l = np.ascontiguousarray(vals_l, np.float32)
    yin = np.ascontiguousarray(yin, np.float32)
    N, K = cols_l.shape
    b = np.atleast_1d(np.ascontiguousarray(bias_l, np.float32)).ravel()
    kind = 0 if b.size == 1 else 2
    out = np.empty((yin.shape[0], N), np.float64)
    rc = lib().oracle_layer_f64(N, K, _p(cols_l, _I32P), _p(vals_l, _F32P), kind, _p(b, _F32P),
                                clip, yin.shape[0], _p(yin, _F32P), _p(out, _F64P))
    assert rc == 0
    return out


def snicit_exact(cols, vals, bias, clip, y0, t, sample_size=32, segments=16, eps=0.03,
                 nthreads=0, want_y=True, ref_segments=False):
    """Exact SNICIT restatement.  Returns a dict with cat, yL, pool, F, centroids, assign,
    live (per pre layer), nerow (per post layer) and info (B', pool, k, nerow at t).
    ref_segments=True sums the pool rows over the reference's overlapping segments
    (quirk Q2, src/SDGC/SNICIT/kernel.hpp:147-148) instead of N/segments-neuron ones."""
    cols = np.ascontiguousarray(cols, np.int32)
    vals = np.ascontiguousarray(vals, np.float32)
    y0 = np.ascontiguousarray(y0, np.float32)
    L, N, K = cols.shape
    B = y0.shape[0]
    kind, b = _bias(bias, L, N)
    cat = np.empty(B, np.int32)
    yL = np.empty((B, N), np.float32) if want_y else None
    pool = np.full(64, -1, np.int32)
    F = np.zeros(64 * segments, np.float32)
    cent = np.full(64, -1, np.int32)
    assign = np.empty(B, np.int32)
    live = np.full(L, -1, np.int64)
    nerow = np.zeros(L, np.int64)
    info = np.zeros(4, np.int64)
    rc = lib().oracle_snicit_exact_ex(N, L, K, _p(cols, _I32P), _p(vals, _F32P), kind,
                                      _p(b, _F32P), clip, B, _p(y0, _F32P), t, sample_size,
                                      segments, eps, int(bool(ref_segments)),
                                      _p(cat, _I32P), _p(yL, _F32P), _p(pool, _I32P),
                                      _p(F, _F32P), _p(cent, _I32P), _p(assign, _I32P),
                                      _p(live, _I64P), _p(nerow, _I64P), _p(info, _I64P),
                                      nthreads)
    assert rc == 0, rc
    Bp, P, k, ne0 = (int(x) for x in info)
    return {"cat": cat, "yL": yL, "pool": pool[:P], "F": F[:P * segments].reshape(P, segments),
            "centroids": cent[:k], "assign": assign, "live": live[:t], "nerow": nerow[t:],
            "live_at_t": Bp, "k": k, "nerow_at_t": ne0}


def _model_args(cols, vals, bias, clip, w_in, b_in, in_clip, w_out, b_out):
    cols = np.ascontiguousarray(cols, np.int32)
    vals = np.ascontiguousarray(vals, np.float32)
    L, N, K = cols.shape
    kind, b = _bias(bias, L, N)
    Nin = 0 if w_in is None else w_in.shape[1]
    Nout = 0 if w_out is None else w_out.shape[0]
    wi = np.ascontiguousarray(w_in if Nin else np.zeros((1, 1)), np.float32)
    bi = np.ascontiguousarray(b_in if Nin else np.zeros(1), np.float32)
    wo = np.ascontiguousarray(w_out if Nout else np.zeros((1, 1)), np.float32)
    bo = np.ascontiguousarray(b_out if Nout else np.zeros(1), np.float32)
    keep = (cols, vals, b, wi, bi, wo, bo)
    args = [Nin, _p(wi, _F32P), _p(bi, _F32P), in_clip, N, L, K, _p(cols, _I32P),
            _p(vals, _F32P), kind, _p(b, _F32P), clip, Nout, _p(wo, _F32P), _p(bo, _F32P)]
    return args, keep, N, Nout


def model_plain(cols, vals, bias, clip, x0, w_in=None, b_in=None, in_clip=1.0, w_out=None,
                b_out=None, nthreads=0):
    """Plain forward of a model with optional dense input/output layers.
    Returns (yL (B,N), logits (B,Nout) or None, cat)."""
    args, keep, N, Nout = _model_args(cols, vals, bias, clip, w_in, b_in, in_clip, w_out, b_out)
    x0 = np.ascontiguousarray(x0, np.float32)
    B = x0.shape[0]
    yL = np.empty((B, N), np.float32)
    lg = np.empty((B, max(Nout, 1)), np.float32)
    cat = np.empty(B, np.int32)
    rc = lib().oracle_model_plain(*args, B, _p(x0, _F32P), _p(yL, _F32P), _p(lg, _F32P),
                                  _p(cat, _I32P), nthreads)
    assert rc == 0, rc
    return yL, (lg if Nout else None), cat


def model_snicit_lossy(cols, vals, bias, clip, x0, t, w_in=None, b_in=None, in_clip=1.0,
                       w_out=None, b_out=None, pool=128, eps_pool=0.03, eps_assign=0.04,
                       eps_post=0.05, positive_only=False, nthreads=0):
    """Lossy SNICIT restatement.  Returns a dict with cat, yL, logits, pool, centroids,
    assign, nerow (per post layer), k, nerow_at_t."""
    args, keep, N, Nout = _model_args(cols, vals, bias, clip, w_in, b_in, in_clip, w_out, b_out)
    x0 = np.ascontiguousarray(x0, np.float32)
    B = x0.shape[0]
    L = keep[0].shape[0]
    cat = np.empty(B, np.int32)
    yL = np.empty((B, N), np.float32)
    lg = np.empty((B, max(Nout, 1)), np.float32)
    pc = np.full(pool, -1, np.int32)
    cc = np.full(pool, -1, np.int32)
    assign = np.empty(B, np.int32)
    nerow = np.zeros(L, np.int64)
    info = np.zeros(4, np.int64)
    rc = lib().oracle_model_snicit_lossy(*args, B, _p(x0, _F32P), t, pool, eps_pool, eps_assign,
                                         eps_post, int(positive_only), _p(cat, _I32P),
                                         _p(yL, _F32P), _p(lg, _F32P), _p(pc, _I32P),
                                         _p(cc, _I32P), _p(assign, _I32P), _p(nerow, _I64P),
                                         _p(info, _I64P), nthreads)
    assert rc == 0, rc
    _, P, k, ne0 = (int(x) for x in info)
    return {"cat": cat, "yL": yL, "logits": lg if Nout else None, "pool": pc[:P],
            "centroids": cc[:k], "assign": assign, "nerow": nerow[t:], "k": k,
            "nerow_at_t": ne0}
