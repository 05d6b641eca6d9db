"""snicit_amd -- MI355X (gfx950) SNICIT sparse-DNN inference layer loop.

Python mirror of the reference's in-process interface over the C-ABI of libsnicit.so
(include/snicit.h).  The reference's classes take the net in the constructor and run the
layer loop in ``infer``:

    SNICIT_SDGC::SNICIT(weight_path, bias, N, L, threshold).infer(input, golden, num_inputs, batch)
        (src/SDGC/SNICIT/SNICIT.hpp:43-56, 87-203)
    SNICIT_SDGC::BF(weight_path, bias, N, L).infer(input, golden, num_inputs)
        (src/SDGC/BF/BF.hpp:55-68, 102-124)
    SNICIT_SDGC::SNIG / XY  (src/SDGC/SNIG/SNIG.hpp:51-66, src/SDGC/XY/XY.hpp:41-54)

Here the classes take a resident :class:`Net` (weights stay in HBM) and ``infer`` takes the
input rows, returning the per-sample categories.  All compute runs in the HIP library;
there is no CPU fallback: importing works without a GPU, but every device call raises if
the library or the device is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsnicit.so")

E_TIMEOUT = -6
MODE_SNICIT, MODE_BF, MODE_SNIG, MODE_XY, MODE_PLAIN_DENSE = 0, 1, 2, 3, 4
MODES = {"SNICIT": MODE_SNICIT, "BF": MODE_BF, "SNIG": MODE_SNIG, "XY": MODE_XY,
         "PLAIN_DENSE": MODE_PLAIN_DENSE}
BIAS_SCALAR, BIAS_PER_LAYER, BIAS_PER_NEURON = 0, 1, 2
CLUSTER_EXACT, CLUSTER_LOSSY = 0, 1

# Graph Challenge biases per N (main/SDGC.cu:49-54).
SDGC_BIAS = {1024: -0.3, 4096: -0.35, 16384: -0.4, 65536: -0.45}


class SnicitError(RuntimeError):
    """A failed library call; ``code`` is the negative SNICIT_E_* status."""

    def __init__(self, msg, code=0):
        super().__init__(msg)
        self.code = code


class RunCfg(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int32), ("threshold", ctypes.c_int32), ("cluster", ctypes.c_int32),
        ("sample_size", ctypes.c_int32), ("segments", ctypes.c_int32), ("eps", ctypes.c_float),
        ("lossy_pool", ctypes.c_int32), ("lossy_eps_pool", ctypes.c_float),
        ("lossy_eps_assign", ctypes.c_float), ("lossy_eps_post", ctypes.c_float),
        ("use_graph", ctypes.c_int32), ("lossy_positive_only", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 6),
    ]


class TuneInfo(ctypes.Structure):
    _fields_ = [("grid_target", ctypes.c_int64), ("narrow", ctypes.c_int32),
                ("ell_lds_kb", ctypes.c_int32), ("ell_chunked", ctypes.c_int32),
                ("from_cache", ctypes.c_int32), ("ms_default", ctypes.c_double),
                ("ms_tuned", ctypes.c_double), ("key", ctypes.c_char * 96),
                ("post_grid_target", ctypes.c_int64), ("post_narrow", ctypes.c_int32),
                ("ms_post_default", ctypes.c_double), ("ms_post_tuned", ctypes.c_double),
                ("ms_pre_default", ctypes.c_double), ("ms_pre_tuned", ctypes.c_double),
                ("ms_run_default", ctypes.c_double), ("ms_run_tuned", ctypes.c_double),
                ("post_form", ctypes.c_int32), ("sample_size", ctypes.c_int32),
                ("placements", ctypes.c_int32), ("placement", ctypes.c_int32),
                ("compact_slack", ctypes.c_int32), ("plain_tiles", ctypes.c_int32)]

    def as_dict(self):
        return {"grid_target": self.grid_target, "narrow": self.narrow,
                "ell_lds_kb": self.ell_lds_kb, "ell_chunked": self.ell_chunked,
                "from_cache": bool(self.from_cache), "ms_default": self.ms_default,
                "ms_tuned": self.ms_tuned, "key": self.key.decode(),
                "post_grid_target": self.post_grid_target, "post_narrow": self.post_narrow,
                "ms_post_default": self.ms_post_default, "ms_post_tuned": self.ms_post_tuned,
                "ms_pre_default": self.ms_pre_default, "ms_pre_tuned": self.ms_pre_tuned,
                "ms_run_default": self.ms_run_default, "ms_run_tuned": self.ms_run_tuned,
                "post_form": self.post_form, "sample_size": self.sample_size,
                "placements": self.placements, "placement": self.placement,
                "compact_slack": self.compact_slack, "plain_tiles": self.plain_tiles}


class ServerCfg(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("slot_samples", ctypes.c_int64),
                ("slot_segments", ctypes.c_int32), ("max_slots_per_net", ctypes.c_int32),
                ("cluster_min", ctypes.c_int32), ("warm", ctypes.c_int32),
                ("group_window_ms", ctypes.c_double), ("group_samples", ctypes.c_int64),
                ("pinned_bytes", ctypes.c_int64), ("round_ms", ctypes.c_double)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("ms_queue", ctypes.c_double), ("ms_run", ctypes.c_double),
                ("ms_total", ctypes.c_double), ("group_samples", ctypes.c_int64),
                ("group_requests", ctypes.c_int32), ("slot", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [("requests_done", ctypes.c_int64), ("samples_done", ctypes.c_int64),
                ("groups", ctypes.c_int64), ("device_bytes", ctypes.c_int64),
                ("slots", ctypes.c_int32), ("queued", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Segment(ctypes.Structure):
    _fields_ = [("net", ctypes.c_int32), ("pad", ctypes.c_int32), ("samples", ctypes.c_int64)]


class Stats(ctypes.Structure):
    _fields_ = [
        ("ms_total", ctypes.c_double), ("ms_pre", ctypes.c_double),
        ("ms_cluster", ctypes.c_double), ("ms_post", ctypes.c_double),
        ("ms_recover", ctypes.c_double), ("edges", ctypes.c_double),
        ("batch", ctypes.c_int64), ("live_at_threshold", ctypes.c_int64),
        ("nerow_first", ctypes.c_int64), ("nerow_last", ctypes.c_int64),
        ("active", ctypes.c_int64), ("centroids", ctypes.c_int32), ("layers", ctypes.c_int32),
        ("rows_processed", ctypes.c_int64), ("reserved", ctypes.c_int64 * 6),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


# Every symbol the header declares, with its ctypes signature.
_P = ctypes.c_void_p
_I32P = ctypes.POINTER(ctypes.c_int32)
_F32P = ctypes.POINTER(ctypes.c_float)
_I64P = ctypes.POINTER(ctypes.c_int64)
SIGNATURES = {
    "snicit_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
    "snicit_ctx_destroy": (ctypes.c_int, [_P]),
    "snicit_ctx_set_grouping": (ctypes.c_int, [_P, ctypes.c_int32]),
    "snicit_last_error": (ctypes.c_char_p, []),
    "snicit_run_cfg_default": (None, [ctypes.POINTER(RunCfg)]),
    "snicit_net_create_ell": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                             _I32P, _F32P, ctypes.c_int32, _F32P, ctypes.c_float,
                                             ctypes.POINTER(_P)]),
    "snicit_net_generate_sdgc": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32,
                                                ctypes.c_uint64, ctypes.c_float, ctypes.c_float,
                                                ctypes.c_float, ctypes.POINTER(_P)]),
    "snicit_net_create_csr": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, _I64P, _I32P,
                                             _F32P, ctypes.c_int32, ctypes.c_int32, _F32P,
                                             ctypes.c_float, ctypes.POINTER(_P)]),
    "snicit_net_set_io": (ctypes.c_int, [_P, ctypes.c_int32, _F32P, _F32P, ctypes.c_float,
                                         ctypes.c_int32, _F32P, _F32P]),
    "snicit_net_stream_weights": (ctypes.c_int, [_P, ctypes.c_int32]),
    "snicit_net_destroy": (ctypes.c_int, [_P]),
    "snicit_net_info": (ctypes.c_int, [_P, _I32P, _I32P, _I32P,
                                       ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(ctypes.c_int64)]),
    "snicit_plan_create": (ctypes.c_int, [_P, ctypes.POINTER(RunCfg), ctypes.c_int64,
                                          ctypes.POINTER(_P)]),
    "snicit_plan_destroy": (ctypes.c_int, [_P]),
    "snicit_plan_set_input": (ctypes.c_int, [_P, _F32P, ctypes.c_int64]),
    "snicit_plan_set_input_images": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_int64,
                                                    ctypes.c_int64]),
    "snicit_plan_run": (ctypes.c_int, [_P, ctypes.POINTER(Stats)]),
    "snicit_plan_run_layers": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(Stats)]),
    "snicit_plan_categories": (ctypes.c_int, [_P, _I32P]),
    "snicit_plan_activations": (ctypes.c_int, [_P, _F32P]),
    "snicit_plan_trace": (ctypes.c_int64, [_P, ctypes.c_int32, _P, ctypes.c_int64]),
    "snicit_plan_profile": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_double), ctypes.c_int64]),
    "snicit_infer": (ctypes.c_int, [_P, ctypes.POINTER(RunCfg), _F32P, ctypes.c_int64,
                                    ctypes.c_int64, _I32P, ctypes.POINTER(Stats)]),
    "snicit_infer_noalloc": (ctypes.c_int, [_P, ctypes.POINTER(RunCfg), _F32P, ctypes.c_int64,
                                            ctypes.c_int64, _I32P, ctypes.POINTER(Stats)]),
    "snicit_plan_autotune": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.POINTER(TuneInfo)]),
    "snicit_plan_autotune_ex": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_uint32,
                                               ctypes.POINTER(TuneInfo)]),
    "snicit_plan_set_shape": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_int32]),
    "snicit_plan_set_post_shape": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32]),
    "snicit_plan_set_post_form": (ctypes.c_int, [_P, ctypes.c_int32]),
    "snicit_plan_set_compact_slack": (ctypes.c_int, [_P, ctypes.c_int32]),
    "snicit_plan_set_plain_tiles": (ctypes.c_int, [_P, ctypes.c_int32]),
    "snicit_batch_autotune": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.POINTER(TuneInfo)]),
    "snicit_server_cfg_default": (None, [ctypes.POINTER(ServerCfg)]),
    "snicit_server_create": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(RunCfg),
                                            ctypes.c_int32, ctypes.POINTER(ServerCfg),
                                            ctypes.POINTER(_P)]),
    "snicit_server_destroy": (ctypes.c_int, [_P]),
    "snicit_server_submit": (ctypes.c_int, [_P, ctypes.c_int32, _F32P, ctypes.c_int64,
                                            ctypes.POINTER(ctypes.c_uint64)]),
    "snicit_server_submit_images": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_uint64,
                                                   ctypes.c_int64, ctypes.c_int64,
                                                   ctypes.POINTER(ctypes.c_uint64)]),
    "snicit_server_wait": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_double, _I32P,
                                          ctypes.POINTER(RequestInfo)]),
    "snicit_server_stats_get": (ctypes.c_int, [_P, ctypes.POINTER(ServerStats)]),
    "snicit_batch_create": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(RunCfg),
                                           ctypes.c_int32, ctypes.POINTER(Segment), ctypes.c_int32,
                                           ctypes.POINTER(_P)]),
    "snicit_batch_destroy": (ctypes.c_int, [_P]),
    "snicit_batch_set_input": (ctypes.c_int, [_P, ctypes.c_int32, _F32P]),
    "snicit_batch_set_input_images": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_uint64,
                                                     ctypes.c_int64]),
    "snicit_batch_run": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(Stats)]),
    "snicit_batch_categories": (ctypes.c_int, [_P, ctypes.c_int32, _I32P]),
    "snicit_batch_activations": (ctypes.c_int, [_P, ctypes.c_int32, _F32P]),
    "snicit_gen_sdgc_ell": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, _I32P]),
    "snicit_gen_images": (ctypes.c_int, [ctypes.c_int32, ctypes.c_uint64, ctypes.c_int64,
                                         ctypes.c_int64, _F32P]),
    "snicit_gen_unstructured": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64,
                                               _I32P, _F32P, _F32P, _F32P]),
    "snicit_gen_features": (ctypes.c_int, [ctypes.c_int32, ctypes.c_uint64, ctypes.c_int64,
                                           ctypes.c_int64, _F32P]),
}

_lib = None


def lib():
    """Loads libsnicit.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SnicitError(f"{LIB_PATH} is missing; run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().snicit_last_error().decode(errors="replace")
        raise SnicitError(f"snicit error {rc}: {msg}", rc)


def _i32(a):
    return a.ctypes.data_as(_I32P)


def _f32(a):
    return a.ctypes.data_as(_F32P)


def default_cfg(mode="SNICIT", threshold=30, **kw):
    cfg = RunCfg()
    lib().snicit_run_cfg_default(ctypes.byref(cfg))
    cfg.mode = MODES[mode] if isinstance(mode, str) else int(mode)
    cfg.threshold = int(threshold)
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


# ---- host generators ---------------------------------------------------------------
def gen_sdgc_cols(N, L, seed):
    """Columns (L, N, 32) of the seeded SDGC-like net (values are all the same weight)."""
    cols = np.empty((L, N, 32), dtype=np.int32)
    _check(lib().snicit_gen_sdgc_ell(N, L, seed, _i32(cols)))
    return cols


def gen_images(N, seed, first_sample, B):
    """(B, N) float32 binary blob images of samples first_sample .. first_sample+B-1."""
    out = np.empty((B, N), dtype=np.float32)
    _check(lib().snicit_gen_images(N, seed, first_sample, B, _f32(out)))
    return out


def gen_unstructured(Nin, N, L, K, Nout, seed):
    """Seeded unstructured MLP: (cols (L,N,K), vals (L,N,K), w_in (N,Nin) or None,
    w_out (Nout,N) or None)."""
    cols = np.empty((L, N, K), np.int32)
    vals = np.empty((L, N, K), np.float32)
    w_in = np.empty((N, max(Nin, 1)), np.float32)
    w_out = np.empty((max(Nout, 1), N), np.float32)
    _check(lib().snicit_gen_unstructured(Nin, N, L, K, Nout, seed, _i32(cols), _f32(vals),
                                         _f32(w_in), _f32(w_out)))
    return cols, vals, (w_in if Nin else None), (w_out if Nout else None)


def gen_features(Nin, seed, first_sample, B):
    out = np.empty((B, Nin), np.float32)
    _check(lib().snicit_gen_features(Nin, seed, first_sample, B, _f32(out)))
    return out


# ---- device objects --------------------------------------------------------------------
class Context:
    """One device and one HIP stream (replaces GpuEnv, src/SDGC/gpu_lib/gpu_env.h:11-89)."""

    def __init__(self, device=0):
        h = _P()
        _check(lib().snicit_ctx_create(device, ctypes.byref(h)))
        self.h = h

    GROUP_NONE, GROUP_CONSECUTIVE, GROUP_SEARCH = 0, 1, 2

    def set_grouping(self, grouping):
        """Panel grouping of short rows in the nets created next (see snicit_ctx_set_grouping)."""
        _check(lib().snicit_ctx_set_grouping(self.h, grouping))

    def close(self):
        if self.h:
            lib().snicit_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Net:
    """A resident net.  ``from_ell`` replaces the TSV weight ingestion of
    Base::_load_weight (src/SDGC/base.hpp:89-141); ``generate_sdgc`` builds the seeded
    SDGC-like net on the device."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle
        n, l, k = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        p, b = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().snicit_net_info(self.h, ctypes.byref(n), ctypes.byref(l), ctypes.byref(k),
                                     ctypes.byref(p), ctypes.byref(b)))
        self.N, self.L, self.K = n.value, l.value, k.value
        self.panels, self.device_bytes = p.value, b.value
        self.Nin, self.Nout = 0, 0

    @classmethod
    def from_ell(cls, ctx, cols, vals, bias, clip=32.0):
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float32)
        L, N, K = cols.shape
        assert vals.shape == cols.shape
        b = np.atleast_1d(np.ascontiguousarray(bias, dtype=np.float32))
        if b.size == 1:
            kind = BIAS_SCALAR
        elif b.size == L:
            kind = BIAS_PER_LAYER
        elif b.size == L * N:
            kind = BIAS_PER_NEURON
        else:
            raise ValueError("bias must have 1, L or L*N entries")
        h = _P()
        _check(lib().snicit_net_create_ell(ctx.h, N, L, K, _i32(cols), _f32(vals), kind,
                                           _f32(b.ravel()), clip, ctypes.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_csr(cls, ctx, N, row_ptr, cols, vals, bias, clip=32.0, rows_are_inputs=False):
        """Hidden layers as CSR: row_ptr (L, N+1) absolute offsets into cols / vals.  With
        rows_are_inputs, row i lists the outputs that input i feeds (the reference's beyond
        file orientation)."""
        rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
        L = rp.shape[0]
        assert rp.shape == (L, N + 1)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float32)
        b = np.atleast_1d(np.ascontiguousarray(bias, dtype=np.float32))
        kind = {1: BIAS_SCALAR, L: BIAS_PER_LAYER, L * N: BIAS_PER_NEURON}.get(b.size)
        if kind is None:
            raise ValueError("bias must have 1, L or L*N entries")
        h = _P()
        _check(lib().snicit_net_create_csr(ctx.h, N, L, rp.ctypes.data_as(_I64P),
                                           _i32(cols if cols.size else np.zeros(1, np.int32)),
                                           _f32(vals if vals.size else np.zeros(1, np.float32)),
                                           int(rows_are_inputs), kind, _f32(b.ravel()), clip,
                                           ctypes.byref(h)))
        return cls(ctx, h)

    @classmethod
    def generate_sdgc(cls, ctx, N, L, seed, bias=None, weight=0.0625, clip=32.0):
        if bias is None:
            bias = SDGC_BIAS[N]
        h = _P()
        _check(lib().snicit_net_generate_sdgc(ctx.h, N, L, seed, weight, bias, clip,
                                              ctypes.byref(h)))
        return cls(ctx, h)

    def set_io(self, w_in=None, b_in=None, in_clip=1.0, w_out=None, b_out=None):
        """Dense input layer (N x Nin, clip [0, in_clip]) and dense output layer (Nout x N,
        linear; category = argmax).  Call before creating plans."""
        Nin = 0 if w_in is None else w_in.shape[1]
        Nout = 0 if w_out is None else w_out.shape[0]
        wi = np.ascontiguousarray(w_in if w_in is not None else np.zeros((1, 1)), np.float32)
        bi = np.ascontiguousarray(b_in if b_in is not None else np.zeros(max(self.N, 1)),
                                  np.float32)
        wo = np.ascontiguousarray(w_out if w_out is not None else np.zeros((1, 1)), np.float32)
        bo = np.ascontiguousarray(b_out if b_out is not None else np.zeros(max(Nout, 1)),
                                  np.float32)
        _check(lib().snicit_net_set_io(self.h, Nin, _f32(wi), _f32(bi), in_clip, Nout, _f32(wo),
                                       _f32(bo)))
        self.Nin, self.Nout = Nin, Nout

    def stream_weights(self, buffers=2):
        """Move the hidden layers' weights to pinned host memory; plans copy each layer into
        one of `buffers` device slots as it runs.  Call before creating plans."""
        _check(lib().snicit_net_stream_weights(self.h, buffers))
        b = ctypes.c_int64()
        _check(lib().snicit_net_info(self.h, None, None, None, None, ctypes.byref(b)))
        self.device_bytes = b.value

    def close(self):
        if self.h:
            lib().snicit_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """Activations of up to ``max_batch`` samples for one mode on one net."""

    (TRACE_LIVE, TRACE_NEROW, TRACE_POOL, TRACE_F, TRACE_CENTROIDS, TRACE_ASSIGN,
     TRACE_LOGITS) = range(7)

    def __init__(self, net, cfg, max_batch):
        self.net = net
        self.cfg = cfg
        h = _P()
        _check(lib().snicit_plan_create(net.h, ctypes.byref(cfg), max_batch, ctypes.byref(h)))
        self.h = h
        self.B = 0

    def set_input(self, y0):
        y0 = np.ascontiguousarray(y0, dtype=np.float32)
        assert y0.ndim == 2 and y0.shape[1] == (self.net.Nin or self.net.N)
        _check(lib().snicit_plan_set_input(self.h, _f32(y0), y0.shape[0]))
        self.B = y0.shape[0]

    def set_input_images(self, seed, first_sample, B):
        _check(lib().snicit_plan_set_input_images(self.h, seed, first_sample, B))
        self.B = B

    def run(self, stop_layer=None):
        st = Stats()
        if stop_layer is None:
            _check(lib().snicit_plan_run(self.h, ctypes.byref(st)))
        else:
            _check(lib().snicit_plan_run_layers(self.h, stop_layer, ctypes.byref(st)))
        return st

    def categories(self):
        out = np.empty(self.B, dtype=np.int32)
        _check(lib().snicit_plan_categories(self.h, _i32(out)))
        return out

    TUNE_SAMPLE, TUNE_PLACEMENT = 1, 2

    def autotune(self, cache_path=None, sample=False, placement=False):
        """Measures the layer-kernel launch shape on this plan's input (or takes it from the
        JSON cache); sample: also the exact clustering sample size (categories checked equal
        to the configured size's); placement: also up to four fresh allocations of the
        activation buffers.  Returns the TuneInfo as a dict."""
        info = TuneInfo()
        path = None if cache_path is None else os.fsencode(cache_path)
        flags = (self.TUNE_SAMPLE if sample else 0) | (self.TUNE_PLACEMENT if placement else 0)
        _check(lib().snicit_plan_autotune_ex(self.h, path, flags, ctypes.byref(info)))
        return info.as_dict()

    def set_shape(self, grid_target, narrow, ell_lds_kb=80, ell_chunked=2):
        _check(lib().snicit_plan_set_shape(self.h, grid_target, narrow, ell_lds_kb, ell_chunked))

    def set_post_shape(self, post_grid_target, post_narrow=64):
        _check(lib().snicit_plan_set_post_shape(self.h, post_grid_target, post_narrow))

    # post layers with at most two distinct weight rows per panel: LDS-staged kernel,
    # wavefront-local kernel, or wavefront-local kernel reading centroid chain records that
    # the launch of the layer before computed (default)
    POST_LDS, POST_WAVE, POST_WAVE_CHAINS = 0, 1, 2

    def set_post_form(self, form):
        _check(lib().snicit_plan_set_post_form(self.h, form))

    def set_plain_tiles(self, tiles):
        """64-column tiles per wavefront of the plain vector kernel: 1, 2 or 4."""
        _check(lib().snicit_plan_set_plain_tiles(self.h, tiles))

    def set_compact_slack(self, slack):
        """Dead columns per 1024 a plain layer's compaction may leave in the list (0: exact)."""
        _check(lib().snicit_plan_set_compact_slack(self.h, slack))

    def activations(self):
        out = np.empty((self.B, self.net.N), dtype=np.float32)
        _check(lib().snicit_plan_activations(self.h, _f32(out)))
        return out

    def trace(self, what):
        n = lib().snicit_plan_trace(self.h, what, None, 0)
        if n < 0:
            _check(int(n))
        dtype = {0: np.int64, 1: np.int64, 2: np.int32, 3: np.float32, 4: np.int32,
                 5: np.int32, 6: np.float32}[what]
        out = np.empty(max(int(n), 1), dtype=dtype)
        m = lib().snicit_plan_trace(self.h, what, out.ctypes.data_as(_P), out.size)
        if m < 0:
            _check(int(m))
        return out[:int(n)]

    def profile(self):
        """Per-layer SpMM launch durations (ms), rows and centroid-launch ms (eager replay)."""
        L = self.net.L
        ms = np.zeros(L, np.float64)
        rows = np.zeros(L, np.int64)
        cms = np.zeros(L, np.float64)
        _check(lib().snicit_plan_profile(self.h, ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                         rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                         cms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), L))
        return ms, rows, cms

    def close(self):
        if self.h:
            lib().snicit_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Model:
    """A beyond-style MLP on one context: optional dense input layer, hidden layers (ELL),
    optional dense output layer with argmax categories (src/beyond/SNICIT/SNICIT.hpp:59-74)."""

    def __init__(self, ctx, cols, vals, bias, clip=1.0, w_in=None, b_in=None, in_clip=1.0,
                 w_out=None, b_out=None):
        self.net = Net.from_ell(ctx, cols, vals, bias, clip=clip)
        if w_in is not None or w_out is not None:
            self.net.set_io(w_in, b_in, in_clip, w_out, b_out)


class Batch:
    """Uneven batch segments over several nets, run as one captured graph.

    ``nets`` and ``cfgs`` pair up; ``segments`` is a list of (net index, samples)."""

    def __init__(self, ctx, nets, cfgs, segments):
        self.nets = list(nets)
        self.segments = [(int(n), int(b)) for n, b in segments]
        hs = (_P * len(nets))(*[n.h for n in nets])
        cf = (RunCfg * len(cfgs))(*cfgs)
        sg = (Segment * len(segments))(*[Segment(n, 0, b) for n, b in self.segments])
        h = _P()
        _check(lib().snicit_batch_create(ctx.h, hs, cf, len(nets), sg, len(segments),
                                         ctypes.byref(h)))
        self.h = h

    def set_input(self, seg, y0):
        net, B = self.segments[seg]
        y0 = np.ascontiguousarray(y0, dtype=np.float32)
        n = self.nets[net]
        assert y0.shape == (B, n.Nin or n.N)
        _check(lib().snicit_batch_set_input(self.h, seg, _f32(y0)))

    def set_input_images(self, seg, seed, first_sample):
        _check(lib().snicit_batch_set_input_images(self.h, seg, seed, first_sample))

    def run(self):
        """Returns (ms, per-net Stats list)."""
        ms = ctypes.c_double()
        st = (Stats * len(self.nets))()
        _check(lib().snicit_batch_run(self.h, ctypes.byref(ms), st))
        return ms.value, list(st)

    def autotune(self, cache_path=None):
        info = (TuneInfo * len(self.nets))()
        path = None if cache_path is None else os.fsencode(cache_path)
        _check(lib().snicit_batch_autotune(self.h, path, info))
        return [i.as_dict() for i in info]

    def categories(self, seg):
        out = np.empty(self.segments[seg][1], np.int32)
        _check(lib().snicit_batch_categories(self.h, seg, _i32(out)))
        return out

    def activations(self, seg):
        net, B = self.segments[seg]
        out = np.empty((B, self.nets[net].N), np.float32)
        _check(lib().snicit_batch_activations(self.h, seg, _f32(out)))
        return out

    def close(self):
        if self.h:
            lib().snicit_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def server_cfg(**kw):
    cfg = ServerCfg()
    lib().snicit_server_cfg_default(ctypes.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


class Server:
    """Continuous-batching request loop over resident nets (one worker thread, slots of
    captured stage graphs per net).  ``submit`` returns a request id; ``wait`` returns
    (categories, info dict)."""

    def __init__(self, ctx, nets, cfgs, cfg=None, **kw):
        self.nets = list(nets)
        self.cfg = cfg if cfg is not None else server_cfg(**kw)
        hs = (_P * len(nets))(*[n.h for n in nets])
        cf = (RunCfg * len(cfgs))(*cfgs)
        h = _P()
        _check(lib().snicit_server_create(ctx.h, hs, cf, len(nets), ctypes.byref(self.cfg),
                                          ctypes.byref(h)))
        self.h = h
        self._sizes = {}

    def submit(self, net, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = self.nets[net]
        assert x.ndim == 2 and x.shape[1] == (n.Nin or n.N)
        rid = ctypes.c_uint64()
        _check(lib().snicit_server_submit(self.h, net, _f32(x), x.shape[0], ctypes.byref(rid)))
        self._sizes[rid.value] = x.shape[0]
        return rid.value

    def submit_images(self, net, seed, first_sample, samples):
        rid = ctypes.c_uint64()
        _check(lib().snicit_server_submit_images(self.h, net, seed, first_sample, samples,
                                                 ctypes.byref(rid)))
        self._sizes[rid.value] = samples
        return rid.value

    def wait(self, rid, timeout_ms=-1.0):
        out = np.empty(self._sizes.get(rid, 0), np.int32)
        info = RequestInfo()
        _check(lib().snicit_server_wait(self.h, rid, timeout_ms, _i32(out), ctypes.byref(info)))
        self._sizes.pop(rid, None)
        return out, info.as_dict()

    def stats(self):
        st = ServerStats()
        _check(lib().snicit_server_stats_get(self.h, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if self.h:
            lib().snicit_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Method:
    """Shared body of the mode classes: a resident net and ``infer`` in rounds of
    ``batch`` (SNICIT::infer, src/SDGC/SNICIT/SNICIT.hpp:121 rounds of batch)."""

    mode = "BF"

    def __init__(self, net, threshold=30, **cfg):
        self.net = net
        self.cfg = default_cfg(self.mode, threshold, **cfg)
        self.stats = None

    def infer(self, inputs, batch=None, noalloc=False):
        """Categories of every input, in rounds of `batch`.  noalloc=True runs
        snicit_infer_noalloc: every device buffer from the library's static scratch."""
        inputs = np.ascontiguousarray(inputs, dtype=np.float32)
        B = inputs.shape[0]
        batch = B if batch is None else int(batch)
        if B % batch:
            raise ValueError("batch must divide the number of inputs")
        cat = np.empty(B, dtype=np.int32)
        st = Stats()
        fn = lib().snicit_infer_noalloc if noalloc else lib().snicit_infer
        _check(fn(self.net.h, ctypes.byref(self.cfg), _f32(inputs), B, batch, _i32(cat),
                  ctypes.byref(st)))
        self.stats = st
        return cat


class SNICIT(_Method):
    """Exact SNICIT (pre-threshold SpMM, clustering at ``threshold``, delta SpMM, recovery)."""
    mode = "SNICIT"


class BF(_Method):
    """Plain layer-by-layer SpMM with all-zero rows skipped (BF-2019)."""
    mode = "BF"


class SNIG(_Method):
    """Plain SpMM under the SNIG mode name (same results as BF)."""
    mode = "SNIG"


class XY(_Method):
    """Plain SpMM under the XY mode name (same results as BF)."""
    mode = "XY"
