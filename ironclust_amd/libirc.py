"""ctypes binding of libirc_hip.so (the C ABI in include/irc_hip.h).

This module only marshals arguments; every computation runs in the gfx950 library.
Loading fails loudly when the library is missing: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "_lib", "libirc_hip.so")

IRC_OK, IRC_E_ARG, IRC_E_NOMEM, IRC_E_HIP, IRC_E_UNSUPPORTED, IRC_E_FORCED = range(6)
PREC_EXACT, PREC_FP32, PREC_BF16_SCREEN, PREC_AUTO, PREC_FP8_SCREEN = range(5)
MEM_HOST, MEM_DEVICE = 0, 1
FLAG_GRAPH, FLAG_AUTOTUNE, FLAG_SITE_LAUNCHES = 4, 8, 16          # irc_options.flags

EXPORTED = [
    "irc_options_default", "irc_ctx_create", "irc_ctx_destroy", "irc_last_error", "irc_version",
    "irc_knn_drift", "irc_delta_drift", "irc_rho_delta_batch", "irc_get_stats",
    "irc_knn_ab", "irc_knn_ab_full", "irc_knn_ab_ib", "irc_delta_knn_ab", "irc_rho_drift_dc", "irc_delta_drift_dc", "irc_selftest_mfma", "irc_selftest_mfma_accum",
    "irc_calc_drift", "irc_assign_chain", "irc_isolation_counts", "irc_autotune", "irc_rho_delta_batch_noalloc",
]

_c_f32p = ctypes.POINTER(ctypes.c_float)
_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_u32p = ctypes.POINTER(ctypes.c_uint32)
_c_i64p = ctypes.POINTER(ctypes.c_int64)
_c_u8p = ctypes.POINTER(ctypes.c_uint8)


class Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("k_over", ctypes.c_int32), ("tile", ctypes.c_int32),
                ("force_fail", ctypes.c_int32), ("mem", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("waves_per_cu", ctypes.c_int32), ("large_ncand", ctypes.c_int32)]


class SiteProblem(ctypes.Structure):
    _fields_ = [("D", ctypes.c_int32), ("k", ctypes.c_int32), ("n_cols", ctypes.c_int64),
                ("F", _c_f32p), ("drift_of_col", _c_i32p), ("id_of_col", _c_i64p),
                ("n_drift", ctypes.c_int32), ("link", _c_u8p),
                ("in_begin", ctypes.c_int64), ("n_in", ctypes.c_int64)]


class Batch(ctypes.Structure):
    _fields_ = [("n_spikes", ctypes.c_int64), ("D", ctypes.c_int32), ("k", ctypes.c_int32),
                ("n_sites", ctypes.c_int32), ("n_drift", ctypes.c_int32), ("n_cols", ctypes.c_int64),
                ("F", _c_f32p), ("col_spike", _c_i32p), ("site_ptr", _c_i64p),
                ("site_n_primary", _c_i64p), ("lim_drift", _c_i64p), ("link", _c_u8p)]


class DriftProblem(ctypes.Structure):
    _fields_ = [("n_spikes", ctypes.c_int64), ("amp", _c_f32p), ("pos", _c_f32p), ("n_amp", ctypes.c_int32),
                ("n_pos", ctypes.c_int32), ("n_drift", ctypes.c_int32), ("n_batch", ctypes.c_int32),
                ("n_time_max", ctypes.c_int32)]


class Outputs(ctypes.Structure):
    _fields_ = [("rho", _c_f32p), ("delta", _c_f32p), ("nneigh", _c_i64p), ("knn", _c_i64p), ("dk", _c_f32p)]


class Stats(ctypes.Structure):
    _fields_ = [("n_queries", ctypes.c_int64), ("n_pairs", ctypes.c_int64),
                ("n_screen_fallback", ctypes.c_int64), ("n_delta_full", ctypes.c_int64),
                ("ms_total", ctypes.c_double), ("ms_knn", ctypes.c_double), ("ms_delta", ctypes.c_double),
                ("ms_screen", ctypes.c_double), ("precision_used", ctypes.c_int32),
                ("ms_rerank", ctypes.c_double), ("n_rerank_overflow", ctypes.c_int64),
                ("ms_graph", ctypes.c_double), ("n_rerank_second", ctypes.c_int64)]


class TuneReport(ctypes.Structure):
    _fields_ = [("n_trials", ctypes.c_int32), ("best_trial", ctypes.c_int32), ("ms_base", ctypes.c_double),
                ("ms_best", ctypes.c_double), ("k_over", ctypes.c_int32 * 16), ("waves_per_cu", ctypes.c_int32 * 16),
                ("tile", ctypes.c_int32 * 16), ("large_ncand", ctypes.c_int32 * 16), ("ms", ctypes.c_double * 16),
                ("identical", ctypes.c_int32 * 16)]


class IrcError(RuntimeError):
    """A non-OK irc_status; `status` holds the code (the MEX maps it to irc:hip errors)."""

    def __init__(self, status, msg):
        super().__init__(f"irc_hip status {status}: {msg}")
        self.status = status


_lib = None


def load(path=None):
    """Load libirc_hip.so; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with ironclust_amd.build.build_lib()")
    lib = ctypes.CDLL(p)
    lib.irc_options_default.argtypes = [ctypes.POINTER(Options)]
    lib.irc_options_default.restype = None
    lib.irc_ctx_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]
    lib.irc_ctx_create.restype = ctypes.c_int
    lib.irc_ctx_destroy.argtypes = [ctypes.c_void_p]
    lib.irc_ctx_destroy.restype = None
    lib.irc_last_error.argtypes = [ctypes.c_void_p]
    lib.irc_last_error.restype = ctypes.c_char_p
    lib.irc_version.argtypes = []
    lib.irc_version.restype = ctypes.c_char_p
    lib.irc_knn_drift.argtypes = [ctypes.c_void_p, ctypes.POINTER(SiteProblem), ctypes.POINTER(Options),
                                  _c_f32p, _c_i32p]
    lib.irc_knn_drift.restype = ctypes.c_int
    lib.irc_delta_drift.argtypes = [ctypes.c_void_p, ctypes.POINTER(SiteProblem), ctypes.POINTER(Options),
                                    _c_f32p, _c_f32p, _c_i32p]
    lib.irc_delta_drift.restype = ctypes.c_int
    lib.irc_rho_delta_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(Batch), ctypes.POINTER(Options),
                                        ctypes.POINTER(Outputs)]
    lib.irc_rho_delta_batch.restype = ctypes.c_int
    lib.irc_rho_delta_batch_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(Batch), ctypes.POINTER(Options),
                                                ctypes.POINTER(Outputs)]
    lib.irc_rho_delta_batch_noalloc.restype = ctypes.c_int
    lib.irc_autotune.argtypes = [ctypes.c_void_p, ctypes.POINTER(Batch), ctypes.POINTER(Options),
                                 ctypes.POINTER(Outputs), ctypes.POINTER(Options), ctypes.POINTER(TuneReport)]
    lib.irc_autotune.restype = ctypes.c_int
    lib.irc_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stats)]
    lib.irc_get_stats.restype = ctypes.c_int
    lib.irc_selftest_mfma.argtypes = [ctypes.c_void_p, _c_f32p]
    lib.irc_selftest_mfma.restype = ctypes.c_int
    lib.irc_selftest_mfma_accum.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_f32p, _c_f32p,
                                            _c_f32p]
    lib.irc_selftest_mfma_accum.restype = ctypes.c_int
    lib.irc_knn_ab.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, _c_f32p, ctypes.c_int64, _c_f32p,
                               ctypes.c_int32, _c_f32p, _c_i32p]
    lib.irc_knn_ab.restype = ctypes.c_int
    lib.irc_knn_ab_full.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, _c_f32p, ctypes.c_int64, _c_f32p,
                                    ctypes.c_int32, _c_f32p, _c_i32p]
    lib.irc_knn_ab_full.restype = ctypes.c_int
    lib.irc_knn_ab_ib.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, _c_f32p, ctypes.c_int64, _c_i32p,
                                  ctypes.c_int64, _c_f32p, ctypes.c_int32, _c_f32p, _c_i32p]
    lib.irc_knn_ab_ib.restype = ctypes.c_int
    lib.irc_delta_knn_ab.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, _c_f32p, _c_f32p,
                                     ctypes.c_int64, _c_f32p, _c_f32p, _c_f32p, _c_i32p]
    lib.irc_delta_knn_ab.restype = ctypes.c_int
    lib.irc_rho_drift_dc.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, _c_i32p,
                                     _c_f32p, _c_f32p, ctypes.c_float, _c_f32p]
    lib.irc_rho_drift_dc.restype = ctypes.c_int
    lib.irc_delta_drift_dc.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, _c_i32p,
                                       _c_f32p, _c_f32p, _c_f32p, ctypes.c_float, _c_f32p, _c_u32p]
    lib.irc_delta_drift_dc.restype = ctypes.c_int
    lib.irc_calc_drift.argtypes = [ctypes.c_void_p, ctypes.POINTER(DriftProblem), _c_i32p, _c_i64p,
                                   ctypes.POINTER(ctypes.c_uint8), _c_i32p]
    lib.irc_calc_drift.restype = ctypes.c_int
    lib.irc_assign_chain.argtypes = [ctypes.c_void_p, ctypes.c_int64, _c_i64p, _c_i32p, ctypes.POINTER(Options)]
    lib.irc_assign_chain.restype = ctypes.c_int
    lib.irc_isolation_counts.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, _c_i64p, _c_i32p,
                                         ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(Options)]
    lib.irc_isolation_counts.restype = ctypes.c_int
    _lib = lib
    return lib


def _p(a, ct):
    if a is None:
        return None
    return a.ctypes.data_as(ct)


def _dptr(t, ct):
    """ctypes pointer of a device tensor (torch) or None."""
    if t is None:
        return None
    return ctypes.cast(ctypes.c_void_p(t.data_ptr()), ct)


def make_options(precision=PREC_AUTO, k_over=0, force_fail=False, mem=MEM_HOST, flags=0, tile=0,
                 waves_per_cu=0, large_ncand=0):
    lib = load()
    o = Options()
    lib.irc_options_default(ctypes.byref(o))
    o.precision = precision
    o.k_over = k_over
    o.force_fail = 1 if force_fail else 0
    o.mem = mem
    o.flags = flags
    o.tile = tile
    o.waves_per_cu = waves_per_cu
    o.large_ncand = large_ncand
    return o


class Context:
    """One irc_ctx (device buffers, stream) -- the role of MATLAB's `persistent CK`."""

    def __init__(self, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        st = self.lib.irc_ctx_create(ctypes.byref(h), int(device))
        if st != IRC_OK:
            raise IrcError(st, f"irc_ctx_create(device={device}) failed")
        self.h = h

    def close(self):
        if self.h:
            self.lib.irc_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != IRC_OK:
            raise IrcError(st, self.lib.irc_last_error(self.h).decode(errors="replace"))

    def stats(self):
        s = Stats()
        self._check(self.lib.irc_get_stats(self.h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def selftest_mfma(self):
        out = np.zeros(6144, np.float32)
        self._check(self.lib.irc_selftest_mfma(self.h, _p(out, _c_f32p)))
        return out

    def selftest_mfma_accum(self, prec, A, B):
        """(T, 32, K) x (T, K, 32) float32 tiles through one MFMA chain of the screen's form prec
        (1 FP32, 2 BF16); returns (T, 32, 32)."""
        A = np.ascontiguousarray(A, np.float32)
        B = np.ascontiguousarray(B, np.float32)
        T, _, K = A.shape
        out = np.zeros((T, 32, 32), np.float32)
        self._check(self.lib.irc_selftest_mfma_accum(self.h, prec, K, T, _p(A, _c_f32p), _p(B, _c_f32p),
                                                     _p(out, _c_f32p)))
        return out

    # ---- per-site forms (search_knn_drift_ / search_delta_drift_) --------------
    def _site_problem(self, F, drift_of_col, link, in_begin, n_in, k, id_of_col):
        F = np.ascontiguousarray(F, dtype=np.float32)
        n_cols, D = F.shape
        drift = np.ascontiguousarray(drift_of_col, dtype=np.int32)
        link = np.asfortranarray(np.asarray(link, dtype=np.uint8))
        ids = None if id_of_col is None else np.ascontiguousarray(id_of_col, dtype=np.int64)
        pr = SiteProblem()
        pr.D, pr.k, pr.n_cols = D, int(k), n_cols
        pr.F = _p(F, _c_f32p)
        pr.drift_of_col = _p(drift, _c_i32p)
        pr.id_of_col = _p(ids, _c_i64p)
        pr.n_drift = link.shape[0]
        lk = np.ascontiguousarray(link.reshape(-1, order="F"))
        pr.link = _p(lk, _c_u8p)
        pr.in_begin, pr.n_in = int(in_begin), int(n_in)
        return pr, (F, drift, ids, lk)

    def knn_drift(self, F, drift_of_col, link, in_begin, n_in, k, id_of_col=None, opts=None):
        """F: (n_cols, D) float32 rows = MATLAB D x n_cols columns. Returns (dk[n_in], knn[n_in, k])."""
        pr, keep = self._site_problem(F, drift_of_col, link, in_begin, n_in, k, id_of_col)
        dk = np.zeros(int(n_in), np.float32)
        knn = np.zeros((int(n_in), int(k)), np.int32)
        o = opts or make_options()
        self._check(self.lib.irc_knn_drift(self.h, ctypes.byref(pr), ctypes.byref(o), _p(dk, _c_f32p),
                                           _p(knn, _c_i32p)))
        return dk, knn

    def delta_drift(self, F, drift_of_col, link, in_begin, n_in, rho_cols, id_of_col=None, opts=None):
        pr, keep = self._site_problem(F, drift_of_col, link, in_begin, n_in, 1, id_of_col)
        rho = np.ascontiguousarray(rho_cols, dtype=np.float32)
        dmin = np.zeros(int(n_in), np.float32)
        nn = np.zeros(int(n_in), np.int32)
        o = opts or make_options()
        self._check(self.lib.irc_delta_drift(self.h, ctypes.byref(pr), ctypes.byref(o), _p(rho, _c_f32p),
                                             _p(dmin, _c_f32p), _p(nn, _c_i32p)))
        return dmin, nn

    # ---- batch form (sort_page_) ------------------------------------------------
    def rho_delta_batch(self, layout, k, opts=None, want_knn=True, want_dk=False):
        """layout: dict from ironclust_amd.irc2.site_layout (host numpy arrays)."""
        F = np.ascontiguousarray(layout["F"], dtype=np.float32)
        col_spike = np.ascontiguousarray(layout["col_spike"], dtype=np.int32)
        site_ptr = np.ascontiguousarray(layout["site_ptr"], dtype=np.int64)
        site_np = np.ascontiguousarray(layout["site_n_primary"], dtype=np.int64)
        lim = np.ascontiguousarray(layout["lim_drift"], dtype=np.int64)
        link = np.asarray(layout["link"], dtype=np.uint8)
        lk = np.ascontiguousarray(link.reshape(-1, order="F"))
        N = int(layout["n_spikes"])
        b = Batch()
        b.n_spikes, b.D, b.k = N, F.shape[1], int(k)
        b.n_sites, b.n_drift, b.n_cols = len(site_np), len(lim) - 1, F.shape[0]
        b.F, b.col_spike = _p(F, _c_f32p), _p(col_spike, _c_i32p)
        b.site_ptr, b.site_n_primary = _p(site_ptr, _c_i64p), _p(site_np, _c_i64p)
        b.lim_drift, b.link = _p(lim, _c_i64p), _p(lk, _c_u8p)
        rho = np.zeros(N, np.float32)
        delta = np.zeros(N, np.float32)
        nn = np.zeros(N, np.int64)
        knn = np.zeros((N, int(k)), np.int64) if want_knn else None
        dk = np.zeros(N, np.float32) if want_dk else None
        out = Outputs(_p(rho, _c_f32p), _p(delta, _c_f32p), _p(nn, _c_i64p), _p(knn, _c_i64p), _p(dk, _c_f32p))
        o = opts or make_options()
        o.mem = MEM_HOST
        self._check(self.lib.irc_rho_delta_batch(self.h, ctypes.byref(b), ctypes.byref(o), ctypes.byref(out)))
        res = {"rho": rho, "delta": delta, "nneigh": nn}
        if want_knn:
            res["knn"] = knn
        if want_dk:
            res["dk"] = dk
        return res

    def _device_structs(self, dev, out, opts):
        site_ptr = np.ascontiguousarray(dev["site_ptr"], dtype=np.int64)
        site_np = np.ascontiguousarray(dev["site_n_primary"], dtype=np.int64)
        lim = np.ascontiguousarray(dev["lim_drift"], dtype=np.int64)
        lk = np.ascontiguousarray(np.asarray(dev["link"], dtype=np.uint8).reshape(-1, order="F"))
        F = dev["F"]
        b = Batch()
        b.n_spikes, b.D, b.k = int(dev["n_spikes"]), int(F.shape[1]), int(dev["k"])
        b.n_sites, b.n_drift, b.n_cols = len(site_np), len(lim) - 1, int(F.shape[0])
        b.F, b.col_spike = _dptr(F, _c_f32p), _dptr(dev["col_spike"], _c_i32p)
        b.site_ptr, b.site_n_primary = _p(site_ptr, _c_i64p), _p(site_np, _c_i64p)
        b.lim_drift, b.link = _p(lim, _c_i64p), _p(lk, _c_u8p)
        o = opts or make_options()
        o.mem = MEM_DEVICE
        outs = Outputs(_dptr(out["rho"], _c_f32p), _dptr(out["delta"], _c_f32p), _dptr(out["nneigh"], _c_i64p),
                       _dptr(out.get("knn"), _c_i64p), _dptr(out.get("dk"), _c_f32p))
        return b, o, outs, (site_ptr, site_np, lim, lk)

    def autotune_device(self, dev, out, opts=None):
        """irc_autotune on device-resident data (as rho_delta_batch_device): returns (options dict of
        the chosen k_over / waves_per_cu / tile / large_ncand, report dict)."""
        b, o, outs, keep = self._device_structs(dev, out, opts)
        best = Options()
        rep = TuneReport()
        self._check(self.lib.irc_autotune(self.h, ctypes.byref(b), ctypes.byref(o), ctypes.byref(outs),
                                          ctypes.byref(best), ctypes.byref(rep)))
        n = rep.n_trials
        trials = [{"k_over": rep.k_over[i], "waves_per_cu": rep.waves_per_cu[i], "tile": rep.tile[i],
                   "large_ncand": rep.large_ncand[i], "ms": rep.ms[i], "identical": bool(rep.identical[i])}
                  for i in range(n)]
        choice = {"k_over": best.k_over, "waves_per_cu": best.waves_per_cu, "tile": best.tile,
                  "large_ncand": best.large_ncand}
        return choice, {"ms_base": rep.ms_base, "ms_best": rep.ms_best, "best_trial": rep.best_trial,
                        "trials": trials}

    def rho_delta_batch_noalloc_device(self, dev, out, opts=None):
        """irc_rho_delta_batch_noalloc on device-resident data (arguments as rho_delta_batch_device)."""
        b, o, outs, keep = self._device_structs(dev, out, opts)
        self._check(self.lib.irc_rho_delta_batch_noalloc(self.h, ctypes.byref(b), ctypes.byref(o), ctypes.byref(outs)))

    def rho_delta_batch_device(self, dev, out, opts=None):
        """Device-resident form.  dev: dict of torch tensors on the context's device
        (F [n_cols, D] float32, col_spike int32) plus host arrays site_ptr, site_n_primary,
        lim_drift, link, n_spikes, k.  out: dict of torch device tensors rho, delta, nneigh
        (int64), optional knn (int64 [N, k]: row = spike, the k ids nearest first) and dk."""
        b, o, outs, keep = self._device_structs(dev, out, opts)
        self._check(self.lib.irc_rho_delta_batch(self.h, ctypes.byref(b), ctypes.byref(o), ctypes.byref(outs)))

    # ---- legacy A/B and dc forms ------------------------------------------------------
    def knn_ab(self, B, A, k, want_index=True):
        B = np.ascontiguousarray(B, dtype=np.float32)
        A = np.ascontiguousarray(A, dtype=np.float32)
        K = np.zeros(A.shape[0], np.float32)
        I = np.zeros((A.shape[0], int(k)), np.int32) if want_index else None
        self._check(self.lib.irc_knn_ab(self.h, B.shape[1], B.shape[0], _p(B, _c_f32p), A.shape[0],
                                        _p(A, _c_f32p), int(k), _p(K, _c_f32p), _p(I, _c_i32p)))
        return K, I

    def knn_ab_full(self, B, A, k):
        """cuda_knn_full form: (K (nA, k) every neighbour's distance, I (nA, k) 1-based into B)."""
        B = np.ascontiguousarray(B, dtype=np.float32)
        A = np.ascontiguousarray(A, dtype=np.float32)
        K = np.zeros((A.shape[0], int(k)), np.float32)
        I = np.zeros((A.shape[0], int(k)), np.int32)
        self._check(self.lib.irc_knn_ab_full(self.h, B.shape[1], B.shape[0], _p(B, _c_f32p), A.shape[0],
                                             _p(A, _c_f32p), int(k), _p(K, _c_f32p), _p(I, _c_i32p)))
        return K, I

    def knn_ab_ib(self, B, IB, A, k, want_index=True):
        """cuda_knn_ib form: candidates are the 1-based columns IB of B; (K (nA,), I (nA, k) B columns)."""
        B = np.ascontiguousarray(B, dtype=np.float32)
        A = np.ascontiguousarray(A, dtype=np.float32)
        ib = np.ascontiguousarray(IB, dtype=np.int32)
        K = np.zeros(A.shape[0], np.float32)
        I = np.zeros((A.shape[0], int(k)), np.int32) if want_index else None
        self._check(self.lib.irc_knn_ab_ib(self.h, B.shape[1], B.shape[0], _p(B, _c_f32p), ib.size, _p(ib, _c_i32p),
                                           A.shape[0], _p(A, _c_f32p), int(k), _p(K, _c_f32p), _p(I, _c_i32p)))
        return K, I

    def delta_knn_ab(self, B, RB, A, RA):
        B = np.ascontiguousarray(B, dtype=np.float32)
        A = np.ascontiguousarray(A, dtype=np.float32)
        RB = np.ascontiguousarray(RB, dtype=np.float32)
        RA = np.ascontiguousarray(RA, dtype=np.float32)
        Dd = np.zeros(A.shape[0], np.float32)
        Nn = np.zeros(A.shape[0], np.int32)
        self._check(self.lib.irc_delta_knn_ab(self.h, B.shape[1], B.shape[0], _p(B, _c_f32p), _p(RB, _c_f32p),
                                              A.shape[0], _p(A, _c_f32p), _p(RA, _c_f32p), _p(Dd, _c_f32p),
                                              _p(Nn, _c_i32p)))
        return Dd, Nn

    def rho_drift_dc(self, n1, miSite12, fet1, fet2, dc2):
        m = np.asfortranarray(np.asarray(miSite12, dtype=np.int32))
        ms = np.ascontiguousarray(m.reshape(-1, order="F"))
        f1 = np.ascontiguousarray(fet1, dtype=np.float32)
        f2 = np.ascontiguousarray(fet2, dtype=np.float32)
        rho = np.zeros(int(n1), np.float32)
        self._check(self.lib.irc_rho_drift_dc(self.h, f1.shape[1], int(n1), f1.shape[0], _p(ms, _c_i32p),
                                              _p(f1, _c_f32p), _p(f2, _c_f32p), float(dc2), _p(rho, _c_f32p)))
        return rho

    def calc_drift(self, amp, pos, n_pos, n_amp, n_drift, n_batch, n_time_max=256, want_counts=False):
        """Drift linkage (calc_drift_).  Returns (link (C, C) bool, link[r, c] = mlDrift(r, c);
        lim (C+1,) 1-based; counts ((n_amp*n_pos), C) int32 or None)."""
        amp = np.ascontiguousarray(amp, dtype=np.float32)
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        pr = DriftProblem()
        pr.n_spikes = amp.size
        pr.amp, pr.pos = _p(amp, _c_f32p), _p(pos, _c_f32p)
        pr.n_amp, pr.n_pos, pr.n_drift, pr.n_batch, pr.n_time_max = n_amp, n_pos, n_drift, n_batch, n_time_max
        C = int(n_drift)
        nC = ctypes.c_int32(0)
        lim = np.zeros(C + 1, np.int64)
        link = np.zeros(C * C, np.uint8)
        counts = np.zeros((C, n_amp * n_pos), np.int32) if want_counts else None
        self._check(self.lib.irc_calc_drift(self.h, ctypes.byref(pr), ctypes.byref(nC), _p(lim, _c_i64p),
                                            link.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                            _p(counts, _c_i32p)))
        C = nC.value
        link = link[:C * C].reshape(C, C, order="F").astype(bool)
        lim = lim[:C + 1]
        if counts is not None:
            counts = np.ascontiguousarray(counts[:C].T) if C > 1 else None
        return link, lim, counts

    def assign_chain(self, nneigh, labels):
        """Cluster of every spike from the peaks' clusters along the nneigh chains (the last loop
        of assignCluster_knn_).  nneigh: 1-based; labels: > 0 at peaks, 0 elsewhere.  Host numpy
        arrays, or torch device tensors (int64 nneigh, int32 labels, updated in place)."""
        if hasattr(labels, "data_ptr"):
            o = make_options()
            o.mem = MEM_DEVICE
            self._check(self.lib.irc_assign_chain(self.h, labels.numel(), _dptr(nneigh, _c_i64p),
                                                  _dptr(labels, _c_i32p), ctypes.byref(o)))
            return labels
        nn = np.ascontiguousarray(nneigh, dtype=np.int64)
        lab = np.array(labels, dtype=np.int32, copy=True)
        self._check(self.lib.irc_assign_chain(self.h, lab.size, _p(nn, _c_i64p), _p(lab, _c_i32p), None))
        return lab

    def isolation_counts(self, knn, clu, n_clu):
        """Neighbour-cluster counts (n_clu, n_clu + 1) uint64: row c-1, column d = kNN entries of
        cluster c's spikes that lie in cluster d (0: none).  knn: k x n 1-based spike ids (host)."""
        K = np.asfortranarray(np.asarray(knn, dtype=np.int64))
        k, n = K.shape
        flat = np.ascontiguousarray(K.reshape(-1, order="F"))
        cl = np.ascontiguousarray(clu, dtype=np.int32)
        out = np.zeros((int(n_clu), int(n_clu) + 1), np.uint64)
        self._check(self.lib.irc_isolation_counts(self.h, n, k, _p(flat, _c_i64p), _p(cl, _c_i32p), int(n_clu),
                                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), None))
        return out

    def delta_drift_dc(self, n1, miSite12, fet1, fet2, rho12, dc2):
        m = np.asfortranarray(np.asarray(miSite12, dtype=np.int32))
        ms = np.ascontiguousarray(m.reshape(-1, order="F"))
        f1 = np.ascontiguousarray(fet1, dtype=np.float32)
        f2 = np.ascontiguousarray(fet2, dtype=np.float32)
        r = np.ascontiguousarray(rho12, dtype=np.float32)
        dl = np.zeros(int(n1), np.float32)
        nn = np.zeros(int(n1), np.uint32)
        self._check(self.lib.irc_delta_drift_dc(self.h, f1.shape[1], int(n1), f1.shape[0], _p(ms, _c_i32p),
                                                _p(f1, _c_f32p), _p(f2, _c_f32p), _p(r, _c_f32p), float(dc2),
                                                _p(dl, _c_f32p), _p(nn, _c_u32p)))
        return dl, nn
