"""ctypes wrapper of the CPU oracle (oracle.c) and the FP32 CPU port (cpu_port.c).

TEST AND MEASUREMENT INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg, always as the checker or the timed CPU baseline.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

_f32 = ctypes.POINTER(ctypes.c_float)
_f64 = ctypes.POINTER(ctypes.c_double)
_i32 = ctypes.POINTER(ctypes.c_int32)
_u32 = ctypes.POINTER(ctypes.c_uint32)
_i64 = ctypes.POINTER(ctypes.c_int64)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_lib = {}


def _load(name):
    if name not in _lib:
        p = os.path.join(BUILD, name)
        if not os.path.exists(p):
            import subprocess
            subprocess.run(["make", "-C", HERE, "-s", "all"], check=True)
        _lib[name] = ctypes.CDLL(p)
    return _lib[name]


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def oracle():
    lib = _load("liboracle.so")
    lib.orc_dist2.argtypes = [_f32, _f32, ctypes.c_int]
    lib.orc_dist2.restype = ctypes.c_double
    lib.orc_rho_from_d2.argtypes = [ctypes.c_double]
    lib.orc_rho_from_d2.restype = ctypes.c_float
    lib.orc_sentinel_delta.argtypes = []
    lib.orc_sentinel_delta.restype = ctypes.c_float
    lib.orc_sort_batch.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   _f32, _f32, _i32, _i32, _i64, _u8, _f32, _f32, _i64, _i64, _f64, ctypes.c_int]
    lib.orc_knn_site.argtypes = [ctypes.c_int, ctypes.c_int64, _f32, _i32, _i64, ctypes.c_int, _u8,
                                 ctypes.c_int64, ctypes.c_int64, ctypes.c_int, _f32, _i32]
    lib.orc_delta_site.argtypes = [ctypes.c_int, ctypes.c_int64, _f32, _i32, _i64, ctypes.c_int, _u8,
                                   ctypes.c_int64, ctypes.c_int64, _f32, _f32, _i32]
    lib.orc_knn_ab.argtypes = [ctypes.c_int, ctypes.c_int64, _f32, ctypes.c_int64, _f32, ctypes.c_int, _f32, _i32]
    lib.orc_knn_ab_list.argtypes = [ctypes.c_int, ctypes.c_int64, _f32, ctypes.c_int64, _i32, ctypes.c_int64,
                                    _f32, ctypes.c_int, _f32, _f32, _i32]
    lib.orc_delta_knn_ab.argtypes = [ctypes.c_int, ctypes.c_int64, _f32, _f32, ctypes.c_int64, _f32, _f32,
                                     _f32, _i32]
    lib.orc_rho_drift_dc.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, _i32, _f32, _f32,
                                     ctypes.c_float, _f32]
    lib.orc_delta_drift_dc.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, _i32, _f32, _f32, _f32,
                                       ctypes.c_float, _f32, _u32]
    return lib


def _np(x, dt):
    if x is None:
        return None
    if hasattr(x, "cpu"):
        x = x.cpu().numpy()
    return np.ascontiguousarray(x, dtype=dt)


def sort_batch(ds, nthreads=8, want_knn=True):
    """Oracle over every site and chunk: dict(rho, delta, nneigh, knn (N, k) 1-based, dk2)."""
    lib = oracle()
    N, D, S, C, k = ds["n_spikes"], ds["D"], ds["n_sites"], ds["n_drift"], ds["k"]
    F1 = _np(ds["F1"], np.float32)
    F2 = _np(ds.get("F2"), np.float32)
    s1 = _np(ds["site1"], np.int32)
    s2 = _np(ds.get("site2"), np.int32)
    lim = _np(ds["lim"], np.int64)
    link = np.ascontiguousarray(np.asarray(ds["link"], dtype=np.uint8).reshape(-1, order="F"))
    rho = np.zeros(N, np.float32)
    delta = np.zeros(N, np.float32)
    nn = np.zeros(N, np.int64)
    knn = np.zeros((N, k), np.int64) if want_knn else None
    dk2 = np.zeros(N, np.float64)
    st = lib.orc_sort_batch(N, D, S, C, k, _p(F1, _f32), _p(F2, _f32), _p(s1, _i32), _p(s2, _i32), _p(lim, _i64),
                            _p(link, _u8), _p(rho, _f32), _p(delta, _f32), _p(nn, _i64), _p(knn, _i64),
                            _p(dk2, _f64), nthreads)
    assert st == 0
    return dict(rho=rho, delta=delta, nneigh=nn, knn=knn, dk2=dk2)


def knn_site(F, drift_of_col, link, in_begin, n_in, k, id_of_col=None):
    lib = oracle()
    F = np.ascontiguousarray(F, np.float32)
    n, D = F.shape
    dr = np.ascontiguousarray(drift_of_col, np.int32)
    ids = None if id_of_col is None else np.ascontiguousarray(id_of_col, np.int64)
    lk = np.ascontiguousarray(np.asarray(link, np.uint8).reshape(-1, order="F"))
    dk = np.zeros(n_in, np.float32)
    I = np.zeros((n_in, k), np.int32)
    assert lib.orc_knn_site(D, n, _p(F, _f32), _p(dr, _i32), _p(ids, _i64), link.shape[0], _p(lk, _u8),
                            in_begin, n_in, k, _p(dk, _f32), _p(I, _i32)) == 0
    return dk, I


def delta_site(F, drift_of_col, link, in_begin, n_in, rho_cols, id_of_col=None):
    lib = oracle()
    F = np.ascontiguousarray(F, np.float32)
    n, D = F.shape
    dr = np.ascontiguousarray(drift_of_col, np.int32)
    ids = None if id_of_col is None else np.ascontiguousarray(id_of_col, np.int64)
    lk = np.ascontiguousarray(np.asarray(link, np.uint8).reshape(-1, order="F"))
    r = np.ascontiguousarray(rho_cols, np.float32)
    dm = np.zeros(n_in, np.float32)
    nn = np.zeros(n_in, np.int32)
    assert lib.orc_delta_site(D, n, _p(F, _f32), _p(dr, _i32), _p(ids, _i64), link.shape[0], _p(lk, _u8),
                              in_begin, n_in, _p(r, _f32), _p(dm, _f32), _p(nn, _i32)) == 0
    return dm, nn


def knn_ab(B, A, k):
    lib = oracle()
    B = np.ascontiguousarray(B, np.float32)
    A = np.ascontiguousarray(A, np.float32)
    K = np.zeros(A.shape[0], np.float32)
    I = np.zeros((A.shape[0], k), np.int32)
    assert lib.orc_knn_ab(B.shape[1], B.shape[0], _p(B, _f32), A.shape[0], _p(A, _f32), k, _p(K, _f32),
                          _p(I, _i32)) == 0
    return K, I


def knn_ab_list(B, A, k, IB=None):
    """cuda_knn_full / cuda_knn_ib, exact: (K_all (nA, k), K (nA,), I (nA, k)); IB 1-based columns of B."""
    lib = oracle()
    B = np.ascontiguousarray(B, np.float32)
    A = np.ascontiguousarray(A, np.float32)
    ib = None if IB is None else np.ascontiguousarray(IB, np.int32)
    Ka = np.zeros((A.shape[0], k), np.float32)
    K = np.zeros(A.shape[0], np.float32)
    I = np.zeros((A.shape[0], k), np.int32)
    assert lib.orc_knn_ab_list(B.shape[1], B.shape[0], _p(B, _f32), 0 if ib is None else len(ib), _p(ib, _i32),
                               A.shape[0], _p(A, _f32), k, _p(Ka, _f32), _p(K, _f32), _p(I, _i32)) == 0
    return Ka, K, I


def delta_knn_ab(B, RB, A, RA):
    lib = oracle()
    B = np.ascontiguousarray(B, np.float32)
    A = np.ascontiguousarray(A, np.float32)
    RB = np.ascontiguousarray(RB, np.float32)
    RA = np.ascontiguousarray(RA, np.float32)
    Dd = np.zeros(A.shape[0], np.float32)
    Nn = np.zeros(A.shape[0], np.int32)
    assert lib.orc_delta_knn_ab(B.shape[1], B.shape[0], _p(B, _f32), _p(RB, _f32), A.shape[0], _p(A, _f32),
                                _p(RA, _f32), _p(Dd, _f32), _p(Nn, _i32)) == 0
    return Dd, Nn


def rho_drift_dc(n1, miSite12, fet1, fet2, dc2):
    lib = oracle()
    ms = np.ascontiguousarray(np.asarray(miSite12, np.int32).reshape(-1, order="F"))
    f1 = np.ascontiguousarray(fet1, np.float32)
    f2 = np.ascontiguousarray(fet2, np.float32)
    rho = np.zeros(n1, np.float32)
    assert lib.orc_rho_drift_dc(f1.shape[1], n1, f1.shape[0], _p(ms, _i32), _p(f1, _f32), _p(f2, _f32),
                                float(dc2), _p(rho, _f32)) == 0
    return rho


def delta_drift_dc(n1, miSite12, fet1, fet2, rho12, dc2):
    lib = oracle()
    ms = np.ascontiguousarray(np.asarray(miSite12, np.int32).reshape(-1, order="F"))
    f1 = np.ascontiguousarray(fet1, np.float32)
    f2 = np.ascontiguousarray(fet2, np.float32)
    r = np.ascontiguousarray(rho12, np.float32)
    dl = np.zeros(n1, np.float32)
    nn = np.zeros(n1, np.uint32)
    assert lib.orc_delta_drift_dc(f1.shape[1], n1, f1.shape[0], _p(ms, _i32), _p(f1, _f32), _p(f2, _f32),
                                  _p(r, _f32), float(dc2), _p(dl, _f32), _p(nn, _u32)) == 0
    return dl, nn


def dist2(q, x):
    lib = oracle()
    q = np.ascontiguousarray(q, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    return lib.orc_dist2(_p(q, _f32), _p(x, _f32), q.shape[0])


def cpu_port():
    lib = _load("libcpuport.so")
    lib.cpu_sort_batch.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   _f32, _f32, _i32, _i32, _i64, _u8, _i32, ctypes.c_int, _f32, _f32, _i64,
                                   _i64, ctypes.c_int]
    lib.cpu_max_threads.restype = ctypes.c_int
    return lib


def cpu_sort_batch(ds, sites=None, nthreads=0):
    """FP32 restatement of the reference CPU path (knn_cpu_ + delta_cpu_), optionally on a
    subset of sites (1-based).  Returns dict(rho, delta, nneigh, knn) (entries of other sites are 0;
    knn is N x k global 1-based ids in ascending FP32 distance)."""
    lib = cpu_port()
    N, D, S, C, k = ds["n_spikes"], ds["D"], ds["n_sites"], ds["n_drift"], ds["k"]
    F1 = _np(ds["F1"], np.float32)
    F2 = _np(ds.get("F2"), np.float32)
    s1 = _np(ds["site1"], np.int32)
    s2 = _np(ds.get("site2"), np.int32)
    lim = _np(ds["lim"], np.int64)
    link = np.ascontiguousarray(np.asarray(ds["link"], dtype=np.uint8).reshape(-1, order="F"))
    sl = None if sites is None else np.ascontiguousarray(sites, np.int32)
    rho = np.zeros(N, np.float32)
    delta = np.zeros(N, np.float32)
    nn = np.zeros(N, np.int64)
    knn = np.zeros((N, k), np.int64)
    st = lib.cpu_sort_batch(N, D, S, C, k, _p(F1, _f32), _p(F2, _f32), _p(s1, _i32), _p(s2, _i32), _p(lim, _i64),
                            _p(link, _u8), _p(sl, _i32), 0 if sl is None else len(sl), _p(rho, _f32),
                            _p(delta, _f32), _p(nn, _i64), _p(knn, _i64), nthreads)
    assert st == 0
    return dict(rho=rho, delta=delta, nneigh=nn, knn=knn)
