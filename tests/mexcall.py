"""Drive the MEX gateway (ironclust_amd/csrc/mex/irc_hip_mex.c) without MATLAB: it is compiled with
gcc against the mock MEX API in tests/mex_mock and linked to libirc_hip.so; mex(cmd, *args) passes
numpy arrays as mxArrays (MATLAB orientation: a 2-D array is m x n, column-major) and returns the
outputs or raises MexError with the identifier and text mexErrMsgIdAndTxt was given."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOCK = os.path.join(ROOT, "tests", "mex_mock")
BUILD = os.path.join(MOCK, "_build")
LIBDIR = os.path.join(ROOT, "ironclust_amd", "_lib")
_lib = None

CLS = {np.dtype(np.bool_): 0, np.dtype(np.float64): 2, np.dtype(np.float32): 3, np.dtype(np.int32): 4,
       np.dtype(np.int64): 5}
NP = {0: np.bool_, 2: np.float64, 3: np.float32, 4: np.int32, 5: np.int64}


class MexError(RuntimeError):
    def __init__(self, ident, text):
        super().__init__(f"{ident}: {text}")
        self.ident = ident


def load():
    global _lib
    if _lib is None:
        os.makedirs(BUILD, exist_ok=True)
        so = os.path.join(BUILD, "libirc_mex_mock.so")
        cmd = ["gcc", "-O1", "-fPIC", "-shared", "-std=c11", "-I", MOCK, "-I", os.path.join(ROOT, "include"),
               os.path.join(ROOT, "ironclust_amd", "csrc", "mex", "irc_hip_mex.c"), os.path.join(MOCK, "mex_mock.c"),
               "-L", LIBDIR, "-lirc_hip", f"-Wl,-rpath,{LIBDIR}", "-o", so]
        subprocess.run(cmd, check=True)
        lib = ctypes.CDLL(so)
        lib.mock_call.restype = ctypes.c_int
        _lib = lib
    return _lib


def mex(cmd, *args, nlhs=2):
    lib = load()
    items = [("char", cmd)] + list(args)
    n = len(items)
    cls = (ctypes.c_int * n)()
    m = (ctypes.c_int64 * n)()
    nn = (ctypes.c_int64 * n)()
    data = (ctypes.c_void_p * n)()
    strs = (ctypes.c_char_p * n)()
    keep = []
    for i, a in enumerate(items):
        if isinstance(a, tuple) and a[0] == "char":
            cls[i], m[i], nn[i] = 1, 1, len(a[1])
            strs[i] = a[1].encode()
            continue
        if isinstance(a, str):
            cls[i], m[i], nn[i] = 1, 1, len(a)
            strs[i] = a.encode()
            continue
        x = np.asarray(a)
        if x.ndim == 0:
            x = x.reshape(1, 1)
        elif x.ndim == 1:
            x = x.reshape(-1, 1)
        x = np.asfortranarray(x)
        keep.append(x)
        cls[i], m[i], nn[i] = CLS[x.dtype], x.shape[0], x.shape[1]
        data[i] = x.ctypes.data
    od = (ctypes.c_void_p * nlhs)()
    om = (ctypes.c_int64 * nlhs)()
    on = (ctypes.c_int64 * nlhs)()
    oc = (ctypes.c_int * nlhs)()
    err = ctypes.create_string_buffer(512)
    rc = lib.mock_call(n, cls, m, nn, data, strs, nlhs, od, om, on, oc, err, 512)
    if rc:
        ident, _, text = err.value.decode().partition("|")
        raise MexError(ident, text)
    outs = []
    for j in range(nlhs):
        if not od[j]:
            outs.append(None)
            continue
        dt = np.dtype(NP[oc[j]])
        cnt = om[j] * on[j]
        buf = (ctypes.c_char * (cnt * dt.itemsize)).from_address(od[j])
        outs.append(np.frombuffer(bytes(buf), dtype=dt).reshape((om[j], on[j]), order="F").copy())
    return outs
