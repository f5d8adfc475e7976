"""MKL PARDISO on the host: the reference's own direct solver, called as it calls it.

Test infrastructure and CPU baseline only (the product never imports this).

The reference solves the full-wave system A(w) X = J of reference()
(src/generateStiff.cpp:568-699) with MKL PARDISO: mtype 13 (complex
unsymmetric), phase 13 (analysis, factorization, solve), nrhs = numPorts,
pardisoinit defaults with iparm[3] = 0 (no iterative refinement scheme
switch), iparm[34] = 1 (0-based indices) and iparm[38] = 1
(:638-672).  Its makefile links MKL's ILP64 interface (makefile:18; myint is
long long, src/fdtd.hpp:30-36), so the indices here are 64-bit as well.

MKL is not part of this repository: the runtime library libmkl_rt is
located at run time (GP_MKL_RT, then the usual conda and system paths) and
called through Intel's documented C entry points (pardisoinit, pardiso,
MKL_Set_Interface_Layer, MKL_Set_Threading_Layer, MKL_Set_Num_Threads).
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os
import time

import numpy as np

_MKL_INTERFACE_ILP64 = 1
_MKL_THREADING_INTEL = 0
_MKL_THREADING_SEQUENTIAL = 1

_lib = None
_path = None


def _candidates():
    env = os.environ.get("GP_MKL_RT")
    if env:
        yield env
    for d in ("/opt/conda/lib", "/usr/lib/x86_64-linux-gnu", "/usr/local/lib", "/opt/intel/oneapi/mkl/latest/lib"):
        for name in ("libmkl_rt.so", "libmkl_rt.so.2", "libmkl_rt.so.1"):
            yield os.path.join(d, name)
    found = ctypes.util.find_library("mkl_rt")
    if found:
        yield found


def available() -> bool:
    try:
        load()
        return True
    except OSError:
        return False


def load():
    """The MKL runtime with the ILP64 interface selected (before any other call)."""
    global _lib, _path
    if _lib is not None:
        return _lib
    err = None
    for p in _candidates():
        if not os.path.exists(p) and "/" in p:
            continue
        try:
            lib = C.CDLL(p, mode=C.RTLD_GLOBAL)
        except OSError as e:
            err = e
            continue
        lib.MKL_Set_Interface_Layer.restype = C.c_int
        lib.MKL_Set_Interface_Layer.argtypes = [C.c_int]
        lib.MKL_Set_Threading_Layer.restype = C.c_int
        lib.MKL_Set_Threading_Layer.argtypes = [C.c_int]
        lib.MKL_Set_Num_Threads.restype = None
        lib.MKL_Set_Num_Threads.argtypes = [C.c_int]
        lib.MKL_Get_Version_String.restype = None
        lib.MKL_Get_Version_String.argtypes = [C.c_char_p, C.c_int]
        i64p = C.POINTER(C.c_int64)
        lib.pardisoinit.restype = None
        lib.pardisoinit.argtypes = [C.c_void_p, i64p, i64p]
        lib.pardiso.restype = None
        lib.pardiso.argtypes = [C.c_void_p, i64p, i64p, i64p, i64p, i64p, C.c_void_p, i64p, i64p, i64p, i64p,
                                i64p, i64p, C.c_void_p, C.c_void_p, i64p]
        if lib.MKL_Set_Interface_Layer(_MKL_INTERFACE_ILP64) != _MKL_INTERFACE_ILP64:
            raise OSError("MKL refused the ILP64 interface layer (already initialised LP64?)")
        _lib, _path = lib, p
        return lib
    raise OSError(f"libmkl_rt not found ({err})")


def version() -> str:
    lib = load()
    buf = C.create_string_buffer(256)
    lib.MKL_Get_Version_String(buf, 256)
    return buf.value.decode(errors="replace").strip()


def set_threads(n: int):
    """n > 1: MKL's threaded layer with n threads; n == 1: its sequential layer
    (what the reference links)."""
    lib = load()
    lib.MKL_Set_Threading_Layer(_MKL_THREADING_SEQUENTIAL if n <= 1 else _MKL_THREADING_INTEL)
    lib.MKL_Set_Num_Threads(max(1, int(n)))


class PardisoError(RuntimeError):
    pass


def solve(A, B):
    """X = A^-1 B for a complex CSR A (scipy.sparse) and B [n, nrhs] complex,
    with the reference's PARDISO call (mtype 13, phase 13, then phase -1).
    Returns (X [n, nrhs], seconds of the phase-13 call)."""
    lib = load()
    A = A.tocsr()
    A.sort_indices()
    n = A.shape[0]
    B = np.asarray(B, np.complex128)
    if B.ndim == 1:
        B = B[:, None]
    nrhs = B.shape[1]
    # PARDISO takes the right-hand sides column after column (J[p*n + i] in reference())
    b = np.ascontiguousarray(B.T)
    x = np.zeros_like(b)
    a = np.ascontiguousarray(A.data, np.complex128)
    ia = np.ascontiguousarray(A.indptr, np.int64)
    ja = np.ascontiguousarray(A.indices, np.int64)
    pt = (C.c_void_p * 64)()
    iparm = (C.c_int64 * 64)()
    i64 = C.c_int64
    mtype, maxfct, mnum, msglvl, err = i64(13), i64(1), i64(1), i64(0), i64(0)
    nn, nr, perm = i64(n), i64(nrhs), i64(0)
    lib.pardisoinit(pt, C.byref(mtype), iparm)
    iparm[38] = 1
    iparm[34] = 1
    iparm[3] = 0
    P = lambda arr: arr.ctypes.data_as(C.c_void_p)  # noqa: E731
    ip = lambda arr: arr.ctypes.data_as(C.POINTER(C.c_int64))  # noqa: E731
    phase = i64(13)
    t = time.perf_counter()
    lib.pardiso(pt, C.byref(maxfct), C.byref(mnum), C.byref(mtype), C.byref(phase), C.byref(nn), P(a), ip(ia),
                ip(ja), C.byref(perm), C.byref(nr), iparm, C.byref(msglvl), P(b), P(x), C.byref(err))
    dt = time.perf_counter() - t
    e13 = err.value
    phase = i64(-1)
    dummy = np.zeros(2, np.complex128)
    lib.pardiso(pt, C.byref(maxfct), C.byref(mnum), C.byref(mtype), C.byref(phase), C.byref(nn), P(dummy), ip(ia),
                ip(ja), C.byref(perm), C.byref(nr), iparm, C.byref(msglvl), P(dummy), P(dummy), C.byref(err))
    if e13 != 0:
        raise PardisoError(f"pardiso phase 13 error {e13}")
    return np.ascontiguousarray(x.T), dt


class Factor:
    """A kept PARDISO factorization (phase 12 once, phase 33 per solve,
    phase -1 on close), with the reference's mtype and iparm; .solve(b) makes
    it usable wherever a SciPy LU object is (oracle.refined_solve)."""

    def __init__(self, A):
        self.lib = load()
        A = A.tocsr()
        A.sort_indices()
        self.n = A.shape[0]
        self.a = np.ascontiguousarray(A.data, np.complex128)
        self.ia = np.ascontiguousarray(A.indptr, np.int64)
        self.ja = np.ascontiguousarray(A.indices, np.int64)
        self.pt = (C.c_void_p * 64)()
        self.iparm = (C.c_int64 * 64)()
        self.mtype = C.c_int64(13)
        self.lib.pardisoinit(self.pt, C.byref(self.mtype), self.iparm)
        self.iparm[38] = 1
        self.iparm[34] = 1
        self.iparm[3] = 0
        self.open = False
        self._call(12, np.zeros(2, np.complex128), np.zeros(2, np.complex128), 1)
        self.open = True

    def _call(self, phase, b, x, nrhs):
        i64 = C.c_int64
        err = i64(0)
        ip = lambda arr: arr.ctypes.data_as(C.POINTER(C.c_int64))  # noqa: E731
        self.lib.pardiso(self.pt, C.byref(i64(1)), C.byref(i64(1)), C.byref(self.mtype), C.byref(i64(phase)),
                         C.byref(i64(self.n)), self.a.ctypes.data_as(C.c_void_p), ip(self.ia), ip(self.ja),
                         C.byref(i64(0)), C.byref(i64(nrhs)), self.iparm, C.byref(i64(0)),
                         b.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p), C.byref(err))
        if err.value != 0:
            raise PardisoError(f"pardiso phase {phase} error {err.value}")

    def solve(self, b):
        b = np.ascontiguousarray(b, np.complex128)
        x = np.zeros_like(b)
        self._call(33, b, x, 1)
        return x

    def close(self):
        if self.open:
            d = np.zeros(2, np.complex128)
            self._call(-1, d, d, 1)
            self.open = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
