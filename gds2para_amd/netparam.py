"""Network-parameter output of the solved Z: conversions between Z, Y and S
(Parasitics::convertParam, src/solnoutclass.hpp:1456-1600) and the CITIfile /
Touchstone writers (SolverDataBase::dumpCITI / dumpTouchstone,
src/solnoutclass.hpp:6040-6215), through libgds2para's host entry points.

Matrices are arrays [nf, np, np] complex with P[f, i, j] = sys->x[(f np + i) np
+ j], the layout Parasitics::saveNetworkParam reads; a Z array from
Layout.solve (shape [nf, np*np]) reshapes to it directly."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi

f64p = C.POINTER(C.c_double)


def _interleave(P):
    P = np.ascontiguousarray(P, np.complex128)
    return np.ascontiguousarray(np.stack([P.real.ravel(), P.imag.ravel()], 1).ravel()), P.shape


def _zsrc(zsource, npt):
    if zsource is None:
        return None
    z = np.ascontiguousarray(np.broadcast_to(np.asarray(zsource, np.float64), (npt,)))
    return z


def convert(P, src: str, dst: str, zsource=50.0):
    """P [nf, np, np] complex from parameter src to dst ('Z', 'Y', 'S')."""
    P = np.asarray(P, np.complex128)
    if P.ndim == 2:
        P = P[None]
    nf, npt = P.shape[0], P.shape[1]
    buf, shape = _interleave(P)
    out = np.zeros_like(buf)
    z = _zsrc(zsource, npt)
    capi.check(capi.load().gp_convert_param(npt, nf, src.encode(), dst.encode(),
                                            None if z is None else z.ctypes.data_as(f64p),
                                            buf.ctypes.data_as(f64p), out.ctypes.data_as(f64p)), "gp_convert_param")
    return (out[0::2] + 1j * out[1::2]).reshape(shape)


def write_citi(path, design, freqs, P, param="Z", timestamp=None):
    P = np.asarray(P, np.complex128).reshape(len(freqs), -1)
    npt = int(round(np.sqrt(P.shape[1])))
    buf, _ = _interleave(P)
    fr = np.ascontiguousarray(freqs, np.float64)
    capi.check(capi.load().gp_write_citi(str(path).encode(), design.encode(),
                                         None if timestamp is None else timestamp.encode(), param.encode(), npt,
                                         len(fr), fr.ctypes.data_as(f64p), buf.ctypes.data_as(f64p)), "gp_write_citi")


def write_touchstone(path, design, freqs, P, param="Z", zsource=50.0, timestamp=None):
    P = np.asarray(P, np.complex128).reshape(len(freqs), -1)
    npt = int(round(np.sqrt(P.shape[1])))
    buf, _ = _interleave(P)
    fr = np.ascontiguousarray(freqs, np.float64)
    z = _zsrc(zsource, npt)
    capi.check(capi.load().gp_write_touchstone(str(path).encode(), design.encode(),
                                               None if timestamp is None else timestamp.encode(), param.encode(),
                                               npt, len(fr), fr.ctypes.data_as(f64p), buf.ctypes.data_as(f64p),
                                               z.ctypes.data_as(f64p)), "gp_write_touchstone")


def write_xyce(path, design, port_names, f0, P0, param="Z", save_thresh=1e-5, timestamp=None):
    """SolverDataBase::dumpXyce: SPICE subcircuit from the lowest frequency's
    np x np block P0 (the Z layout of a solve: index resp + np*src) at f0 Hz
    (gp_write_xyce)."""
    P0 = np.asarray(P0, np.complex128).reshape(1, -1)
    npt = int(round(np.sqrt(P0.shape[1])))
    if len(port_names) != npt:
        raise ValueError("one name per port")
    buf, _ = _interleave(P0)
    names = (C.c_char_p * npt)(*[str(s).encode() for s in port_names])
    capi.check(capi.load().gp_write_xyce(str(path).encode(), design.encode(),
                                         None if timestamp is None else timestamp.encode(), param.encode(), npt,
                                         names, float(f0), buf.ctypes.data_as(f64p), float(save_thresh)),
               "gp_write_xyce")


def write_spef(path, design, port_names, f0, P0, param="Z", port_dirs=None, save_thresh=1e-5, timestamp=None):
    """SolverDataBase::dumpSPEF: the same reduction as write_xyce written as an
    IEEE 1481-1998 SPEF net (gp_write_spef); port_dirs is a string of 'I',
    'O' or 'B', one per port (default all 'B')."""
    P0 = np.asarray(P0, np.complex128).reshape(1, -1)
    npt = int(round(np.sqrt(P0.shape[1])))
    if len(port_names) != npt or (port_dirs is not None and len(port_dirs) != npt):
        raise ValueError("one name and one direction per port")
    buf, _ = _interleave(P0)
    names = (C.c_char_p * npt)(*[str(s).encode() for s in port_names])
    capi.check(capi.load().gp_write_spef(str(path).encode(), design.encode(),
                                         None if timestamp is None else timestamp.encode(), param.encode(), npt,
                                         names, None if port_dirs is None else port_dirs.encode(), float(f0),
                                         buf.ctypes.data_as(f64p), float(save_thresh)),
               "gp_write_spef")
