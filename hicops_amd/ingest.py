"""MS2 text ingest: the reference's MSQuery file reading (msquery.cpp:96-170,
300-505) in libhcx_host.so, giving the raw spectra that Engine.pickpeaks
normalises on the device (MSQuery::pickpeaks, msquery.cpp:509-569)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import native as N
from .workload import RawSpectra


def read_ms2(path: str, scale: int = 100) -> RawSpectra:
    """Every spectrum of an .ms2 file, as MSQuery::readspectrum leaves it:
    peak bins (uint)(m/z * scale), intensities (uint)(I * 1000), the Z line's
    charge and [M+H]+ value and the I RTime value (carried over from the
    previous spectrum when a spectrum has no such line)."""
    lib = N.host()
    h = C.c_void_p()
    N.check(lib.hcx_ms2_read(str(path).encode(), C.c_uint32(scale), C.byref(h)), lib)
    try:
        n, k, mx = C.c_int32(), C.c_uint64(), C.c_uint32()
        N.check(lib.hcx_ms2_info(h, C.byref(n), C.byref(k), C.byref(mx)), lib)
        n, k = n.value, k.value
        r = RawSpectra(np.zeros(k, np.uint32), np.zeros(k, np.uint32), np.zeros(n + 1, np.uint32),
                       np.zeros(n, np.float64), np.zeros(n, np.int32), np.zeros(n, np.float64))
        N.check(lib.hcx_ms2_copy(h, N.ptr(r.mz), N.ptr(r.intensity), N.ptr(r.idx), N.ptr(r.prec_mz),
                                 N.ptr(r.z), N.ptr(r.rtime)), lib)
    finally:
        lib.hcx_ms2_free(h)
    return r
