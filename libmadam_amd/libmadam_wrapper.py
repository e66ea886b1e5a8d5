"""Drop-in for the reference's Python wrapper (python/libmadam_wrapper/__init__.py:1-181).

Same functions, arguments, dtypes and in-place semantics; ``comm`` may be an
mpi4py communicator (``comm.py2f()`` is passed like the reference does) or
None.  Errors end the process like the reference's MPI_Abort unless the
environment variable MADAM_RAISE is set, in which case a MadamError is raised.
"""
import ctypes as ct

import numpy as np

from . import (MadamError, PIXEL_TYPE, PSD_TYPE, SIGNAL_TYPE, TIMESTAMP_TYPE, WEIGHT_TYPE, dets2detstring,
               dict2parstring, lib)

REPEATED_KEYS = ["detset", "detset_nopol", "survey"]
available = True


def _fcomm(comm):
    if comm is None:
        return 0
    return comm.py2f() if hasattr(comm, "py2f") else int(comm)


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


def _raise_if_failed():
    L = lib()
    if L.madam_last_status() != 0:
        raise MadamError(L.madam_last_error().decode())


def destripe(comm, pars, dets, weights, timestamps, pixels, pixweights, signal, periods, npsd, psdstarts,
             psdfreqs, psdvals):
    """Solve and subtract baseline offsets from signal (reference __init__.py:97-150)."""
    L = lib()
    parstring = dict2parstring(pars)
    ndet = len(dets)
    detstring = dets2detstring(dets)
    nsamp = timestamps.size
    nperiod = periods.size
    nnz = pixweights.size // nsamp // ndet
    npsdbin = psdfreqs.size
    npsdval = psdvals.size
    npsdtot = npsdval // npsdbin
    w = np.ascontiguousarray(weights, np.float64)
    ts = np.ascontiguousarray(timestamps.reshape(-1), dtype=TIMESTAMP_TYPE)
    pix = np.ascontiguousarray(pixels.reshape(-1), dtype=PIXEL_TYPE)
    pw = np.ascontiguousarray(pixweights.reshape(-1), dtype=WEIGHT_TYPE)
    sig = np.ascontiguousarray(signal.reshape(-1), dtype=SIGNAL_TYPE)
    per = np.ascontiguousarray(periods, np.int64)
    nps = np.ascontiguousarray(npsd, np.int64)
    pst = np.ascontiguousarray(psdstarts, np.float64)
    pfr = np.ascontiguousarray(psdfreqs, np.float64)
    pv = np.ascontiguousarray(psdvals.reshape(-1), dtype=PSD_TYPE)
    L.destripe(ct.c_int(_fcomm(comm)), ct.c_char_p(parstring), ct.c_long(ndet), ct.c_char_p(detstring), _p(w),
               ct.c_long(nsamp), ct.c_long(nnz), _p(ts), _p(pix), _p(pw), _p(sig), ct.c_long(nperiod), _p(per),
               _p(nps), ct.c_long(npsdtot), _p(pst), ct.c_long(npsdbin), _p(pfr), ct.c_long(npsdval), _p(pv))
    _raise_if_failed()


def destripe_with_cache(comm, timestamps, pixels, pixweights, signal, outpath):
    """Solve with the cached configuration (reference __init__.py:153-174)."""
    L = lib()
    nsamp = timestamps.size
    ndet = signal.size // nsamp
    nnz = pixweights.size // nsamp // ndet
    ts = np.ascontiguousarray(timestamps.reshape(-1), dtype=TIMESTAMP_TYPE)
    pix = np.ascontiguousarray(pixels.reshape(-1), dtype=PIXEL_TYPE)
    pw = np.ascontiguousarray(pixweights.reshape(-1), dtype=WEIGHT_TYPE)
    sig = np.ascontiguousarray(signal.reshape(-1), dtype=SIGNAL_TYPE)
    if isinstance(outpath, str):
        outpath = outpath.encode("ascii")
    L.destripe_with_cache(ct.c_int(_fcomm(comm)), ct.c_long(ndet), ct.c_long(nsamp), ct.c_long(nnz), _p(ts),
                          _p(pix), _p(pw), _p(sig), ct.c_char_p(outpath))
    _raise_if_failed()


def clear_caches():
    """Clear Monte Carlo caches after a destripe() call with mcmode=True."""
    lib().clear_caches()
