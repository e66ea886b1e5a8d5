"""Oracle (test infrastructure only): ctypes bridge to the FP64 CPU restatement.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this package.  The product (``libmadam_amd``) never does.
"""
import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmadam_oracle.so")
_lib = None


def build(force=False):
    """Compile the oracle library with its Makefile (gcc + OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE] + (["-B"] if force else []))
    return _LIB_PATH


class _Result(ct.Structure):
    _fields_ = [
        ("status", ct.c_int32), ("msg", ct.c_char * 512),
        ("nmap", ct.c_int32), ("ndet", ct.c_int32), ("nside_map", ct.c_int32),
        ("nside_cross", ct.c_int32), ("nthreads", ct.c_int32),
        ("npix_map", ct.c_int64), ("npix_cross", ct.c_int64), ("nsamp", ct.c_int64),
        ("nbase", ct.c_int64),
        ("map", ct.POINTER(ct.c_double)), ("binmap", ct.POINTER(ct.c_double)),
        ("cc", ct.POINTER(ct.c_double)), ("crit", ct.POINTER(ct.c_double)),
        ("hits", ct.POINTER(ct.c_int32)), ("mask", ct.POINTER(ct.c_int32)),
        ("aa", ct.POINTER(ct.c_double)), ("yba", ct.POINTER(ct.c_double)),
        ("nna", ct.POINTER(ct.c_double)),
        ("base_start", ct.POINTER(ct.c_int64)), ("base_len", ct.POINTER(ct.c_int64)),
        ("niter", ct.c_int32), ("rrinit", ct.c_double), ("rzinit", ct.c_double),
        ("rr", ct.POINTER(ct.c_double)),
        ("nof", ct.c_int32), ("nband_max", ct.c_int32), ("nchol", ct.c_int32),
        ("ncgfallback", ct.c_int32), ("precond_kind", ct.c_int32),
        ("t_setup", ct.c_double), ("t_bin", ct.c_double), ("t_init", ct.c_double),
        ("t_precond", ct.c_double), ("t_iter", ct.c_double), ("t_final", ct.c_double),
        ("t_total", ct.c_double),
    ]


class _ScanParams(ct.Structure):
    _fields_ = [("fsample", ct.c_double), ("spin_period", ct.c_double),
                ("prec_period", ct.c_double), ("opening_turns", ct.c_double),
                ("lon0_turns", ct.c_double)]


class _ScanDet(ct.Structure):
    _fields_ = [("psi_turns", ct.c_double), ("open_offset_turns", ct.c_double),
                ("spin_offset_turns", ct.c_double)]


def lib():
    global _lib
    if _lib is None:
        build()
        L = ct.CDLL(_LIB_PATH)
        P = np.ctypeslib.ndpointer
        L.oracle_destripe.restype = ct.POINTER(_Result)
        L.oracle_destripe.argtypes = [
            ct.c_int, ct.c_char_p, ct.c_long, ct.c_char_p,
            P(np.float64, flags="C"), ct.c_long, ct.c_long,
            P(np.float64, flags="C"), P(np.int32, flags="C"), P(np.float32, flags="C"),
            P(np.float64, flags="C"), ct.c_long, P(np.int64, flags="C"), P(np.int64, flags="C"),
            ct.c_long, P(np.float64, flags="C"), ct.c_long, P(np.float64, flags="C"),
            ct.c_long, P(np.float64, flags="C")]
        L.oracle_free.argtypes = [ct.POINTER(_Result)]
        L.oracle_invert_eig.restype = ct.c_int
        L.oracle_invert_eig.argtypes = [P(np.float64, flags="C"), ct.c_int,
                                        ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]
        L.oracle_band_cholesky.restype = ct.c_int
        L.oracle_band_cholesky.argtypes = [P(np.float64, flags="C"), ct.c_long, ct.c_int]
        L.oracle_band_solve.argtypes = [P(np.float64, flags="C"), ct.c_long, ct.c_int,
                                        P(np.float64, flags="C")]
        L.oracle_build_filter.restype = ct.c_int
        L.oracle_build_filter.argtypes = [ct.c_double, ct.c_int, ct.c_long, ct.c_int, ct.c_double,
                                          ct.c_long, P(np.float64, flags="C"),
                                          P(np.float64, flags="C"), P(np.float64, flags="C")]
        L.oracle_interpolate_psd.argtypes = [ct.c_long, P(np.float64, flags="C"),
                                             P(np.float64, flags="C"), ct.c_long,
                                             P(np.float64, flags="C"), P(np.float64, flags="C")]
        L.oracle_convolve.argtypes = [ct.c_long, P(np.float64, flags="C"), P(np.float64, flags="C"), ct.c_long,
                                      P(np.float64, flags="C")]
        L.offt_create.restype = ct.c_void_p
        L.offt_create.argtypes = [ct.c_long]
        L.offt_destroy.argtypes = [ct.c_void_p]
        L.offt_r2c.argtypes = [ct.c_void_p, P(np.float64, flags="C"), P(np.float64, flags="C"),
                               P(np.float64, flags="C")]
        L.offt_c2r.argtypes = [ct.c_void_p, P(np.float64, flags="C"), P(np.float64, flags="C"),
                               P(np.float64, flags="C")]
        L.oscan_sincos_turns.argtypes = [ct.c_double, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]
        L.oscan_atan2_pos.restype = ct.c_double
        L.oscan_atan2_pos.argtypes = [ct.c_double, ct.c_double]
        L.oscan_vec2pix_nest.restype = ct.c_int64
        L.oscan_vec2pix_nest.argtypes = [ct.c_int64, ct.c_double, ct.c_double, ct.c_double]
        L.oscan_expand.argtypes = [ct.POINTER(_ScanParams), ct.POINTER(_ScanDet), ct.c_int64,
                                   ct.c_int64, ct.c_int64, P(np.int32, flags="C"),
                                   P(np.float32, flags="C")]
        _lib = L
    return _lib


def dict2parstring(d):
    """Same encoding as the reference wrapper (python/libmadam_wrapper/__init__.py:73-84)."""
    s = ""
    for key, value in d.items():
        if key in ("detset", "detset_nopol", "survey"):
            for v in value:
                s += "{} = {};".format(key, v)
        else:
            s += "{} = {};".format(key, value)
    return s


class OracleError(RuntimeError):
    pass


def destripe(pars, dets, weights, timestamps, pixels, pixweights, signal, periods,
             npsd, psdstarts, psdfreqs, psdvals, subchunk_pass=0):
    """Run the FP64 CPU restatement.  ``pixels`` and ``signal`` are modified in
    place exactly like the reference modifies the caller's buffers.  Returns a
    dict of numpy arrays (maps are [npix, nmap])."""
    L = lib()
    parstring = (pars if isinstance(pars, str) else dict2parstring(pars)).encode("ascii")
    detstring = "".join("{};".format(d) for d in dets).encode("ascii")
    ndet = len(dets)
    nsamp = timestamps.size
    nnz = pixweights.size // max(nsamp * ndet, 1)
    npsdbin = psdfreqs.size
    npsdval = psdvals.size
    npsdtot = npsdval // max(npsdbin, 1)
    args = [
        int(subchunk_pass), parstring, ndet, detstring, np.ascontiguousarray(weights, np.float64),
        nsamp, nnz, np.ascontiguousarray(timestamps, np.float64), pixels, pixweights, signal,
        periods.size, np.ascontiguousarray(periods, np.int64), np.ascontiguousarray(npsd, np.int64),
        npsdtot, np.ascontiguousarray(psdstarts, np.float64), npsdbin,
        np.ascontiguousarray(psdfreqs, np.float64), npsdval, np.ascontiguousarray(psdvals, np.float64)]
    for a, dt in ((pixels, np.int32), (pixweights, np.float32), (signal, np.float64)):
        if a.dtype != dt or not a.flags.c_contiguous:
            raise TypeError("pixels/pixweights/signal must be C-contiguous int32/float32/float64")
    rp = L.oracle_destripe(*args)
    r = rp.contents
    try:
        if r.status != 0:
            raise OracleError(r.msg.decode())
        npix, nmap, nb = r.npix_map, r.nmap, r.nbase

        def arr(ptr, n, shape=None):
            a = np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n > 0 else np.zeros(0)
            return a.reshape(shape) if shape else a

        out = dict(
            nmap=nmap, nbase=nb, npix=npix, niter=r.niter, rrinit=r.rrinit, rzinit=r.rzinit,
            map=arr(r.map, npix * nmap, (npix, nmap)), binmap=arr(r.binmap, npix * nmap, (npix, nmap)),
            cc=arr(r.cc, npix * nmap * nmap, (npix, nmap, nmap)), crit=arr(r.crit, npix),
            hits=arr(r.hits, npix), mask=arr(r.mask, npix),
            aa=arr(r.aa, ndet * nb, (ndet, nb)), yba=arr(r.yba, ndet * nb, (ndet, nb)),
            nna=arr(r.nna, ndet * nb, (ndet, nb)),
            base_start=arr(r.base_start, nb), base_len=arr(r.base_len, nb),
            rr=arr(r.rr, r.niter), nof=r.nof, nband_max=r.nband_max, nchol=r.nchol,
            ncgfallback=r.ncgfallback, precond_kind=r.precond_kind, nthreads=r.nthreads,
            t_setup=r.t_setup, t_bin=r.t_bin, t_init=r.t_init, t_precond=r.t_precond,
            t_iter=r.t_iter, t_final=r.t_final, t_total=r.t_total)
    finally:
        L.oracle_free(rp)
    return out


# ----------------------------------------------------------------- helpers


def invert_eig(c):
    L = lib()
    c = np.asarray(c, dtype=np.float64).ravel(order="F").copy()
    n = int(round(np.sqrt(c.size)))
    cdet, rc = ct.c_double(), ct.c_double()
    sing = L.oracle_invert_eig(c, n, ct.byref(cdet), ct.byref(rc))
    return c.reshape(n, n, order="F"), cdet.value, rc.value, bool(sing)


def band_cholesky(band):
    L = lib()
    b = np.ascontiguousarray(band, np.float64).copy()
    n, nband = b.shape
    info = L.oracle_band_cholesky(b.reshape(-1), n, nband)
    return b, info


def band_solve(band, rhs):
    L = lib()
    b = np.ascontiguousarray(band, np.float64)
    z = np.ascontiguousarray(rhs, np.float64).copy()
    L.oracle_band_solve(b.reshape(-1), b.shape[0], b.shape[1], z)
    return z


def rfft(x):
    L = lib()
    x = np.ascontiguousarray(x, np.float64)
    n = x.size
    p = L.offt_create(n)
    X = np.zeros(2 * (n // 2 + 1))
    L.offt_r2c(p, x, X, np.zeros(n))
    L.offt_destroy(p)
    return X[0::2] + 1j * X[1::2]


def irfft_unnorm(X, n):
    L = lib()
    Xi = np.zeros(2 * (n // 2 + 1))
    Xi[0::2] = X.real
    Xi[1::2] = X.imag
    p = L.offt_create(n)
    x = np.zeros(n)
    L.offt_c2r(p, Xi, x, np.zeros(n))
    L.offt_destroy(p)
    return x


def build_filter(fsample, nshort, nof, sigma, psdfreqs, psd, radiometers=True):
    L = lib()
    out = np.zeros(nof // 2 + 1)
    L.oracle_build_filter(fsample, nshort, nof, int(radiometers), sigma, psdfreqs.size,
                          np.ascontiguousarray(psdfreqs, np.float64),
                          np.ascontiguousarray(psd, np.float64), out)
    return out


def convolve(fcov, x, nof):
    """One segment's prior convolution: y = conv(x - mean, c2r(fcov) / nof) - mean
    (src/noise_routines.f90:816-857)."""
    L = lib()
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros_like(x)
    L.oracle_convolve(nof, np.ascontiguousarray(fcov, np.float64), x, x.size, y)
    return y


def interpolate_psd(freq, psd, target):
    L = lib()
    out = np.zeros(len(target))
    L.oracle_interpolate_psd(len(freq), np.ascontiguousarray(freq, np.float64),
                             np.ascontiguousarray(psd, np.float64), len(target),
                             np.ascontiguousarray(target, np.float64), out)
    return out


def sincos_turns(u):
    L = lib()
    s, c = ct.c_double(), ct.c_double()
    L.oscan_sincos_turns(u, ct.byref(s), ct.byref(c))
    return s.value, c.value


def atan2_pos(y, x):
    return lib().oscan_atan2_pos(y, x)


def vec2pix_nest(nside, x, y, z):
    return lib().oscan_vec2pix_nest(nside, x, y, z)


def expand_pointing(scan, det, nside, i0, n):
    """CPU pointing expansion of the synthetic ring scan (row P0)."""
    L = lib()
    sp = _ScanParams(scan["fsample"], scan["spin_period"], scan["prec_period"],
                     scan["opening_turns"], scan["lon0_turns"])
    dp = _ScanDet(det["psi_turns"], det.get("open_offset_turns", 0.0), det.get("spin_offset_turns", 0.0))
    pix = np.zeros(n, np.int32)
    w = np.zeros(3 * n, np.float32)
    L.oscan_expand(ct.byref(sp), ct.byref(dp), nside, i0, n, pix, w)
    return pix, w.reshape(n, 3)
