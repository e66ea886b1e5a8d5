"""MI355X-native destriping solve with the interface of libmadam.

``libmadam_amd.libmadam_wrapper`` mirrors the reference's ctypes wrapper
(python/libmadam_wrapper/__init__.py): ``destripe``, ``destripe_with_cache``,
``clear_caches``.  ``Madam`` is the handle-based extension (in-memory maps,
device-resident buffers).  The native library ``libmadam.so`` (HIP, gfx950)
is required; there is no CPU fallback.
"""
import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmadam.so")
_lib = None

TIMESTAMP_TYPE = np.float64
SIGNAL_TYPE = np.float64
PIXEL_TYPE = np.int32
WEIGHT_TYPE = np.float32
PSD_TYPE = np.float64

# order of madam_get_stats
STAT_NAMES = ["niter", "rrinit", "rzinit", "t_pointing", "t_bin", "t_init", "t_precond", "t_iter", "t_final",
              "t_scatter", "t_gather", "t_prior", "t_apply_prec", "n_scatter", "n_gather", "precond_kind",
              "nband", "nseg", "nof", "rr_final", "npix_hit", "nactive", "t_total", "t_step", "ncg",
              "band_chunks", "band_chunks_unique", "band_blocks", "band_blocks_unique", "rec_bytes",
              "band_rows_factored", "prec_reused", "pointing_reused", "nwide"]


class MadamError(RuntimeError):
    pass


def lib():
    """Load libmadam.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MadamError("libmadam.so not found at %s; build it with python -m libmadam_amd.build" % LIB_PATH)
        L = ct.CDLL(LIB_PATH)
        vp, cp, lg, i32 = ct.c_void_p, ct.c_char_p, ct.c_long, ct.c_int
        L.madam_create.restype = vp
        L.madam_create.argtypes = [i32]
        L.madam_destroy.argtypes = [vp]
        L.madam_configure.restype = i32
        L.madam_configure.argtypes = [vp, cp, lg, cp, vp, lg, lg, vp, lg, vp, vp, lg, vp, lg, vp, lg, vp, i32]
        L.madam_set_pointing.restype = i32
        L.madam_set_pointing.argtypes = [vp, vp, vp, i32]
        L.madam_solve.restype = i32
        L.madam_solve.argtypes = [vp, vp, i32]
        L.madam_restore_pixels.restype = i32
        L.madam_restore_pixels.argtypes = [vp, vp]
        L.madam_get_sizes.restype = i32
        L.madam_get_sizes.argtypes = [vp, vp]
        L.madam_get_maps.restype = i32
        L.madam_get_maps.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.madam_get_baselines.restype = i32
        L.madam_get_baselines.argtypes = [vp, vp, vp, vp]
        L.madam_get_stats.restype = i32
        L.madam_get_stats.argtypes = [vp, vp, i32]
        L.madam_get_rr_history.restype = i32
        L.madam_get_rr_history.argtypes = [vp, vp, i32]
        L.madam_iterate.restype = i32
        L.madam_iterate.argtypes = [vp, i32, i32]
        L.madam_solve_batch.restype = i32
        L.madam_solve_batch.argtypes = [vp, i32, vp, i32, i32, i32]
        L.madam_iterate_batch.restype = i32
        L.madam_iterate_batch.argtypes = [vp, i32]
        L.madam_get_batch_maps.restype = i32
        L.madam_get_batch_maps.argtypes = [vp, i32, vp, vp]
        L.madam_get_batch_baselines.restype = i32
        L.madam_get_batch_baselines.argtypes = [vp, i32, vp, vp]
        L.madam_get_batch_info.restype = i32
        L.madam_get_batch_info.argtypes = [vp, i32, vp, i32]
        L.madam_get_hits.restype = i32
        L.madam_get_hits.argtypes = [vp, vp]
        L.madam_get_csr.restype = i32
        L.madam_get_csr.argtypes = [vp, vp, vp, vp]
        L.madam_server_create.restype = vp
        L.madam_server_create.argtypes = [i32, ct.c_double, i32]
        L.madam_server_destroy.argtypes = [vp]
        L.madam_server_submit.restype = lg
        L.madam_server_submit.argtypes = [vp, cp, lg, vp, lg, lg, vp, vp, vp, vp, i32, lg, vp, vp, lg, vp, lg, vp,
                                          lg, vp, ct.c_ulonglong, i32, i32]
        L.madam_server_wait.restype = i32
        L.madam_server_wait.argtypes = [vp, lg, ct.c_double]
        L.madam_server_error.restype = cp
        L.madam_server_error.argtypes = [vp, lg]
        L.madam_server_result.restype = i32
        L.madam_server_result.argtypes = [vp, lg, vp, vp, vp, vp, vp, vp]
        L.madam_server_release.restype = i32
        L.madam_server_release.argtypes = [vp, lg]
        L.madam_server_stats.restype = i32
        L.madam_server_stats.argtypes = [vp, vp, i32]
        L.madam_get_filter.restype = i32
        L.madam_get_filter.argtypes = [vp, i32, vp]
        L.madam_apply.restype = i32
        L.madam_apply.argtypes = [vp, i32, vp, vp]
        L.madam_set_kernel_variant.restype = i32
        L.madam_set_kernel_variant.argtypes = [vp, i32]
        L.madam_reset_timers.restype = i32
        L.madam_reset_timers.argtypes = [vp]
        L.madam_set_profiling.restype = i32
        L.madam_set_profiling.argtypes = [vp, i32]
        L.madam_expand_pointing.restype = i32
        L.madam_expand_pointing.argtypes = [vp, vp, i32, lg, lg, lg, vp, vp, i32]
        L.madam_last_error.restype = cp
        L.madam_last_status.restype = i32
        L.madam_cache_handle.restype = i32
        L.madam_cache_handle.argtypes = [ct.POINTER(vp)]
        L.destripe.restype = None
        L.destripe_with_cache.restype = None
        L.clear_caches.restype = None
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, int):
        return ct.c_void_p(a)
    if hasattr(a, "data_ptr"):          # torch tensor on the device
        return ct.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ct.c_void_p)


def _check(rc):
    if rc != 0:
        raise MadamError(lib().madam_last_error().decode())


def dict2parstring(d):
    """Parameter string in the reference format (python/libmadam_wrapper/__init__.py:73-84)."""
    if isinstance(d, (str, bytes)):
        return d.encode("ascii") if isinstance(d, str) else d
    s = ""
    for key, value in d.items():
        if key in ("detset", "detset_nopol", "survey"):
            for v in value:
                s += "{} = {};".format(key, v)
        else:
            s += "{} = {};".format(key, value)
    return s.encode("ascii")


def dets2detstring(dets):
    return "".join("{};".format(d) for d in dets).encode("ascii")


class Madam:
    """One destriping configuration on one GPU (handle over the C-ABI)."""

    def __init__(self, device=0):
        self._h = lib().madam_create(device)
        if not self._h:
            raise MadamError(lib().madam_last_error().decode())
        self._keep = []

    def close(self):
        if self._h:
            lib().madam_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def configure(self, pars, dets, weights, timestamps, periods, npsd, psdstarts, psdfreqs, psdvals, nnz,
                  subchunk_pass=0):
        self.ndet = len(dets)
        self.nsamp = int(np.asarray(timestamps).size)
        self.nnz = int(nnz)
        arrs = [np.ascontiguousarray(weights, np.float64), np.ascontiguousarray(timestamps, np.float64),
                np.ascontiguousarray(periods, np.int64), np.ascontiguousarray(npsd, np.int64),
                np.ascontiguousarray(psdstarts, np.float64), np.ascontiguousarray(psdfreqs, np.float64),
                np.ascontiguousarray(psdvals, np.float64).reshape(-1)]
        w, ts, per, nps, pst, pfr, pv = arrs
        npsdbin = pfr.size
        npsdval = pv.size
        npsdtot = npsdval // max(npsdbin, 1)
        _check(lib().madam_configure(self._h, dict2parstring(pars), self.ndet, dets2detstring(dets), _ptr(w),
                                     self.nsamp, self.nnz, _ptr(ts), per.size, _ptr(per), _ptr(nps), npsdtot,
                                     _ptr(pst), npsdbin, _ptr(pfr), npsdval, _ptr(pv), int(subchunk_pass)))
        return self

    def set_pointing(self, pixels, pixweights, on_device=False):
        if not on_device:
            assert pixels.dtype == np.int32 and pixweights.dtype == np.float32
            assert pixels.flags.c_contiguous and pixweights.flags.c_contiguous
        self._pix = pixels
        self._keep = [pixels, pixweights]
        _check(lib().madam_set_pointing(self._h, _ptr(pixels), _ptr(pixweights), int(on_device)))

    def solve(self, signal, on_device=False):
        if not on_device:
            assert signal.dtype == np.float64 and signal.flags.c_contiguous
        self._keep.append(signal)
        _check(lib().madam_solve(self._h, _ptr(signal), int(on_device)))

    def restore_pixels(self, host_pix):
        _check(lib().madam_restore_pixels(self._h, _ptr(host_pix)))

    def sizes(self):
        out = np.zeros(6, np.int64)
        _check(lib().madam_get_sizes(self._h, _ptr(out)))
        return dict(nmap=int(out[0]), npix=int(out[1]), nbase=int(out[2]), ndet=int(out[3]), niter=int(out[4]),
                    nside_map=int(out[5]))

    def maps(self):
        sz = self.sizes()
        npix, nmap = sz["npix"], sz["nmap"]
        m = np.zeros((npix, nmap))
        b = np.zeros((npix, nmap))
        h = np.zeros(npix, np.int32)
        cc = np.zeros((npix, nmap, nmap))
        mk = np.zeros(npix, np.int32)
        cr = np.zeros(npix)
        _check(lib().madam_get_maps(self._h, _ptr(m), _ptr(b), _ptr(h), _ptr(cc), _ptr(mk), _ptr(cr)))
        return dict(map=m, binmap=b, hits=h, cc=cc, mask=mk, crit=cr)

    def baselines(self):
        sz = self.sizes()
        shape = (sz["ndet"], sz["nbase"])
        aa, yba, nna = np.zeros(shape), np.zeros(shape), np.zeros(shape)
        _check(lib().madam_get_baselines(self._h, _ptr(aa), _ptr(yba), _ptr(nna)))
        return dict(aa=aa, yba=yba, nna=nna)

    def stats(self):
        out = np.zeros(len(STAT_NAMES))
        _check(lib().madam_get_stats(self._h, _ptr(out), len(STAT_NAMES)))
        return dict(zip(STAT_NAMES, out.tolist()))

    def csr(self):
        """The samples grouped by map pixel: off [npix + 1], src [nactive]
        (sample index det * nsamp + t of each entry) and bord [ndet * nbase]
        (the baseline of each slot of the projections' baseline order)."""
        sz = self.sizes()
        off = np.zeros(sz["npix"] + 1, np.int64)
        src = np.zeros(int(self.stats()["nactive"]), np.uint32)
        bord = np.zeros(sz["ndet"] * sz["nbase"], np.int32)
        _check(lib().madam_get_csr(self._h, _ptr(off), _ptr(src), _ptr(bord)))
        return dict(off=off, src=src, bord=bord)

    def rr_history(self):
        n = self.sizes()["niter"]
        out = np.zeros(max(n, 1))
        _check(lib().madam_get_rr_history(self._h, _ptr(out), n))
        return out[:n]

    def iterate(self, niter, restart=False):
        _check(lib().madam_iterate(self._h, int(niter), int(restart)))

    def apply(self, which, x):
        """which: 0 noise prior, 1 preconditioner, 2 destriping operator A (on a
        host vector of ndet * nbase baselines, after solve())."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        y = np.zeros_like(x)
        _check(lib().madam_apply(self._h, int(which), _ptr(x), _ptr(y)))
        return y

    def solve_batch(self, signals, precision=0, on_device=False):
        """Solve nrhs signals [nrhs][ndet][nsamp] (float64 or float32) sharing this
        handle's pointing; precision 0 = FP64 PCG, 1 = FP32 operator inside FP64
        iterative refinement."""
        if on_device:
            nrhs = int(signals.shape[0])
            f32 = str(signals.dtype) in ("torch.float32", "float32")
        else:
            signals = np.ascontiguousarray(signals)
            assert signals.dtype in (np.float64, np.float32)
            nrhs = signals.shape[0]
            f32 = signals.dtype == np.float32
        self._keep.append(signals)
        self.nrhs = nrhs
        _check(lib().madam_solve_batch(self._h, nrhs, _ptr(signals), int(f32), int(on_device), int(precision)))

    def iterate_batch(self, niter):
        _check(lib().madam_iterate_batch(self._h, int(niter)))

    def batch_result(self, irhs):
        """map, binmap [npix][nmap], aa, yba [ndet][nbase] and the solve info of right-hand side irhs."""
        sz = self.sizes()
        npix, nmap = sz["npix"], sz["nmap"]
        m = np.zeros((npix, nmap))
        b = np.zeros((npix, nmap))
        _check(lib().madam_get_batch_maps(self._h, int(irhs), _ptr(m), _ptr(b)))
        aa = np.zeros((sz["ndet"], sz["nbase"]))
        yba = np.zeros((sz["ndet"], sz["nbase"]))
        _check(lib().madam_get_batch_baselines(self._h, int(irhs), _ptr(aa), _ptr(yba)))
        keys = ["steps", "status", "rr_inner", "passes", "rr_final", "t_setup", "t_solve", "pcg_steps", "t_pcg",
                "piece", "t_scatter", "t_prior", "t_gather", "t_prec", "t_step", "n_prof", "rle_records_gather",
                "rle_records_scatter"]
        info = np.zeros(len(keys))
        _check(lib().madam_get_batch_info(self._h, int(irhs), _ptr(info), len(keys)))
        return dict(map=m, binmap=b, aa=aa, yba=yba, info=dict(zip(keys, info.tolist())))

    def filter(self, ipsd=0):
        """The noise filter fcov[nof / 2 + 1] of global PSD ipsd (after solve())."""
        out = np.zeros(int(self.stats()["nof"]) // 2 + 1)
        _check(lib().madam_get_filter(self._h, int(ipsd), _ptr(out)))
        return out

    def set_kernel_variant(self, v):
        _check(lib().madam_set_kernel_variant(self._h, int(v)))

    def reset_timers(self):
        _check(lib().madam_reset_timers(self._h))

    def set_profiling(self, on=True):
        _check(lib().madam_set_profiling(self._h, int(on)))


class Server:
    """Request loop on one GPU (madam_server_*): submit destripe jobs, wait for
    them, read their maps.  Jobs with the same non-zero scan_key share one
    resident set-up and are batched when they wait together."""

    STAT_NAMES = ["done", "failed", "batches", "batched_jobs", "resident_jobs", "evictions", "peak_bytes",
                  "max_inflight", "resident_scans", "memory_evictions", "min_free_bytes"]

    def __init__(self, device=0, hbm_budget=200e9, max_inflight=4):
        self._h = lib().madam_server_create(int(device), float(hbm_budget), int(max_inflight))
        self._jobs = {}

    def close(self):
        if self._h:
            lib().madam_server_destroy(self._h)
            self._h = None
            self._jobs = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, pars, weights, timestamps, pixels, pixweights, signal, periods, npsd, psdstarts, psdfreqs,
               psdvals, nnz=3, scan_key=0, precision=0, on_device=False):
        """signal [ndet][nsamp] float64 or float32; the arrays are kept until the job is released.
        on_device: pixels (int32), pixweights (float32) and signal are contiguous device tensors."""
        w = np.ascontiguousarray(weights, np.float64)
        ts = np.ascontiguousarray(timestamps, np.float64)
        if on_device:
            pix, pw, sig = pixels, pixweights, signal
            assert pix.is_contiguous() and pw.is_contiguous() and sig.is_contiguous()
            assert str(pix.dtype) == "torch.int32" and str(pw.dtype) == "torch.float32"
            f32 = str(sig.dtype) == "torch.float32"
            assert f32 or str(sig.dtype) == "torch.float64"
        else:
            pix = np.ascontiguousarray(pixels, np.int32)
            pw = np.ascontiguousarray(pixweights, np.float32)
            sig = np.ascontiguousarray(signal)
            assert sig.dtype in (np.float64, np.float32)
            f32 = sig.dtype == np.float32
        per = np.ascontiguousarray(periods, np.int64)
        nps = np.ascontiguousarray(npsd, np.int64)
        pst = np.ascontiguousarray(psdstarts, np.float64)
        pfr = np.ascontiguousarray(psdfreqs, np.float64)
        pv = np.ascontiguousarray(psdvals, np.float64).reshape(-1)
        keep = (w, ts, pix, pw, sig, per, nps, pst, pfr, pv)
        ndet, nsamp = w.size, ts.size
        npsdbin = pfr.size
        jid = lib().madam_server_submit(self._h, dict2parstring(pars), ndet, _ptr(w), nsamp, int(nnz), _ptr(ts),
                                        _ptr(pix), _ptr(pw), _ptr(sig), int(f32), per.size,
                                        _ptr(per), _ptr(nps), pv.size // max(npsdbin, 1), _ptr(pst), npsdbin,
                                        _ptr(pfr), pv.size, _ptr(pv), int(scan_key), int(precision),
                                        int(bool(on_device)))
        if jid < 0:
            raise MadamError("job exceeds the server's HBM budget")
        self._jobs[jid] = (keep, ndet)
        return jid

    def wait(self, jid, timeout=-1.0):
        rc = lib().madam_server_wait(self._h, int(jid), float(timeout))
        if rc == 1:
            raise MadamError(lib().madam_server_error(self._h, int(jid)).decode())
        return rc == 0

    def result(self, jid, maps=True):
        """maps=False: sizes, batch, resident flag and times only (no map copies)"""
        sz = np.zeros(6, np.int64)
        tm = np.zeros(4)
        _check(lib().madam_server_result(self._h, int(jid), None, None, None, None, _ptr(sz), _ptr(tm)))
        if not maps:
            return dict(batch=int(sz[4]), resident=int(sz[5]), t_submit=tm[0], t_start=tm[1], t_done=tm[2],
                        latency=tm[2] - tm[0], steps=tm[3])
        npix, nmap, nbase, ndet = int(sz[0]), int(sz[1]), int(sz[2]), int(sz[3])
        m = np.zeros((npix, nmap))
        b = np.zeros((npix, nmap))
        h = np.zeros(npix, np.int32)
        aa = np.zeros((ndet, nbase))
        _check(lib().madam_server_result(self._h, int(jid), _ptr(m), _ptr(b), _ptr(h), _ptr(aa), None, None))
        return dict(map=m, binmap=b, hits=h, aa=aa, batch=int(sz[4]), resident=int(sz[5]), t_submit=tm[0],
                    t_start=tm[1], t_done=tm[2], latency=tm[2] - tm[0], steps=tm[3])

    def release(self, jid):
        lib().madam_server_release(self._h, int(jid))
        self._jobs.pop(jid, None)

    def stats(self):
        out = np.zeros(len(self.STAT_NAMES))
        lib().madam_server_stats(self._h, _ptr(out), len(out))
        return dict(zip(self.STAT_NAMES, out.tolist()))


def cache_stats():
    """Statistics of the context that destripe(mcmode=T) cached (None if empty)."""
    h = ct.c_void_p()
    lib().madam_cache_handle(ct.byref(h))
    if not h.value:
        return None
    out = np.zeros(len(STAT_NAMES))
    _check(lib().madam_get_stats(h, _ptr(out), len(STAT_NAMES)))
    return dict(zip(STAT_NAMES, out.tolist()))


def run(pars, dets, weights, timestamps, pixels, pixweights, signal, periods, npsd, psdstarts, psdfreqs, psdvals,
        subchunk_pass=0, device=0, kernel_variant=0):
    """Full solve with in-memory outputs.  Mutates ``pixels`` and ``signal`` the
    way the reference's destripe() mutates the caller's buffers."""
    ndet = len(dets)
    nsamp = timestamps.size
    nnz = pixweights.size // max(nsamp * ndet, 1)
    m = Madam(device)
    try:
        m.configure(pars, dets, weights, timestamps, periods, npsd, psdstarts, psdfreqs, psdvals, nnz,
                    subchunk_pass)
        m.set_kernel_variant(kernel_variant)
        m.set_pointing(pixels, pixweights)
        m.solve(signal)
        m.restore_pixels(pixels)
        out = m.maps()
        out.update(m.baselines())
        out["stats"] = m.stats()
        out["rr"] = m.rr_history()
        out["niter"] = int(out["stats"]["niter"])
        return out
    finally:
        m.close()


def expand_pointing(scan, dets, nside, i0, n):
    """GPU pointing expansion of the synthetic ring scan (host outputs)."""
    scan5 = np.array([scan["fsample"], scan["spin_period"], scan["prec_period"], scan["opening_turns"],
                      scan["lon0_turns"]], np.float64)
    d3 = np.array([[d["psi_turns"], d.get("open_offset_turns", 0.0), d.get("spin_offset_turns", 0.0)]
                   for d in dets], np.float64)
    pix = np.zeros((len(dets), n), np.int32)
    w = np.zeros((len(dets), n, 3), np.float32)
    _check(lib().madam_expand_pointing(_ptr(scan5), _ptr(d3), len(dets), nside, i0, n, _ptr(pix), _ptr(w), 0))
    return pix, w


def expand_pointing_device(scan, dets, nside, i0, n, pix_ptr, w_ptr):
    """GPU pointing expansion written straight into device buffers
    pix int32[ndet][n] and w float32[ndet][n][3] (device pointers or tensors)."""
    scan5 = np.array([scan["fsample"], scan["spin_period"], scan["prec_period"], scan["opening_turns"],
                      scan["lon0_turns"]], np.float64)
    d3 = np.array([[d["psi_turns"], d.get("open_offset_turns", 0.0), d.get("spin_offset_turns", 0.0)]
                   for d in dets], np.float64)
    _check(lib().madam_expand_pointing(_ptr(scan5), _ptr(d3), len(dets), nside, i0, n, _ptr(pix_ptr), _ptr(w_ptr), 1))
