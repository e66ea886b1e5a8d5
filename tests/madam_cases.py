"""Synthetic destriping problems and dense brute-force solutions for the tests.

Inputs follow the reference ABI layouts (src/smadam.F90:151-154): pixels int32
[ndet*nsamp], weights float32 [ndet*nsamp*nnz], signal float64 [ndet*nsamp],
detector-major.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

SCAN = dict(fsample=50.0, spin_period=60.0, prec_period=360 * 3600.0,
            opening_turns=85.0 / 360.0, lon0_turns=0.0)


def reference_test_case(mc=False):
    """Inputs of python/tests/test_libmadam.py:41-120 (or _mc.py:41-130)."""
    name = "ref_test_libmadam_mc.npz" if mc else "ref_test_libmadam.npz"
    g = np.load(os.path.join(GOLDEN, name))
    nside = 8
    fsample = 32.5
    nsamp = 1000
    ndet = 4
    pars = dict(base_first=1.0, fsample=fsample, nside_map=nside,
                nside_cross=nside if mc else nside // 2,
                nside_submap=nside if mc else nside // 4,
                write_map=True, write_binmap=True, write_matrix=True, write_wcov=True,
                write_hits=True, write_leakmatrix=True, kfilter=True,
                file_root="madam_pytest", path_output="./pymaps/", info=0)
    if mc:
        pars["nsubchunk"] = 1
        pars["mcmode"] = False
    else:
        pars.update(diagfilter=0, iter_max=100, nsubchunk=2, allreduce=True,
                    survey=["hm1 : {} - {}".format(0, nsamp / 2)], bin_subsets=True)
    case = dict(
        pars=pars, dets=["LFI27M", "LFI27S", "LFI28M", "LFI28S"],
        weights=np.ones(ndet), timestamps=np.arange(nsamp, dtype=np.float64),
        pixels=g["pixels"].copy(), pixweights=np.ones(ndet * nsamp, np.float32),
        signal=g["signal"].copy(),
        periods=np.array([0, 250, 500, 750], np.int64),
        npsd=np.ones(ndet, np.int64), psdstarts=np.zeros(ndet),
        psdfreqs=np.arange(10) * fsample / 10, psdvals=np.ones(10 * ndet))
    return case, g


def psd_model(freqs, sigma, fsample, fknee, alpha):
    return sigma ** 2 / fsample * (1.0 + (freqs / fknee) ** alpha)


def one_over_f(rng, n, fsample, sigma, fknee, alpha):
    """White + 1/f noise shaped by FFT: sqrt(1 + (f/fknee)^alpha) on white spectra."""
    white = rng.standard_normal(n) * sigma
    f = np.fft.rfftfreq(n, 1.0 / fsample)
    shape = np.ones_like(f)
    shape[1:] = np.sqrt(1.0 + (f[1:] / fknee) ** alpha)
    shape[0] = 0.0
    return np.fft.irfft(np.fft.rfft(white) * shape, n)


def ring_case(ndet=4, nsamp=4096, nside=8, fsample=50.0, base_first=1.0, nnz=3,
              seed=1, prior=False, sigma=1e-4, fknee=0.1, alpha=-1.0, nperiod=1,
              sky_sigma=(100e-6, 5e-6, 5e-6), noise=True, offsets=None, spin_period=2.0,
              prec_period=None, extra_pars=None, pointing=None):
    """Small ring-scan problem.  Pointing comes from ``pointing(idet, nsamp)``
    (the CPU P0 expansion in the tests) or a deterministic fallback."""
    rng = np.random.default_rng(seed)
    npix = 12 * nside ** 2
    pix = np.zeros((ndet, nsamp), np.int32)
    w = np.zeros((ndet, nsamp, 3), np.float32)
    for d in range(ndet):
        p, ww = pointing(d, nsamp)
        pix[d] = p
        w[d] = ww
    sky = np.zeros((npix, 3))
    for j in range(3):
        sky[:, j] = rng.standard_normal(npix) * sky_sigma[j]
    if nnz == 1:
        sig = sky[pix, 0].copy()
        w = w[:, :, :1].copy()
        w[:] = 1.0
    else:
        sig = np.einsum("dsk,dsk->ds", sky[pix], w.astype(np.float64))
    if noise:
        for d in range(ndet):
            sig[d] += one_over_f(rng, nsamp, fsample, sigma, fknee, alpha)
    if offsets is not None:
        sig += offsets
    periods = np.linspace(0, nsamp, nperiod + 1)[:-1].astype(np.int64)
    freqs = np.logspace(-5, np.log10(fsample / 2), 100)
    psd = psd_model(freqs, sigma, fsample, fknee, alpha)
    pars = dict(base_first=base_first, fsample=fsample, nside_map=nside, nside_cross=nside,
                nside_submap=max(1, nside // 4), kfilter=prior, info=0,
                pixlim_cross=1e-3, pixlim_map=1e-3)
    if extra_pars:
        pars.update(extra_pars)
    case = dict(pars=pars, dets=["d%02d" % d for d in range(ndet)],
                weights=np.full(ndet, 1.0 / sigma ** 2), timestamps=np.arange(nsamp) / fsample,
                pixels=pix.reshape(-1).copy(), pixweights=w.reshape(-1).copy(),
                signal=sig.reshape(-1).copy(), periods=periods,
                npsd=np.ones(ndet, np.int64), psdstarts=np.zeros(ndet), psdfreqs=freqs,
                psdvals=np.tile(psd, ndet), sky=sky, nnz=nnz)
    return case


def call_args(case):
    return [case[k] for k in ("pars", "dets", "weights", "timestamps", "pixels", "pixweights",
                              "signal", "periods", "npsd", "psdstarts", "psdfreqs", "psdvals")]


def copy_case(case):
    out = dict(case)
    for k in ("pixels", "pixweights", "signal"):
        out[k] = case[k].copy()
    return out


# ------------------------------------------------------------------ dense brute force


def dense_problem(case, res, prior_kernel=None, wsamp=None):
    """Dense normal equations of the destriping problem for order-0 baselines.

    Uses the same sample selection as the reference: samples whose cross pixel
    was rejected (cca == 0) are dropped from F (src/smadam_routines.f90:542-544).
    Returns (A, b, P_good, F_good, W) in baseline coordinates [ndet*nbase].
    ``prior_kernel(idet, kstart, noba)`` returns the dense C_a^-1 block.
    ``wsamp`` [ndet, nsamp] is the detector weight of every sample (default:
    the detector's weight everywhere).
    """
    ndet = len(case["dets"])
    nsamp = case["timestamps"].size
    nnz = case["pixweights"].size // (ndet * nsamp)
    nb = res["nbase"]
    bstart, blen = res["base_start"], res["base_len"]
    base_of = np.zeros(nsamp, np.int64)
    for k in range(nb):
        base_of[bstart[k]:bstart[k] + blen[k]] = k
    pix = case["pixels"].reshape(ndet, nsamp)
    w = case["pixweights"].reshape(ndet, nsamp, nnz).astype(np.float64)
    sig = case["signal"].reshape(ndet, nsamp)
    dw = np.asarray(case["weights"], np.float64)
    npix = res["npix"]
    nmap = res["nmap"]
    # pixel matrix and binned map
    Wd = np.repeat(dw, nsamp) if wsamp is None else np.asarray(wsamp, np.float64).reshape(-1)
    rows = np.arange(ndet * nsamp)
    ipix = pix.reshape(-1)
    ok = ipix >= 0
    ww = w.reshape(-1, nnz)[:, :nmap] if nmap == 3 else np.ones((ndet * nsamp, 1))
    P = np.zeros((ndet * nsamp, npix * nmap))
    for j in range(nmap):
        P[rows[ok], ipix[ok] * nmap + j] = ww[ok, j]
    PtWP = P.T @ (Wd[:, None] * P)
    # good pixels = those the oracle did not reject
    goodpix = np.zeros(npix * nmap, bool)
    for q in range(npix):
        if np.any(res["cc"][q] != 0):
            goodpix[q * nmap:(q + 1) * nmap] = True
    gs = ok & goodpix[np.where(ok, ipix, 0) * nmap]
    Pg = P[gs][:, goodpix]
    Wg = Wd[gs]
    F = np.zeros((ndet * nsamp, ndet * nb))
    det_of = np.repeat(np.arange(ndet), nsamp)
    F[rows, det_of * nb + np.tile(base_of, ndet)] = 1.0
    Fg = F[gs]
    M = Pg.T @ (Wg[:, None] * Pg)
    Minv = np.linalg.inv(M)
    FtWP = Fg.T @ (Wg[:, None] * Pg)
    dg = sig.reshape(-1)[gs]
    A = Fg.T @ (Wg[:, None] * Fg) - FtWP @ Minv @ FtWP.T
    b = Fg.T @ (Wg * dg) - FtWP @ (Minv @ (Pg.T @ (Wg * dg)))
    if prior_kernel is not None:
        for d in range(ndet):
            for (ks, no) in prior_kernel["segments"][d]:
                blk = prior_kernel["block"](d, ks, no)
                A[d * nb + ks:d * nb + ks + no, d * nb + ks:d * nb + ks + no] += blk
    return A, b, PtWP


def prior_block(invcov, noba):
    """Dense C_a^-1 for one trimmed segment: mean removal, zero-padded circular
    convolution with invcov = c2r(fcov)/nof, mean removal
    (src/noise_routines.f90:816-857)."""
    nof = invcov.size
    idx = (np.arange(noba)[:, None] - np.arange(noba)[None, :]) % nof
    T = invcov[idx]
    Pm = np.eye(noba) - np.ones((noba, noba)) / noba
    return Pm @ T @ Pm


def project_out_null(A, x, tol=1e-9):
    """Component of x orthogonal to the numerical null space of symmetric A."""
    ev, V = np.linalg.eigh(A)
    null = np.abs(ev) < tol * np.abs(ev).max()
    N = V[:, null]
    return x - N @ (N.T @ x), N


def with_psd_switch(case, switch_times, levels=(1.0, 3.0), fknees=(0.1, 0.3), alphas=(-1.0, -1.5), sigma=1e-4):
    """Two PSDs per detector (src/inputparam.f90:73-163): detector d uses PSD 0
    before ``switch_times[d]`` and PSD 1 from then on; PSD j is
    levels[j] * sigma^2 / fs * (1 + (f / fknees[j])^alphas[j])."""
    ndet = len(case["dets"])
    fs = case["pars"]["fsample"]
    freqs = case["psdfreqs"]
    rows = []
    for d in range(ndet):
        for j in range(2):
            rows.append(levels[j] * psd_model(freqs, sigma, fs, fknees[j], alphas[j]))
    case["npsd"] = np.full(ndet, 2, np.int64)
    case["psdstarts"] = np.array([[0.0, t] for t in switch_times], np.float64).reshape(-1)
    case["psdvals"] = np.concatenate(rows)
    return case
