"""Pins for the FP64 CPU oracle (oracle/madam_oracle.c).

Each check is chosen so that a plausible slip (a dropped term, wrong sign,
transposed operand, wrong normalisation) fails it:
  * known answers of the reference's own tests (hits = bincount, binned map
    = per-pixel mean, subchunk hit additivity);
  * dense brute-force normal equations for tiny IQU problems, with and
    without the noise prior, and with every preconditioner;
  * closed forms / library routines for the FFT, eigen-inverse, band
    Cholesky, PSD interpolation and the noise filter.
"""
import numpy as np
import pytest

import oracle
from tests import madam_cases as mc


def run(case, **kw):
    c = mc.copy_case(case)
    r = oracle.destripe(*mc.call_args(c), **kw)
    return r, c


# --------------------------------------------------------------- known answers


def test_reference_known_answers():
    case, g = mc.reference_test_case()
    r, c = run(case)
    # python/tests/test_libmadam.py:155-169: hits and binned map (nnz = 1)
    assert np.array_equal(r["hits"], g["hmap"])
    np.testing.assert_allclose(r["binmap"][:, 0], g["bmap"], rtol=0, atol=1e-12)
    assert np.array_equal(r["binmap"][:, 0].astype(np.float32), g["bmap32"])
    # 171-174: hits = hits_sub1of2 + hits_sub2of2
    r1, _ = run(case, subchunk_pass=1)
    r2, _ = run(case, subchunk_pass=2)
    assert np.array_equal(r1["hits"] + r2["hits"], r["hits"])
    assert r1["hits"].sum() == r2["hits"].sum() == 2000


def test_reference_mc_known_answers():
    case, g = mc.reference_test_case(mc=True)
    r, _ = run(case)
    assert np.array_equal(r["hits"], g["hmap"])
    np.testing.assert_allclose(r["binmap"][:, 0], g["bmap"], rtol=0, atol=1e-12)


def test_reference_buffers_mutated_like_reference():
    case, g = mc.reference_test_case()
    c = mc.copy_case(case)
    # flag one sample by zero weight, one by pixel -1
    c["pixweights"][5] = 0
    c["pixels"][7] = -1
    sig0 = c["signal"].copy()
    r = oracle.destripe(*mc.call_args(c))
    # flagged samples end as 12*nside_max**2 (src/pointing.f90:168-191)
    assert c["pixels"][5] == 12 * 8 ** 2 and c["pixels"][7] == 12 * 8 ** 2
    assert np.array_equal(np.delete(c["pixels"], [5, 7]), np.delete(g["pixels"], [5, 7]))
    # survey defined -> clean_tod subtracts F a from the caller's signal
    nb, bs, bl = r["nbase"], r["base_start"], r["base_len"]
    exp = sig0.reshape(4, -1).copy()
    for k in range(nb):
        exp[:, bs[k]:bs[k] + bl[k]] -= r["aa"][:, k:k + 1]
    np.testing.assert_allclose(c["signal"], exp.reshape(-1), rtol=0, atol=1e-12)


def test_parameter_errors():
    case, _ = mc.reference_test_case(mc=True)
    case = mc.copy_case(case)
    case["pars"] = dict(case["pars"], no_such_key=1)
    with pytest.raises(oracle.OracleError, match="Unknown parameter"):
        oracle.destripe(*mc.call_args(case))
    case, _ = mc.reference_test_case(mc=True)
    case = mc.copy_case(case)
    case["pixels"][3] = 12 * 64 ** 2
    with pytest.raises(oracle.OracleError, match="Too large pixel"):
        oracle.destripe(*mc.call_args(case))


# --------------------------------------------------------------- dense brute force

SCAN_SMALL = dict(fsample=50.0, spin_period=2.0, prec_period=400.0,
                  opening_turns=85.0 / 360.0, lon0_turns=0.0)


def scan_pointing(nside, psis=(0.0, 0.125, 0.25, 0.375), offs=(0.0, 0.01, 0.0, 0.01)):
    def f(d, n):
        return oracle.expand_pointing(SCAN_SMALL, dict(psi_turns=psis[d % 4],
                                                       open_offset_turns=offs[d % 4]), nside, 0, n)
    return f


def small_case(prior=False, nnz=3, extra=None, nperiod=1, seed=3, **kw):
    return mc.ring_case(ndet=kw.pop("ndet", 4), nsamp=kw.pop("nsamp", 1600), nside=4,
                        fsample=50.0, base_first=1.0, nnz=nnz, seed=seed, prior=prior,
                        nperiod=nperiod, pointing=scan_pointing(4),
                        extra_pars=dict(dict(cglimit=1e-28, iter_max=400), **(extra or {})), **kw)


def prior_kernel_for(case, res):
    """Dense prior blocks rebuilt from the oracle's own filter function."""
    pars = case["pars"]
    fs = pars["fsample"]
    nshort = int(round(pars["base_first"] * fs))
    nof = res["nof"]
    fcov = oracle.build_filter(fs, nshort, nof, 1.0 / np.sqrt(case["weights"][0]),
                               case["psdfreqs"], case["psdvals"][:case["psdfreqs"].size])
    invcov = np.fft.irfft(fcov, nof)
    ndet = len(case["dets"])
    nb = res["nbase"]
    segs = []
    for d in range(ndet):
        nz = np.nonzero(res["nna"][d])[0]
        segs.append([(int(nz[0]), int(nz[-1] - nz[0] + 1))] if nz.size else [])
    return dict(segments=segs, block=lambda d, ks, no: mc.prior_block(invcov, no)), invcov


@pytest.mark.parametrize("prior", [False, True])
@pytest.mark.parametrize("nnz", [1, 3])
def test_dense_brute_force(prior, nnz):
    case = small_case(prior=prior, nnz=nnz)
    r, _ = run(case)
    assert r["niter"] > 3
    pk = prior_kernel_for(case, r)[0] if prior else None
    A, b, _ = mc.dense_problem(case, r, pk)
    a = r["aa"].reshape(-1)
    # the oracle's right-hand side is the dense b
    np.testing.assert_allclose(r["yba"].reshape(-1), b, rtol=0, atol=1e-9 * np.abs(b).max())
    # residual of the converged solve
    res = np.linalg.norm(b - A @ a) / np.linalg.norm(b)
    assert res < 1e-10, res
    # equals the minimum-norm solution up to the null space
    x = np.linalg.lstsq(A, b, rcond=1e-12)[0]
    d, N = mc.project_out_null(A, a - x)
    assert np.linalg.norm(d) < 1e-8 * np.linalg.norm(x)
    if not prior:
        assert N.shape[1] >= 1     # the global offset is unconstrained without a prior


def test_dense_map_equals_binned_residual():
    """map = (P^T W P)^-1 P^T W (d - F a) at nside_map (src/smadam_routines.f90:1489-1539)."""
    case = small_case(prior=False)
    r, _ = run(case)
    ndet = len(case["dets"])
    nsamp = case["timestamps"].size
    pix = case["pixels"].reshape(ndet, nsamp)
    w = case["pixweights"].reshape(ndet, nsamp, 3).astype(np.float64)
    d = case["signal"].reshape(ndet, nsamp).copy()
    for k in range(r["nbase"]):
        s, l = r["base_start"][k], r["base_len"][k]
        d[:, s:s + l] -= r["aa"][:, k:k + 1]
    npix = r["npix"]
    cc = np.zeros((npix, 3, 3))
    rhs = np.zeros((npix, 3))
    dw = case["weights"]
    for i in range(ndet):
        np.add.at(cc, pix[i], dw[i] * w[i][:, :, None] * w[i][:, None, :])
        np.add.at(rhs, pix[i], dw[i] * w[i] * d[i][:, None])
    m = np.zeros((npix, 3))
    good = r["mask"] > 0
    m[good] = np.linalg.solve(cc[good], rhs[good][:, :, None])[:, :, 0]
    np.testing.assert_allclose(r["map"], m, rtol=0, atol=1e-9 * np.abs(m).max())


def test_noise_free_sky_only():
    """No noise: baselines are one constant (the null space) and the map is the sky."""
    # unit detector weights: the rounding-level RHS stays below the reference's
    # rzinit < 1e-30 cut (src/smadam_routines.f90:749-756), so no CG step is taken
    case = small_case(prior=False, noise=False, sigma=1.0)
    r, _ = run(case)
    a = r["aa"]
    assert r["niter"] == 0 and np.all(a == 0)
    good = r["mask"] > 0
    sky = case["sky"]
    dm = r["map"][good] - sky[good]
    dm[:, 0] -= dm[:, 0].mean()
    assert np.abs(dm).max() < 1e-12


def test_injected_offsets_recovered():
    rng = np.random.default_rng(7)
    ndet, nsamp = 4, 1600
    offs = np.repeat(rng.standard_normal((ndet, nsamp // 50)) * 1e-3, 50, axis=1)
    case = small_case(prior=False, noise=False, offsets=offs, extra=dict(cglimit=1e-26))
    r, _ = run(case)
    got = r["aa"] - r["aa"].mean()
    exp = offs[:, ::50] - offs[:, ::50].mean()
    np.testing.assert_allclose(got, exp, rtol=0, atol=1e-12)


def test_operator_symmetry_and_preconditioner_invariance():
    """Band Cholesky, diagonal (width 1) and inner-CG preconditioners all solve
    the same dense normal equations (residual against the brute-force A, b).
    The solutions themselves may differ along near-null directions of these
    tiny problems, so the residual is what is pinned.  The Fourier
    preconditioner (use_fprecond) removes each segment's mean
    (src/noise_routines.f90:977-993, 816-857), so it cannot move per-segment
    offsets and is only required to run."""
    kinds = []
    for extra in ({}, dict(precond_width_min=1, precond_width_max=1), dict(use_cgprecond=True)):
        case = small_case(prior=True, extra=extra)
        r, _ = run(case)
        kinds.append((r["precond_kind"], r["nchol"] > 0))
        A, b, _ = mc.dense_problem(case, r, prior_kernel_for(case, r)[0])
        a = r["aa"].reshape(-1)
        assert np.linalg.norm(b - A @ a) / np.linalg.norm(b) < 1e-10
        np.testing.assert_allclose(A, A.T, rtol=0, atol=1e-9 * np.abs(A).max())
    assert kinds == [(3, True), (1, False), (3, False)]
    r, _ = run(small_case(prior=True, extra=dict(use_fprecond=True, iter_max=20)))
    assert r["precond_kind"] == 2 and r["niter"] == 20


def test_multiple_periods_and_flags():
    case = small_case(prior=True, nperiod=3)
    case["pixels"][100:160] = -1            # a gap inside one period
    case["pixweights"][3 * 900:3 * 905] = 0   # zero-weight flags
    r, c = run(case)
    assert r["niter"] > 0
    assert np.all(c["pixels"][100:160] == 12 * 4 ** 2)
    pk = prior_kernel_for(case, r)
    nb = r["nbase"]
    # segments per (det, period) from the oracle's nna
    segs = []
    pb = np.searchsorted(r["base_start"], case["periods"])
    for d in range(len(case["dets"])):
        sd = []
        for ip in range(len(pb)):
            k0, k1 = pb[ip], (pb[ip + 1] if ip + 1 < len(pb) else nb)
            nz = np.nonzero(r["nna"][d, k0:k1])[0]
            if nz.size:
                sd.append((int(k0 + nz[0]), int(nz[-1] - nz[0] + 1)))
        segs.append(sd)
    flagged = mc.copy_case(case)
    flagged["pixels"] = np.where(c["pixels"] >= 12 * 4 ** 2, -1, c["pixels"]).astype(np.int32)
    A, b, _ = mc.dense_problem(flagged, r,
                               dict(segments=segs, block=lambda d, ks, no: mc.prior_block(pk[1], no)))
    a = r["aa"].reshape(-1)
    assert np.linalg.norm(b - A @ a) / np.linalg.norm(b) < 1e-9


# --------------------------------------------------------------- building blocks


@pytest.mark.parametrize("n", [4, 8, 64, 1024, 2 ** 16])
def test_fft_vs_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    X = oracle.rfft(x)
    np.testing.assert_allclose(X, np.fft.rfft(x), rtol=0, atol=1e-12 * np.sqrt(n))
    y = oracle.irfft_unnorm(X, n) / n
    np.testing.assert_allclose(y, x, rtol=0, atol=1e-13)
    # c2r ignores the imaginary parts of X_0 and X_{n/2} (FFTW semantics)
    X2 = X.copy()
    X2[0] += 3j
    X2[-1] -= 2j
    np.testing.assert_allclose(oracle.irfft_unnorm(X2, n) / n, x, rtol=0, atol=1e-13)


def test_invert_eig_cases():
    rng = np.random.default_rng(0)
    for _ in range(50):
        a = rng.standard_normal((3, 3))
        c = a @ a.T + 0.1 * np.eye(3)
        up = np.triu(c)
        inv, cdet, rc, sing = oracle.invert_eig(up)
        ev = np.linalg.eigvalsh(c)
        assert not sing
        np.testing.assert_allclose(inv, np.linalg.inv(c), rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(rc, ev.min() / ev.max(), rtol=1e-10)
        np.testing.assert_allclose(cdet, np.prod(ev), rtol=1e-10)
    # isotropic angles 0/45/90/135 deg -> sum w w^T = diag(4, 2, 2) -> rcond 1/2
    psi = np.deg2rad([0, 45, 90, 135])
    w = np.stack([np.ones(4), np.cos(2 * psi), np.sin(2 * psi)], 1)
    inv, cdet, rc, sing = oracle.invert_eig(np.triu(w.T @ w))
    assert abs(rc - 0.5) < 1e-15 and not sing
    # one hit, one angle -> singular, rejected
    inv, cdet, rc, sing = oracle.invert_eig(np.triu(np.outer(w[0], w[0])))
    assert sing and rc == 0 and np.all(inv == 0)


def test_band_cholesky_vs_dense():
    rng = np.random.default_rng(1)
    n, nband = 300, 7
    c = np.zeros(nband)
    c[0] = 10.0
    c[1:] = rng.standard_normal(nband - 1) * 0.5
    nna = 1 + rng.random(n)
    band = np.tile(c, (n, 1))
    band[:, 0] += nna
    dense = np.zeros((n, n))
    for m in range(nband):
        idx = np.arange(n - m)
        dense[idx + m, idx] = c[m]
        dense[idx, idx + m] = c[m]
    dense[np.arange(n), np.arange(n)] += nna
    fac, info = oracle.band_cholesky(band)
    assert info == 0
    L = np.linalg.cholesky(dense)
    for m in range(nband):
        np.testing.assert_allclose(fac[:n - m, m], np.diag(L, -m), rtol=1e-12, atol=1e-12)
    rhs = rng.standard_normal(n)
    np.testing.assert_allclose(oracle.band_solve(fac, rhs), np.linalg.solve(dense, rhs), rtol=1e-11)
    # not positive definite -> info = failing column (1-based)
    bad = band.copy()
    bad[5, 0] = -1.0
    assert oracle.band_cholesky(bad)[1] > 0


def test_interpolate_psd_loglog():
    f = np.array([0.0, 0.1, 1.0, 10.0])
    p = np.array([5.0, 100.0, 10.0, 1.0])    # a power law 10/f between 0.1 and 10
    t = np.array([0.0, 0.2, 0.5, 3.0, 20.0])
    out = oracle.interpolate_psd(f, p, t)
    np.testing.assert_allclose(out[1:], 10.0 / t[1:], rtol=1e-13)
    assert out[0] == 0.0


def test_filter_closed_form():
    """fcov = 1 / (fa * aliased(PSD - plateau) * sinc^2), fcov[0] = 0."""
    fs, nshort, nof = 10.0, 5, 64
    fa = fs / nshort
    freqs = np.logspace(-4, np.log10(fs / 2), 400)
    sig = 1e-3
    psd = sig ** 2 / fs * (1 + (freqs / 0.5) ** -2.0)
    fcov = oracle.build_filter(fs, nshort, nof, sig, freqs, psd)
    plateau = oracle.interpolate_psd(freqs, psd, np.array([fs / 2]))[0]
    asp = np.zeros(nof // 2 + 1)
    for k in range(3):
        f = k * fa + np.arange(nof) * fa / nof
        s = oracle.interpolate_psd(freqs, psd, f)
        margin = np.float64(np.float32(1e-6))
        s = np.where(s > plateau * (1 + margin), s - plateau, s * margin)
        if k == 0:
            s[0] = 0.0
        x = np.pi * f / fa
        g = np.where(x < np.float32(1e-30), 1.0, np.sin(x) ** 2 / np.where(x == 0, 1, x) ** 2)
        if k == 0:
            asp[0] = s[0]
        else:
            asp[0] += 2 * s[0] * g[0]
        i = np.arange(1, nof // 2 + 1)
        asp[i] += s[i] * g[i] + s[nof - i] * g[nof - i]
    exp = 1.0 / (fa * asp)
    exp[0] = 0.0
    np.testing.assert_allclose(fcov, exp, rtol=1e-12)


def test_filter_against_ou_closed_form():
    """build_filter (src/noise_routines.f90:367-490) pinned without its alias
    loop.  For an Ornstein-Uhlenbeck process (two-sided PSD A / (1 + (f/f0)^2)
    on the white plateau) the covariance of baseline means (boxcars of length
    T = 1 / fa) is closed form, C(0) = A pi f0 2 (b - 1 + e^-b) / b^2 and
    C(l) = A pi f0 (2 cosh b - 2) / b^2 e^(-b |l|) with b = 2 pi f0 T, and so
    is its transform Chat(w) = sum_l C(l) e^(-i w l).  By Poisson summation
    Chat / fa is the sum over every alias m of (PSD - plateau) sinc^2; the
    filter keeps the aliases m = -3 .. 2 (k = 0..2 and their mirrors), so
    fcov = 1 / (fa * (Chat / fa - tail - c (1 - tail_sinc))) with the excluded
    tails summed from the analytic PSD and c the plateau's Lorentzian part
    (sum_m sinc^2 = 1).  The table is log-spaced; its log-log interpolation
    error is ~2e-6, far below the k = 2 aliases' share (up to 1.6e-3), so a
    dropped or misplaced alias term, sinc^2 or plateau fails by 100x.
    The time-domain statement: circulant(irfft(fcov)) inverts the Toeplitz
    covariance of the baseline means up to the truncation bound (7.2e-4 for
    this PSD; 1 - 1/nof on the diagonal: the mean mode is left free)."""
    fs, nshort, nof = 100.0, 50, 256
    fa, f0, sigma = fs / nshort, 0.2, 1e-3
    plateau = sigma ** 2 / fs
    A = 100 * plateau
    freqs = np.logspace(-5, 2, 4001)
    lor = lambda f: A / (1 + (f / f0) ** 2)
    psd = plateau + lor(freqs)
    fcov = oracle.build_filter(fs, nshort, nof, sigma, freqs, psd)
    c = oracle.interpolate_psd(freqs, psd, np.array([fs / 2]))[0] - plateau   # plateau at fs / 2
    b = 2 * np.pi * f0 / fa
    C0 = A * np.pi * f0 * 2 * (b - 1 + np.exp(-b)) / b ** 2
    K = A * np.pi * f0 * (2 * np.cosh(b) - 2) / b ** 2
    nu = np.arange(1, nof // 2 + 1) / nof
    z = np.exp(-b - 2j * np.pi * nu)
    chat = C0 + 2 * K * np.real(z / (1 - z))
    sinc2 = lambda x: np.sin(np.pi * x) ** 2 / (np.pi * x) ** 2
    S = lambda x: lor(fa * np.abs(x)) * sinc2(x)
    j = np.arange(3, 200000)[:, None]
    tail = (S(j + nu) + S(j + 1 - nu)).sum(0)
    tail_s = (sinc2(j + nu) + sinc2(j + 1 - nu)).sum(0)
    asp = chat / fa - tail - c * (1 - tail_s)
    assert fcov[0] == 0
    np.testing.assert_allclose(fcov[1:], 1 / (fa * asp), rtol=1e-5)
    # the k = 2 aliases matter at the 1e-3 level: the check has the power to see them
    k2 = (S(2 + nu) + S(3 - nu) - c * (sinc2(2 + nu) + sinc2(3 - nu))) / asp
    assert np.abs(k2).max() > 100 * 1e-5
    # time domain: circulant(h) C_per = I - 1/nof, up to the truncation bound
    bound = np.abs((tail - c * tail_s) / asp).max()
    assert bound < 1e-3
    h = np.fft.irfft(fcov, nof)
    l = np.arange(nof)
    cl = lambda l: np.where(l == 0, C0, K * np.exp(-b * np.abs(l)))
    cper = cl(l) + cl(nof - l) * (l > 0)
    conv = np.real(np.fft.ifft(np.fft.fft(h) * np.fft.fft(cper)))
    err = conv - (np.eye(1, nof)[0] - 1 / nof)
    assert np.abs(err).max() <= bound, (np.abs(err).max(), bound)


def _multi_psd_truth(case, r, perturb=False):
    """Per-sample weights and per-interval prior blocks from the PSD each
    interval's first sample falls in (src/noise_routines.f90:193-236), with
    noise weights from the PSD plateau at fs / 2 (w = 1 / (plateau fs),
    src/inputparam.f90:120-140, src/parameter_control.f90:196-204).  perturb:
    PSD 0 everywhere (the mistake the pin must catch)."""
    pars = case["pars"]
    fs = pars["fsample"]
    nshort = int(round(pars["base_first"] * fs))
    ndet = len(case["dets"])
    nsamp = case["timestamps"].size
    freqs = case["psdfreqs"]
    nbin = freqs.size
    rows = case["psdvals"].reshape(-1, nbin)
    starts = case["psdstarts"].reshape(ndet, 2)
    t = case["timestamps"]
    per = list(case["periods"]) + [nsamp]
    pb = np.searchsorted(r["base_start"], case["periods"])
    nb = r["nbase"]
    wsamp = np.zeros((ndet, nsamp))
    blocks = {}
    segs = []
    for d in range(ndet):
        sd = []
        for i in range(len(per) - 1):
            # weights: psd_index_det, the last PSD with start <= t; filter:
            # psd_index, the first PSD with start <= t <= its stop (the next
            # start), so at a start exactly on t the filter keeps the old PSD
            jw = 0 if perturb else int(np.searchsorted(starts[d], t[per[i]], side="right") - 1)
            jf = 0 if perturb else int(np.searchsorted(starts[d][1:], t[per[i]], side="left"))
            pl = oracle.interpolate_psd(freqs, rows[2 * d + jw], np.array([fs / 2]))[0]
            wsamp[d, per[i]:per[i + 1]] = 1.0 / (pl * fs)
            psd = rows[2 * d + jf]
            plf = oracle.interpolate_psd(freqs, psd, np.array([fs / 2]))[0]
            fcov = oracle.build_filter(fs, nshort, r["nof"], np.sqrt(plf * fs), freqs, psd)
            k0, k1 = pb[i], (pb[i + 1] if i + 1 < len(pb) else nb)
            blocks[(d, int(k0))] = np.fft.irfft(fcov, r["nof"])
            sd.append((int(k0), int(k1 - k0)))
        segs.append(sd)
    pk = dict(segments=segs, block=lambda d, ks, no: mc.prior_block(blocks[(d, ks)], no))
    return wsamp, pk


def test_multi_psd_selection_dense():
    """npsd = 2 (src/noise_routines.f90:193-236): detector 0 switches PSD
    inside period 1, detector 1 exactly at the start of period 2, where the
    reference's two lookups part: the weights take the new PSD and the filter
    the old one (its stop time is inclusive).  Noise weights come from each
    PSD.  The oracle's solution satisfies the dense normal equations built
    with each interval's PSDs (weights and prior block), and does not satisfy
    the ones built with PSD 0 throughout."""
    case = small_case(prior=True, nperiod=4, ndet=2, extra=dict(noise_weights_from_psd=True))
    t, per = case["timestamps"], case["periods"]
    mc.with_psd_switch(case, [t[per[1] + 150], t[per[2]]])
    r, _ = run(case)
    assert r["niter"] > 3
    a = r["aa"].reshape(-1)
    for perturb in (False, True):
        wsamp, pk = _multi_psd_truth(case, r, perturb)
        A, b, _ = mc.dense_problem(case, r, pk, wsamp=wsamp)
        res = np.linalg.norm(b - A @ a) / np.linalg.norm(b)
        brel = np.linalg.norm(r["yba"].reshape(-1) - b) / np.linalg.norm(b)
        print("perturb", perturb, "residual", res, "rhs", brel)
        if not perturb:
            assert res < 1e-10 and brel < 1e-12
        else:
            assert res > 1e-3 or brel > 1e-3


def test_sparse_period_flagged():
    """A detector-period with fewer than 0.1 % valid samples is flagged as a
    whole (src/read_data.f90:241-272): here 3 of det 1's 4000 samples in
    period 1 stay valid (0.075 %), and they come back as the dummy pixel and
    drop out of the hits; 5 valid samples (0.125 %) are kept."""
    for nvalid, flagged in ((3, True), (5, False)):
        case = small_case(prior=False, nperiod=4, nsamp=16000)
        pix = case["pixels"].reshape(4, -1)
        pix[1, 4000 + nvalid:8000] = -1
        r, c = run(case)
        cp = c["pixels"].reshape(4, -1)
        assert np.all(cp[1, 4000 + nvalid:8000] == 12 * 4 ** 2)
        assert np.all(cp[1, 4000:4000 + nvalid] == 12 * 4 ** 2) == flagged
        assert r["hits"].sum() == 4 * 16000 - 4000 + (0 if flagged else nvalid)


@pytest.mark.parametrize("nof,noba", [(64, 30), (4096, 2000), (2 ** 16, 32769)])
def test_convolve_segment_vs_circulant(nof, noba):
    """One segment's convolution (src/noise_routines.f90:816-857): mean
    removed, zero-padded to nof, circular convolution with invcov =
    irfft(fcov), mean removed again.  Against the explicit circulant product
    for the small case and numpy's FFT for the large ones."""
    fs = 50.0
    f = np.logspace(-5, np.log10(fs / 2), 100)
    psd = mc.psd_model(f, 1e-4, fs, 0.1, -1.0)
    fc = oracle.build_filter(fs, 1, nof, 1e-4, f, psd)
    x = np.random.default_rng(nof).standard_normal(noba)
    y = oracle.convolve(fc, x, nof)
    h = np.fft.irfft(fc, nof)
    xm = np.zeros(nof)
    xm[:noba] = x - x.mean()
    if nof <= 64:
        idx = (np.arange(noba)[:, None] - np.arange(noba)[None, :]) % nof
        ref = h[idx] @ (x - x.mean())
    else:
        ref = np.fft.irfft(np.fft.rfft(h) * np.fft.rfft(xm), nof)[:noba]
    ref -= ref.mean()
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
