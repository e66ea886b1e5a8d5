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
