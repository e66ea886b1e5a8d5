"""GPU parity: the HIP path (libmadam.so through its C-ABI) against the FP64
CPU oracle on the same seeded inputs.

Bars (stated here, checked below):
  * hits and flagged pixels: bit-exact;
  * pointing expansion (P0): bit-exact pixels and weights;
  * FP64 binned maps, destriped maps, baselines, RHS and nna: relative rms
    error <= 1e-10 against the oracle (FP64 accumulations in a different
    order; PCG iterates inherit those rounding differences).
"""
import numpy as np
import pytest

import oracle
from tests import madam_cases as mc

pytestmark = pytest.mark.gpu

RTOL_RMS = 1e-10


@pytest.fixture(scope="module", autouse=True)
def _torch_gpu_first():
    """In a process that uses both, torch's GPU runtime is initialised before
    libmadam.so's, the order bench.py uses (the reverse order leaves torch
    without a device); only the full-size test builds its inputs with torch."""
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except ImportError:
        pass
    yield


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.sqrt(np.mean(b ** 2))
    if den == 0:
        return float(np.sqrt(np.mean(a ** 2)))
    return float(np.sqrt(np.mean((a - b) ** 2)) / den)


def both(case, **kw):
    import libmadam_amd
    c1 = mc.copy_case(case)
    c2 = mc.copy_case(case)
    g = libmadam_amd.run(*mc.call_args(c1), **kw)
    o = oracle.destripe(*mc.call_args(c2), **kw)
    return g, o, c1, c2


def check_parity(g, o, c1, c2, maps=True, tol=RTOL_RMS):
    print("niter gpu/oracle", g["niter"], o["niter"], "stats", {k: g["stats"][k] for k in ("precond_kind", "nseg", "nband")})
    assert np.array_equal(g["hits"], o["hits"])
    assert np.array_equal(c1["pixels"], c2["pixels"])
    assert np.array_equal(g["mask"], o["mask"])
    errs = dict(binmap=rel_rms(g["binmap"], o["binmap"]), yba=rel_rms(g["yba"], o["yba"]),
                nna=rel_rms(g["nna"], o["nna"]), cc=rel_rms(g["cc"], o["cc"]))
    if maps:
        errs.update(map=rel_rms(g["map"], o["map"]), aa=rel_rms(g["aa"], o["aa"]),
                    signal=rel_rms(c1["signal"], c2["signal"]))
    print("relative rms errors", errs)
    for k, v in errs.items():
        assert v <= tol, (k, v)
    assert abs(g["niter"] - o["niter"]) <= 1


SCAN_SMALL = dict(fsample=50.0, spin_period=2.0, prec_period=400.0, opening_turns=85.0 / 360.0, lon0_turns=0.0)


def pointing_fn(nside, fs=50.0, spin=2.0, prec=400.0):
    scan = dict(SCAN_SMALL, fsample=fs, spin_period=spin, prec_period=prec)

    def f(d, n):
        return oracle.expand_pointing(scan, dict(psi_turns=(0.0, 0.125, 0.25, 0.375)[d % 4],
                                                 open_offset_turns=0.004 * (d // 4)), nside, 0, n)
    return f


def ring(nside=8, ndet=4, nsamp=20000, prior=False, nnz=3, nperiod=1, extra=None, base=1.0, seed=11, **kw):
    return mc.ring_case(ndet=ndet, nsamp=nsamp, nside=nside, fsample=50.0, base_first=base, nnz=nnz, seed=seed,
                        prior=prior, nperiod=nperiod, pointing=pointing_fn(nside), extra_pars=extra, **kw)


def test_pointing_expansion_bit_exact():
    import libmadam_amd
    scan = dict(fsample=100.0, spin_period=60.0, prec_period=360 * 3600.0, opening_turns=85 / 360, lon0_turns=0.01)
    dets = [dict(psi_turns=a, open_offset_turns=o, spin_offset_turns=s)
            for a, o, s in ((0.0, 0.0, 0.0), (0.125, 0.001, 0.0003), (0.25, -0.002, 0.0), (0.375, 0.0, -0.01))]
    for nside in (64, 512, 4096):
        n = 200000
        i0 = 123456789 if nside == 4096 else 0
        gp, gw = libmadam_amd.expand_pointing(scan, dets, nside, i0, n)
        for d, det in enumerate(dets):
            op, ow = oracle.expand_pointing(scan, det, nside, i0, n)
            assert np.array_equal(gp[d], op), (nside, d, np.sum(gp[d] != op))
            assert np.array_equal(gw[d].view(np.uint32), ow.view(np.uint32)), (nside, d)


def test_reference_known_answers_on_gpu():
    """python/tests/test_libmadam.py:155-174 known answers, through the GPU path."""
    import libmadam_amd
    case, gold = mc.reference_test_case()
    case = mc.copy_case(case)
    case["pars"] = dict(case["pars"], kfilter=False)
    g = libmadam_amd.run(*mc.call_args(case))
    assert np.array_equal(g["hits"], gold["hmap"])
    np.testing.assert_allclose(g["binmap"][:, 0], gold["bmap"], rtol=0, atol=1e-12)
    assert np.array_equal(g["binmap"][:, 0].astype(np.float32), gold["bmap32"])
    case1, _ = mc.reference_test_case()
    case1["pars"] = dict(case1["pars"], kfilter=False)
    h1 = libmadam_amd.run(*mc.call_args(mc.copy_case(case1)), subchunk_pass=1)["hits"]
    h2 = libmadam_amd.run(*mc.call_args(mc.copy_case(case1)), subchunk_pass=2)["hits"]
    assert np.array_equal(h1 + h2, g["hits"])


@pytest.mark.parametrize("nnz", [1, 3])
def test_no_prior_diagonal_preconditioner(nnz):
    g, o, c1, c2 = both(ring(nnz=nnz))
    assert g["stats"]["precond_kind"] == 1
    check_parity(g, o, c1, c2)


def test_noise_prior_band_preconditioner():
    g, o, c1, c2 = both(ring(prior=True, nsamp=30000, nperiod=2))
    assert g["stats"]["precond_kind"] == 3
    check_parity(g, o, c1, c2)


@pytest.mark.parametrize("flags", ["none", "head", "bursts"])
def test_band_factor_periodic_continuation(flags):
    """Where nna is constant, 64-row blocks of the band factor that repeat
    bitwise are copied instead of recomputed; the solve must be bitwise the
    one of the full column loop (kernel variant + 16).  "head" flags random
    samples of the first 40 % of every detector, so nna varies there and the
    repeat starts later; "bursts" flags 200-sample bursts every 30000 samples,
    so each segment has several constant runs, each with its own repeat."""
    import libmadam_amd
    case = ring(prior=True, ndet=4, nsamp=400000, nperiod=2)
    pix = case["pixels"].reshape(4, -1)
    if flags == "head":
        rng = np.random.default_rng(3)
        head = pix[:, :160000]
        head[rng.random(head.shape) < 0.05] = -1
    elif flags == "bursts":
        for b0 in range(10000, 400000, 30000):
            pix[:, b0:b0 + 200] = -1
    c1, c2 = mc.copy_case(case), mc.copy_case(case)
    g = libmadam_amd.run(*mc.call_args(c1))
    f = libmadam_amd.run(*mc.call_args(c2), kernel_variant=16)
    rows = f["stats"]["band_rows_factored"]
    print("band rows factored", g["stats"]["band_rows_factored"], "of", rows, "niter", g["niter"], f["niter"])
    assert g["stats"]["precond_kind"] == 3 and rows == 4 * 8000
    assert g["stats"]["band_rows_factored"] < rows
    assert g["niter"] == f["niter"]
    for k in ("aa", "map", "binmap", "yba"):
        assert np.array_equal(g[k], f[k]), k
    assert np.array_equal(c1["signal"], c2["signal"])


@pytest.mark.parametrize("width", [20, 30, 40, 50])
def test_band_preconditioner_wide(width):
    """Wider bands (kd classes 19, 29, 39 and 49: the factor's window no longer fits
    one 16-lane DPP row) against the oracle, and the periodic continuation of
    the factorization bitwise against the full column loop at that width."""
    import libmadam_amd
    case = ring(prior=True, ndet=4, nsamp=200000, nperiod=2,
                extra=dict(precond_width_min=width, precond_width_max=width))
    g, o, c1, c2 = both(case)
    assert g["stats"]["precond_kind"] == 3 and g["stats"]["nband"] == width
    check_parity(g, o, c1, c2)
    f = libmadam_amd.run(*mc.call_args(mc.copy_case(case)), kernel_variant=16)
    print("band rows factored", g["stats"]["band_rows_factored"], "of", f["stats"]["band_rows_factored"])
    assert g["stats"]["band_rows_factored"] < f["stats"]["band_rows_factored"]
    for k in ("aa", "map"):
        assert np.array_equal(g[k], f[k]), k


def test_repeated_solve_reuses_preconditioner():
    """A second solve on the same pointing with new TOD reuses the segments,
    filters and band preconditioner (nna is bitwise unchanged) and gives
    bitwise the results of a fresh context; a new pointing rebuilds them."""
    import libmadam_amd
    case = ring(prior=True, nsamp=30000, nperiod=2)
    case2 = ring(prior=True, nsamp=30000, nperiod=2, seed=12)
    fresh = libmadam_amd.run(*mc.call_args(mc.copy_case(case2)))
    a = mc.call_args(mc.copy_case(case))
    b = mc.call_args(mc.copy_case(case2))
    ndet = len(a[1])
    nnz = a[5].size // (a[3].size * ndet)
    m = libmadam_amd.Madam(0)
    try:
        m.configure(*a[:4], a[7], a[8], a[9], a[10], a[11], nnz, 0)
        m.set_pointing(a[4], a[5])
        m.solve(a[6])
        assert m.stats()["prec_reused"] == 0
        m.solve(b[6])
        st = m.stats()
        assert st["prec_reused"] == 1 and st["precond_kind"] == 3
        out = m.maps()
        out.update(m.baselines())
        for k in ("aa", "map", "binmap", "yba", "nna"):
            assert np.array_equal(out[k], fresh[k]), k
        assert int(st["niter"]) == fresh["niter"]
        m.set_pointing(b[4], b[5])
        m.solve(mc.copy_case(case2)["signal"])
        assert m.stats()["prec_reused"] == 0
    finally:
        m.close()


def test_plain_fp64_4det_2e20_nside64():
    """The plain FP64 destripe at full size: 4 detectors (0/45/90/135 deg) x
    2^20 samples at 50 Hz, ring scan with a 60 s spin and the spin axis
    advancing 1 deg/hour, Nside 64 IQU, 1 s baselines, no noise prior, PCG
    tolerance 1e-12, rcond cut 1e-3 on both pixel-matrix checks."""
    case = mc.ring_case(ndet=4, nsamp=2 ** 20, nside=64, fsample=50.0, base_first=1.0, nnz=3, seed=7, prior=False,
                        pointing=pointing_fn(64, fs=50.0, spin=60.0, prec=360 * 3600.0),
                        extra_pars=dict(cglimit=1e-12))
    g, o, c1, c2 = both(case)
    assert g["stats"]["precond_kind"] == 1
    check_parity(g, o, c1, c2)


def test_full_size_sky_and_offsets_recovered():
    """16 detectors x 2^24 samples at 100 Hz, Nside 512 IQU, 0.25 s baselines in
    two periods, no noise prior: a TOD of a known sky plus one random offset
    per baseline and no noise.  The destriping equations are then solved
    exactly by the sky and the offsets up to a common constant (the offsets'
    mean trades against the I monopole), so at full size the solve must
    return Q and U of the sky and I up to that constant, to the accuracy the
    1e-12 PCG tolerance leaves; the hits must add up to the samples.  That
    accuracy scales with the offsets over the sky and with the conditioning
    of the scan (the oracle leaves 3e-4 of Q and U at Nside 8 and 16 with
    1 mK offsets; this scan, with 2 uK offsets, must stay below 1e-2), and a
    wrong sign or index fails it by orders of magnitude."""
    import torch
    import libmadam_amd
    from bench import SCAN, SIGMA, detectors, psd_table
    ndet, nsamp, nside, fs, nbs = 16, 2 ** 24, 512, 100.0, 25
    npix, half = 12 * nside * nside, nsamp // 2
    pix = torch.empty(ndet, nsamp, dtype=torch.int32, device="cuda")
    w = torch.empty(ndet, nsamp, 3, dtype=torch.float32, device="cuda")
    libmadam_amd.expand_pointing_device(SCAN | dict(fsample=fs), detectors(ndet), nside, 0, nsamp, pix, w)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    sky = torch.randn(npix, 3, generator=g, device="cuda", dtype=torch.float64)
    sky *= torch.tensor([100e-6, 5e-6, 5e-6], device="cuda", dtype=torch.float64)
    nbp = (half + nbs - 1) // nbs   # baselines per period (the last one shorter)
    offs = torch.randn(ndet, 2 * nbp, generator=g, device="cuda", dtype=torch.float64) * 2e-6
    i = torch.arange(nsamp, device="cuda")
    k = (i // half) * nbp + (i % half) // nbs
    tod = torch.empty(ndet, nsamp, device="cuda", dtype=torch.float64)
    for d in range(ndet):
        tod[d] = (sky[pix[d].long()] * w[d].double()).sum(-1) + offs[d][k]
    torch.cuda.synchronize()
    pars = dict(base_first=0.25, fsample=fs, nside_map=nside, nside_cross=nside, nside_submap=16, kfilter=False,
                cglimit=1e-12, iter_max=1000, info=0, pixlim_map=1e-3, pixlim_cross=1e-3)
    m = libmadam_amd.Madam(0)
    try:
        freqs, psd = psd_table(fs)
        m.configure(pars, ["d%02d" % j for j in range(ndet)], np.full(ndet, 1.0 / SIGMA ** 2),
                    np.arange(nsamp) / fs, np.array([0, half], np.int64), np.ones(ndet, np.int64), np.zeros(ndet),
                    freqs, np.tile(psd, ndet), 3)
        m.set_pointing(pix, w, on_device=True)
        m.solve(tod, on_device=True)
        out = m.maps()
        st = m.stats()
    finally:
        m.close()
    skyh = sky.cpu().numpy()
    good = out["mask"] > 0
    print("niter", st["niter"], "rr", st["rr_final"], "good pixels", int(good.sum()))
    assert int(out["hits"].sum()) == ndet * nsamp
    assert st["rr_final"] < 1e-12 and good.sum() > 0.2 * npix
    d = out["map"][good] - skyh[good]
    d[:, 0] -= d[:, 0].mean()
    rel = np.sqrt(np.mean(d ** 2, axis=0)) / np.sqrt(np.mean(skyh[good] ** 2, axis=0))
    print("relative rms error I (less its mean), Q, U", rel)
    assert rel[0] < 1e-3 and rel[1] < 1e-2 and rel[2] < 1e-2, rel


def test_odd_baseline_count():
    """401 baselines (one detector, temperature only, with and without the
    noise prior): vector lengths that are not a multiple of the dot products'
    8-element chunks, of the 64-baseline gather blocks or of the band's 64-row
    chunks."""
    for prior in (False, True):
        g, o, c1, c2 = both(ring(ndet=1, nsamp=20050, nnz=1, prior=prior))
        assert g["yba"].size == 401
        check_parity(g, o, c1, c2)


def test_long_pixel_chains():
    """Nside 2: about 16,000 hits per pixel, so each pixel spans some 32 tiles of
    the pixel-side reduction and its partial sums are joined by the wave-wide
    path of the fix-up (chains longer than 16 tiles)."""
    g, o, c1, c2 = both(ring(nside=2, ndet=4, nsamp=200000))
    assert g["stats"]["npix_hit"] <= 48
    check_parity(g, o, c1, c2)


def test_pointing_record_layouts():
    """The projections read 12-byte records {index, w1, w2} when every
    intensity weight is exactly 1, 4-byte records for temperature-only maps,
    and the full 16-byte records otherwise (here a per-sample gain on all three
    weights); each layout against the oracle."""
    g, o, c1, c2 = both(ring(prior=True, nsamp=30000, nperiod=2))
    assert g["stats"]["rec_bytes"] == 12
    check_parity(g, o, c1, c2)
    g, o, c1, c2 = both(ring(prior=True, nsamp=30000, nperiod=2, nnz=1))
    assert g["stats"]["rec_bytes"] == 4
    check_parity(g, o, c1, c2)
    case = ring(prior=True, nsamp=30000, nperiod=2)
    w = case["pixweights"].reshape(-1, 3)
    gain = 1.0 + 0.05 * np.random.default_rng(5).standard_normal((w.shape[0], 1))
    w[:] = (w * gain).astype(np.float32)
    g, o, c1, c2 = both(case)
    assert g["stats"]["rec_bytes"] == 16
    check_parity(g, o, c1, c2)


@pytest.mark.parametrize("variant", [1, 2, 3])
def test_noise_prior_band_preconditioner_other_kernels(variant):
    """the alternative projection kernels: 1 wave per pixel / lane group per
    baseline, 2 LDS tiles on both sides, 3 segmented tiles on both sides"""
    import libmadam_amd
    case = ring(prior=True, nsamp=30000, nperiod=2)
    c1, c2 = mc.copy_case(case), mc.copy_case(case)
    g = libmadam_amd.run(*mc.call_args(c1), kernel_variant=variant)
    o = oracle.destripe(*mc.call_args(c2))
    check_parity(g, o, c1, c2)


def test_noise_prior_fourier_preconditioner():
    g, o, c1, c2 = both(ring(prior=True, nsamp=30000, nperiod=2, extra=dict(use_fprecond=True, iter_max=30)))
    assert g["stats"]["precond_kind"] == 2
    check_parity(g, o, c1, c2, tol=1e-8)


def test_flags_degrade_and_write_tod():
    case = ring(prior=True, nside=16, nsamp=40000, nperiod=3,
                extra=dict(nside_cross=8, write_tod=True, nside_submap=4))
    case["pixels"][500:900] = -1
    case["pixweights"][3 * 7000:3 * 7100] = 0
    g, o, c1, c2 = both(case)
    assert np.sum(c1["pixels"] == 12 * 16 ** 2) >= 500
    check_parity(g, o, c1, c2)


def test_subchunk_pass():
    case = ring(nsamp=24000, nperiod=2, extra=dict(nsubchunk=3))
    for sub in (1, 3):
        g, o, c1, c2 = both(case, subchunk_pass=sub)
        check_parity(g, o, c1, c2)


def test_wrapper_drop_in_mutates_like_reference(tmp_path):
    import os
    from libmadam_amd import libmadam_wrapper as madam
    os.environ["MADAM_RAISE"] = "1"
    case = ring(prior=True, nsamp=20000, extra=dict(write_tod=True, path_output=str(tmp_path)))
    case["pixels"][100:200] = -1
    c1 = mc.copy_case(case)
    c2 = mc.copy_case(case)
    madam.destripe(None, c1["pars"], c1["dets"], c1["weights"], c1["timestamps"], c1["pixels"], c1["pixweights"],
                   c1["signal"], c1["periods"], c1["npsd"], c1["psdstarts"], c1["psdfreqs"], c1["psdvals"])
    oracle.destripe(*mc.call_args(c2))
    assert np.array_equal(c1["pixels"], c2["pixels"])
    assert rel_rms(c1["signal"], c2["signal"]) < RTOL_RMS
    # unknown parameters are rejected
    c3 = mc.copy_case(case)
    with pytest.raises(Exception, match="Unknown parameter"):
        madam.destripe(None, dict(c3["pars"], bogus=1), c3["dets"], c3["weights"], c3["timestamps"], c3["pixels"],
                       c3["pixweights"], c3["signal"], c3["periods"], c3["npsd"], c3["psdstarts"], c3["psdfreqs"],
                       c3["psdvals"])


def test_reference_case_with_noise_filter():
    """The reference's own test configuration (kfilter=T, flat PSDs, nside_cross <
    nside_map, nsubchunk=2, surveys): the band criterion fails and the inner-CG
    preconditioner is used (src/noise_routines.f90:1113-1120, 1305-1409)."""
    case, gold = mc.reference_test_case()
    g, o, c1, c2 = both(case)
    assert g["stats"]["ncg"] > 0
    assert np.array_equal(g["hits"], gold["hmap"])
    check_parity(g, o, c1, c2)


def test_inner_cg_preconditioner_forced():
    g, o, c1, c2 = both(ring(prior=True, nsamp=20000, nperiod=2, extra=dict(use_cgprecond=True)))
    assert g["stats"]["ncg"] > 0 and g["stats"]["nband"] == 0
    check_parity(g, o, c1, c2)


def test_mc_mode_cache_cycle(tmp_path):
    """destripe(mcmode=T) -> destripe_with_cache(new signal) -> clear_caches, as in
    python/tests/test_libmadam_mc.py:132-191: the caller passes the same pix
    and weight arrays again (pix as the first call left them: its flagged
    samples hold the dummy pixel), so the cached call reuses the pointing
    set-up and, nna being bitwise unchanged, the filters and preconditioner.
    The cached solve equals a fresh oracle solve of the new signal (clean
    right-hand side)."""
    import os
    import libmadam_amd
    from libmadam_amd import libmadam_wrapper as madam
    os.environ["MADAM_RAISE"] = "1"
    case = ring(prior=True, nsamp=20000, extra=dict(mcmode=True, write_tod=True, path_output=str(tmp_path)))
    case["pixels"][300:420] = -1
    c1 = mc.copy_case(case)
    madam.destripe(None, c1["pars"], c1["dets"], c1["weights"], c1["timestamps"], c1["pixels"], c1["pixweights"],
                   c1["signal"], c1["periods"], c1["npsd"], c1["psdstarts"], c1["psdfreqs"], c1["psdvals"])
    assert np.all(c1["pixels"][300:420] == 12 * 8 ** 2)
    rng = np.random.default_rng(3)
    sig2 = case["signal"] + rng.standard_normal(case["signal"].size) * 1e-4
    c1["signal"][:] = sig2
    (tmp_path / "mc").mkdir()
    madam.destripe_with_cache(None, c1["timestamps"], c1["pixels"], c1["pixweights"], c1["signal"],
                              str(tmp_path / "mc"))
    st = libmadam_amd.cache_stats()
    assert st["pointing_reused"] == 1 and st["prec_reused"] == 1, st
    assert (tmp_path / "mc" / "madam_map.fits").exists()
    madam.clear_caches()
    madam.clear_caches()   # a second call only warns
    assert libmadam_amd.cache_stats() is None
    c3 = mc.copy_case(case)
    c3["signal"] = sig2.copy()
    c3["pars"] = dict(case["pars"], mcmode=False)
    c3["pars"].pop("path_output")
    oracle.destripe(*mc.call_args(c3))
    assert rel_rms(c1["signal"], c3["signal"]) < RTOL_RMS
    assert np.array_equal(c1["pixels"], c3["pixels"])


def test_mc_mode_new_pointing():
    """destripe_with_cache with a different pointing (a pixel moved): the
    pointing set-up is rebuilt, the dummy pixels of the caller's array still
    count as flags, and the result equals the oracle's solve of that input."""
    import os
    import libmadam_amd
    from libmadam_amd import libmadam_wrapper as madam
    os.environ["MADAM_RAISE"] = "1"
    case = ring(prior=True, nsamp=20000, extra=dict(mcmode=True, write_tod=True, write_map=False))
    case["pixels"][300:420] = -1
    c1 = mc.copy_case(case)
    madam.destripe(None, c1["pars"], c1["dets"], c1["weights"], c1["timestamps"], c1["pixels"], c1["pixweights"],
                   c1["signal"], c1["periods"], c1["npsd"], c1["psdstarts"], c1["psdfreqs"], c1["psdvals"])
    c1["pixels"][5000] = (c1["pixels"][5000] + 1) % (12 * 8 ** 2)
    c1["signal"][:] = case["signal"]
    moved = c1["pixels"].copy()
    try:
        madam.destripe_with_cache(None, c1["timestamps"], c1["pixels"], c1["pixweights"], c1["signal"], "")
        st = libmadam_amd.cache_stats()
        assert st["pointing_reused"] == 0, st
    finally:
        madam.clear_caches()
    c3 = mc.copy_case(case)
    c3["pixels"] = np.where(moved == 12 * 8 ** 2, -1, moved).astype(np.int32)
    c3["pars"] = dict(case["pars"], mcmode=False)
    oracle.destripe(*mc.call_args(c3))
    assert rel_rms(c1["signal"], c3["signal"]) < RTOL_RMS
    assert np.array_equal(c1["pixels"], c3["pixels"])


def test_bitwise_reproducible():
    """The pixel grouping is a stable sort, so two solves of the same input give
    identical bits (maps, baselines, iteration count)."""
    import libmadam_amd
    case = ring(prior=True, nsamp=30000, nperiod=2)
    g1 = libmadam_amd.run(*mc.call_args(mc.copy_case(case)))
    g2 = libmadam_amd.run(*mc.call_args(mc.copy_case(case)))
    assert g1["niter"] == g2["niter"]
    for k in ("map", "binmap", "aa", "yba", "cc"):
        assert np.array_equal(np.asarray(g1[k]).view(np.uint64), np.asarray(g2[k]).view(np.uint64)), k


def test_noise_prior_long_fft():
    """One-sample baselines: 70000 baselines per period, nof = 2^18, so the
    convolution runs with 256 x 512 four-step plans (the shape class of the
    2^20 production filters)."""
    case = ring(prior=True, ndet=2, nsamp=140000, nperiod=2, base=0.02, extra=dict(iter_max=20, use_fprecond=True))
    g, o, c1, c2 = both(case)
    assert g["stats"]["nof"] == 2 ** 18 and g["stats"]["precond_kind"] == 2
    check_parity(g, o, c1, c2, tol=1e-8)


def test_band_preconditioner_chunk_groups():
    """5000-baseline segments: 79 chunks of 64 rows in two chain groups per
    segment, so the grouped carry chain (group products, top-level chain,
    second local pass) is exercised."""
    g, o, c1, c2 = both(ring(prior=True, ndet=2, nsamp=500000, nperiod=2))
    assert g["stats"]["precond_kind"] == 3 and g["stats"]["nband"] >= 2
    check_parity(g, o, c1, c2)


# ------------------------------------------------------------------ long filters, multi-PSD, flags, wide bands


@pytest.mark.parametrize("log2nof,nsamp", [(20, 300000), (21, 600000), (22, 1100000)])
def test_prior_convolution_long_fft(log2nof, nsamp):
    """The noise-prior convolution at nof = 2^20 (512 x 1024 four-step plans,
    74 KB LDS column passes), 2^21 and 2^22 (1024-point columns, 147 KB of
    LDS): one detector, one-sample baselines, one period, so the segment is
    the whole period.
      * convolution: the GPU's result against the oracle's convolution of the
        same segment with the GPU's own filter, to FFT rounding (1e-13);
      * filter (build_filter): the GPU's against the oracle's, bin by bin, to
        1e-12 relative where the PSD is at least 1 % above the plateau.  Near
        fs / 2 the PSD approaches the plateau (measured there), PSD - plateau
        loses up to ten digits and the device's and the host's exp / log
        differ in the last bit: those bins agree to ~1e-6 relative, and the
        whole filter to 1e-8 rms."""
    import libmadam_amd
    case = ring(prior=True, ndet=1, nsamp=nsamp, nperiod=1, base=0.02, nnz=1,
                extra=dict(use_fprecond=True, iter_max=1))
    a = mc.call_args(mc.copy_case(case))
    m = libmadam_amd.Madam(0)
    try:
        m.configure(*a[:4], a[7], a[8], a[9], a[10], a[11], 1, 0)
        m.set_pointing(a[4], a[5])
        m.solve(a[6])
        st = m.stats()
        assert st["nof"] == 2 ** log2nof and st["nseg"] == 1
        x = np.random.default_rng(log2nof).standard_normal(nsamp)
        y = m.apply(0, x)
        fc_gpu = m.filter(0)
    finally:
        m.close()
    nof = 2 ** log2nof
    ref = oracle.convolve(fc_gpu, x, nof)
    err = rel_rms(y, ref)
    fs = case["pars"]["fsample"]
    freqs, psd = case["psdfreqs"], case["psdvals"][:case["psdfreqs"].size]
    fc = oracle.build_filter(fs, 1, nof, 1.0 / np.sqrt(case["weights"][0]), freqs, psd)
    fk = np.arange(nof // 2 + 1) * fs / nof
    pl = oracle.interpolate_psd(freqs, psd, np.array([fs / 2]))[0]
    well = oracle.interpolate_psd(freqs, psd, np.maximum(fk, 1e-9)) > 1.01 * pl
    well[0] = False
    ferr = np.abs(fc_gpu[well] / fc[well] - 1).max()
    print("nof 2^%d: convolution %.3g, filter %.3g on %d well-conditioned bins, %.3g rms over all"
          % (log2nof, err, ferr, well.sum(), rel_rms(fc_gpu, fc)))
    assert err <= 1e-13
    assert well.sum() > nof // 16 and ferr <= 1e-12
    assert rel_rms(fc_gpu, fc) <= 1e-8


@pytest.mark.timeout(900)
def test_config2_shape_noise_prior_band():
    """Two of the benchmark's detectors at its full per-detector shape: 2^24
    samples at 100 Hz in two periods of 2^23, 0.25 s baselines (335,545 per
    segment: 5,243 band chunks, 82 carry-chain groups), Nside 512 IQU, noise
    prior from the PSD (sigma 2e-4 K, fknee 0.05 Hz, alpha -1.5) with nof =
    2^20, band-Cholesky preconditioner, PCG tolerance 1e-12.  Full solve
    against the oracle: hits bit-exact, maps, baselines and TOD to 1e-10."""
    import libmadam_amd
    from bench import SCAN, SIGMA, FKNEE, ALPHA, psd_table, params
    ndet, nsamp, nside, fs = 2, 2 ** 24, 512, 100.0
    print("config-2 shape: 2 det x 2^24, pointing on the GPU", flush=True)
    pix, w = libmadam_amd.expand_pointing(SCAN | dict(fsample=fs), [dict(psi_turns=0.0), dict(psi_turns=0.125)],
                                          nside, 0, nsamp)
    rng = np.random.default_rng(5)
    sky = rng.standard_normal((12 * nside ** 2, 3)) * np.array([100e-6, 5e-6, 5e-6])
    sig = np.einsum("dsk,dsk->ds", sky[pix], w.astype(np.float64))
    for d in range(ndet):
        sig[d] += mc.one_over_f(rng, nsamp, fs, SIGMA, FKNEE, ALPHA)
    freqs, psd = psd_table(fs)
    # rcond cut 1e-3 on both pixel-matrix checks: the default map-level cut
    # (1e-6) keeps pixels whose inverse carries eps / rcond ~ 1e-10 of rounding
    pars = dict(params(nside, fs, 0.25), pixlim_map=1e-3, pixlim_cross=1e-3)
    case = dict(pars=pars, dets=["a", "b"], weights=np.full(ndet, 1 / SIGMA ** 2),
                timestamps=np.arange(nsamp) / fs, pixels=pix.reshape(-1), pixweights=w.reshape(-1),
                signal=sig.reshape(-1), periods=np.array([0, nsamp // 2], np.int64), npsd=np.ones(ndet, np.int64),
                psdstarts=np.zeros(ndet), psdfreqs=freqs, psdvals=np.tile(psd, ndet))
    case["pars"]["write_tod"] = True
    print("solving on the GPU and the oracle", flush=True)
    g, o, c1, c2 = both(case)
    assert g["stats"]["nof"] == 2 ** 20 and g["stats"]["precond_kind"] == 3 and g["stats"]["ncg"] == 0
    print("band chunks", g["stats"]["band_chunks"])
    assert 4 * 5200 <= g["stats"]["band_chunks"] <= 4 * 5243
    check_parity(g, o, c1, c2)


def test_multi_psd_switch():
    """npsd = 2 per detector with noise weights from each PSD: detector 0
    switches PSD inside period 1 after a 65 s gap (61 empty baselines end the
    period's first segment, so its second segment starts past the switch: the
    preconditioner takes the PSD at the segment start, the convolution the one
    at the period start); detector 1 switches exactly at the start of period
    2 (weights take the new PSD, the filter the old one)."""
    case = ring(prior=True, ndet=2, nsamp=45000, nperiod=3, extra=dict(noise_weights_from_psd=True))
    t, per = case["timestamps"], case["periods"]
    pix = case["pixels"].reshape(2, -1)
    pix[0, per[1] + 2000:per[1] + 2000 + 3250] = -1
    mc.with_psd_switch(case, [t[per[1] + 4000], t[per[2]]])
    g, o, c1, c2 = both(case)
    assert g["stats"]["precond_kind"] == 3 and g["stats"]["nseg"] == 7
    check_parity(g, o, c1, c2)


def test_sparse_period_flagged():
    """A detector-period with fewer than 0.1 % valid samples is flagged whole
    (src/read_data.f90:241-272); one with 0.125 % is kept."""
    for nvalid in (3, 5):
        case = ring(ndet=4, nsamp=16000, nperiod=4)
        pix = case["pixels"].reshape(4, -1)
        pix[1, 4000 + nvalid:8000] = -1
        g, o, c1, c2 = both(case)
        cp = c1["pixels"].reshape(4, -1)
        assert np.all(cp[1, 4000:4000 + nvalid] == 12 * 8 ** 2) == (nvalid == 3)
        check_parity(g, o, c1, c2)


def test_band_preconditioner_wider_than_kernels():
    """precond_width_min = 60: wider than the band kernels' widest class
    (50), so those segments use the inner-CG preconditioner.  With a
    different preconditioner the PCG takes another path through the
    operator's null space (a global offset traded between the baselines and
    the I map), so with both solves converged to 1e-26 the baselines and maps
    are compared less that offset; hits, masks, pixel matrices and the
    right-hand side are compared as everywhere."""
    case = ring(prior=True, ndet=2, nsamp=100000, nperiod=2,
                extra=dict(precond_width_min=60, precond_width_max=100, cglimit=1e-26, iter_max=2000))
    g, o, c1, c2 = both(case)
    print("niter gpu/oracle", g["niter"], o["niter"], "oracle band", o["nband_max"])
    assert g["stats"]["nwide"] > 0 and g["stats"]["ncg"] > 0 and o["nband_max"] >= 60
    g["niter"] = o["niter"]
    check_parity(g, o, c1, c2, maps=False)
    good = o["mask"] > 0
    da = rel_rms(g["aa"] - g["aa"].mean(), o["aa"] - o["aa"].mean())
    mg, mo = g["map"][good].copy(), o["map"][good].copy()
    mg[:, 0] -= mg[:, 0].mean()
    mo[:, 0] -= mo[:, 0].mean()
    dm = rel_rms(mg, mo)
    print("less the offset: baselines %.3g, map %.3g" % (da, dm))
    assert da <= 1e-9 and dm <= 1e-9
