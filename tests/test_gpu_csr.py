"""GPU: the grouping of the samples by map pixel (the hand-written stable radix
sort, the hit counts from its runs, the CSR fill) and the baseline slot order,
checked exactly against numpy's stable sort of the same keys.

The reference visits the samples of a pixel in time order inside each
thread's pixel range (src/smadam_routines.f90:111-422); the CSR must give every
pixel its samples in ascending sample index, so that every per-pixel FP64 sum
has one fixed order.  The sort's passes (one per 8 bits of the pixel index),
its blocks (several tiles each above 2^24 samples) and flagged samples (which
sort after every pixel) are all exercised.
"""
import numpy as np
import pytest

from tests import madam_cases as mc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _torch_gpu_first():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except ImportError:
        pass
    yield


SCAN = dict(fsample=50.0, spin_period=2.0, prec_period=400.0, opening_turns=85.0 / 360.0, lon0_turns=0.0)


def expected(pix, ndet, nsamp, npix, L):
    """CSR and slot order from numpy's stable sorts; pix [ndet][nsamp], -1 flagged;
    one period of nsamp / L baselines of L samples"""
    keys = np.where(pix >= 0, pix, npix).reshape(-1).astype(np.int64)
    order = np.argsort(keys, kind="stable")
    hits = np.bincount(keys[keys < npix], minlength=npix)
    off = np.concatenate([[0], np.cumsum(hits)])
    total = int(off[-1])
    nb = nsamp // L
    k = np.repeat(np.arange(nb), ndet)        # the slot sort's input order: (baseline, detector)
    det = np.tile(np.arange(ndet), nb)
    pm = pix[det, k * L + (L - 1) // 2]
    cell = np.where(pm >= 0, pm, npix)
    bord = (det * nb + k)[np.argsort(cell, kind="stable")]
    return off, order[:total].astype(np.uint32), bord.astype(np.int32), hits


def run_case(ndet, nsamp, nside, flag_frac, seed):
    import libmadam_amd
    rng = np.random.default_rng(seed)
    dets = [dict(psi_turns=(0.0, 0.125, 0.25, 0.375)[d % 4], open_offset_turns=0.004 * (d // 4))
            for d in range(ndet)]
    pix, w = libmadam_amd.expand_pointing(SCAN, dets, nside, 0, nsamp)
    pix = pix.copy()
    if flag_frac > 0:
        pix[rng.random(pix.shape) < flag_frac] = -1
    npix = 12 * nside * nside
    fs, L = 50.0, 50
    freqs = np.logspace(-5, np.log10(fs / 2), 50)
    psd = mc.psd_model(freqs, 1e-4, fs, 0.1, -1.0)
    pars = dict(base_first=L / fs, fsample=fs, nside_map=nside, nside_cross=nside, nside_submap=max(1, nside // 4),
                kfilter=False, info=0)
    m = libmadam_amd.Madam(0)
    try:
        m.configure(pars, ["d%d" % d for d in range(ndet)], np.full(ndet, 1e8), np.arange(nsamp) / fs,
                    np.array([0], np.int64), np.ones(ndet, np.int64), np.zeros(ndet), freqs, np.tile(psd, ndet), 3)
        m.set_pointing(pix.reshape(-1).copy(), w.reshape(-1).copy())
        assert m.sizes()["nbase"] == nsamp // L
        got = m.csr()
        hits = np.zeros(npix, np.int32)
        libmadam_amd.lib().madam_get_hits(m._h, libmadam_amd._ptr(hits))
    finally:
        m.close()
    off, src, bord, ehits = expected(pix, ndet, nsamp, npix, L)
    assert np.array_equal(hits, ehits)
    assert np.array_equal(got["off"], off)
    assert np.array_equal(got["src"], src)
    assert np.array_equal(got["bord"], bord)


@pytest.mark.parametrize("nside,flag", [(4, 0.0), (16, 0.05), (256, 0.01)])
def test_csr_small(nside, flag):
    """one, two and three sort passes (8, 12 and 20 key bits), with and without
    flagged samples, partial last tiles"""
    run_case(3, 40000 + 50 * 7, nside, flag, seed=nside)


def test_csr_several_tiles_per_block():
    """2 x (2^24 - 16) samples: above 2^24 elements each sort block walks several tiles"""
    run_case(2, 50 * 335544, 512, 0.002, seed=5)
