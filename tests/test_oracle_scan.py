"""Pins for the CPU pointing expansion (row P0, oracle/oracle_scan.c).

HEALPix closed forms (Gorski et al. 2005): pix2ang of the nested scheme is
written out here independently (ring index, ring phi) and ang2pix must invert
it at every pixel centre; nside=1 gives the 12 base pixels; random points
land in a pixel whose centre is within the maximum pixel radius; equal-area
pixels get equal counts for uniform points.  The +,-,*,/ sin/cos/atan used by
the expansion are checked against libm.
"""
import numpy as np
import pytest

import oracle

JRLL = np.array([2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4])
JPLL = np.array([1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7])


def compress_bits(v):
    r = np.zeros_like(v)
    for b in range(20):
        r |= ((v >> (2 * b)) & 1) << b
    return r


def pix2vec_nest(nside, pix):
    pix = np.asarray(pix, np.int64)
    npface = nside * nside
    face = pix // npface
    ipf = pix % npface
    ix = compress_bits(ipf)
    iy = compress_bits(ipf >> 1)
    jr = JRLL[face] * nside - ix - iy - 1
    z = np.empty(pix.shape)
    nr = np.empty(pix.shape, np.int64)
    kshift = np.zeros(pix.shape, np.int64)
    north = jr < nside
    south = jr > 3 * nside
    eq = ~(north | south)
    nr[north] = jr[north]
    z[north] = 1 - nr[north] ** 2 / (3.0 * npface)
    nr[south] = 4 * nside - jr[south]
    z[south] = nr[south] ** 2 / (3.0 * npface) - 1
    nr[eq] = nside
    z[eq] = (2 * nside - jr[eq]) * 2.0 / (3.0 * nside)
    kshift[eq] = (jr[eq] - nside) & 1
    jp = (JPLL[face] * nr + ix - iy + 1 + kshift) // 2
    jp = np.where(jp > 4 * nside, jp - 4 * nside, jp)
    jp = np.where(jp < 1, jp + 4 * nside, jp)
    phi = (jp - (kshift + 1) * 0.5) * (np.pi / 2) / nr
    st = np.sqrt(np.maximum(0.0, 1 - z * z))
    return np.stack([st * np.cos(phi), st * np.sin(phi), z], -1)


@pytest.mark.parametrize("nside", [1, 2, 4, 16, 64, 512])
def test_pixel_centre_round_trip(nside):
    npix = 12 * nside * nside
    pix = np.arange(npix) if npix <= 50000 else np.random.default_rng(0).integers(0, npix, 50000)
    v = pix2vec_nest(nside, pix)
    got = np.array([oracle.vec2pix_nest(nside, *x) for x in v])
    assert np.array_equal(got, pix)


def test_base_pixels_nside1():
    # face centres: north faces at z = 2/3, phi = pi/4 + k pi/2; equator z=0, phi = k pi/2;
    # south faces z = -2/3
    for k in range(4):
        phi = np.pi / 4 + k * np.pi / 2
        st = np.sqrt(1 - (2 / 3) ** 2)
        assert oracle.vec2pix_nest(1, st * np.cos(phi), st * np.sin(phi), 2 / 3) == k
        assert oracle.vec2pix_nest(1, st * np.cos(phi), st * np.sin(phi), -2 / 3) == 8 + k
        phi = k * np.pi / 2
        assert oracle.vec2pix_nest(1, np.cos(phi), np.sin(phi), 0.0) == 4 + k
    # poles
    assert oracle.vec2pix_nest(1, 0.0, 0.0, 1.0) in range(4)
    assert oracle.vec2pix_nest(1, 0.0, 0.0, -1.0) in range(8, 12)


def test_random_points_within_pixel_radius_and_equal_area():
    nside = 16
    rng = np.random.default_rng(5)
    v = rng.standard_normal((40000, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    pix = np.array([oracle.vec2pix_nest(nside, *x) for x in v])
    c = pix2vec_nest(nside, pix)
    ang = np.arccos(np.clip(np.sum(c * v, 1), -1, 1))
    assert ang.max() < 1.5 * np.sqrt(np.pi / 3) / nside
    counts = np.bincount(pix // (nside * nside), minlength=12)
    assert counts.min() > 0.9 * counts.mean()


def test_sincos_atan_against_libm():
    for u in np.concatenate([np.linspace(-3, 3, 2001), [0.125, 0.25, 0.5, 0.75, 1e-9]]):
        s, c = oracle.sincos_turns(u)
        f = u - np.round(u)            # exact reduction to [-1/2, 1/2] turns
        assert abs(s - np.sin(2 * np.pi * f)) < 6e-16 and abs(c - np.cos(2 * np.pi * f)) < 6e-16
    rng = np.random.default_rng(1)
    for y, x in rng.standard_normal((5000, 2)):
        a = oracle.atan2_pos(y, x)
        e = np.arctan2(y, x) % (2 * np.pi)
        assert abs(a - e) < 2e-15   # 2 ulp at 2 pi


def test_scan_weights_unit_and_orientation():
    scan = dict(fsample=50.0, spin_period=60.0, prec_period=360 * 3600.0,
                opening_turns=85 / 360, lon0_turns=0.0)
    pix0, w0 = oracle.expand_pointing(scan, dict(psi_turns=0.0), 64, 0, 3000)
    pix1, w1 = oracle.expand_pointing(scan, dict(psi_turns=0.125), 64, 0, 3000)
    assert np.array_equal(pix0, pix1)       # same boresight
    np.testing.assert_allclose(w0[:, 1] ** 2 + w0[:, 2] ** 2, 1, atol=1e-6)
    assert np.all(w0[:, 0] == 1)
    # rotating the detector by 45 deg rotates (cos2psi, sin2psi) by 90 deg, with
    # one sense of rotation for every sample (here psi = psi_scan - psi0)
    np.testing.assert_allclose(w1[:, 1], w0[:, 2], atol=2e-6)
    np.testing.assert_allclose(w1[:, 2], -w0[:, 1], atol=2e-6)
    # a ring scan repeats its pixel sequence after one spin (small precession)
    pix2, _ = oracle.expand_pointing(scan, dict(psi_turns=0.0), 64, 3000, 3000)
    assert np.mean(pix2 == pix0) > 0.9
    # chunked expansion equals one-shot expansion
    pa, wa = oracle.expand_pointing(scan, dict(psi_turns=0.3), 64, 0, 2000)
    pb, wb = oracle.expand_pointing(scan, dict(psi_turns=0.3), 64, 1000, 1000)
    assert np.array_equal(pa[1000:], pb) and np.array_equal(wa[1000:], wb)
