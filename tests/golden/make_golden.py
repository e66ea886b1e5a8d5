"""Regenerate the known-answer fixtures of the reference's own Python tests.

The reference tests (python/tests/test_libmadam.py:41-174 and
python/tests/test_libmadam_mc.py:41-207) build their inputs with NumPy's
legacy MT19937 stream and check the hit map (bincount of pixels) and the
binned map (per-pixel mean of the signal) against the library's FITS output.
This script rebuilds those exact inputs and the brute-force expected outputs
and stores them as small .npz fixtures (data only).  Run:

    python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def reference_case(seed, second_signal=False):
    np.random.seed(seed)
    nside = 8
    npix = 12 * nside ** 2
    nsamp = 1000
    ndet = 4
    pixels = (np.arange(ndet * nsamp) % npix).astype(np.int32)
    signal = pixels.astype(np.float64)
    signal = signal + np.random.randn(nsamp * ndet)
    out = dict(pixels=pixels, signal=signal)
    if second_signal:
        signal2 = (pixels % 12).astype(np.float64)
        signal2 = signal2 + np.random.randn(nsamp * ndet)
        out["signal2"] = signal2
    hmap = np.bincount(pixels, minlength=npix).astype(np.int64)
    bmap = np.bincount(pixels, weights=signal, minlength=npix)
    good = hmap != 0
    bmap[good] /= hmap[good]
    out["hmap"] = hmap.astype(np.int32)
    out["bmap"] = bmap
    out["bmap32"] = bmap.astype(np.float32)
    return out


def main():
    a = reference_case(123456)
    np.savez(os.path.join(HERE, "ref_test_libmadam.npz"), **a)
    b = reference_case(9876543, second_signal=True)
    np.savez(os.path.join(HERE, "ref_test_libmadam_mc.npz"), **b)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
