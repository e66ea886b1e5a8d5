"""The entry point that allocates nothing (madam_expand_pointing_noalloc)
against madam_expand_pointing, through the C-ABI: bit-exact pixels and
weights on two calls in a row, the second with fewer detectors and samples
(the staging array keeps the first call's table), and the size limit."""
import ctypes as ct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scan():
    scan5 = np.array([100.0, 60.0, 360 * 3600.0, 85 / 360, 0.01], np.float64)
    return scan5


def _dets(ndet):
    return np.array([[0.125 * (j % 4), 0.001 * (j // 4), 0.0003 * j] for j in range(ndet)], np.float64)


def test_expand_pointing_noalloc_matches():
    import torch
    import libmadam_amd
    assert torch.cuda.is_available()
    torch.cuda.init()
    L = libmadam_amd.lib()
    f = L.madam_expand_pointing_noalloc
    f.restype = ct.c_int
    f.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_long, ct.c_long, ct.c_long, ct.c_void_p, ct.c_void_p,
                  ct.c_void_p]
    scan5 = _scan()
    for ndet, nside, i0, n in ((24, 1024, 0, 300000), (3, 256, 98765, 70001)):
        d3 = _dets(ndet)
        pix = torch.full((ndet, n), -7, dtype=torch.int32, device="cuda")
        w = torch.full((ndet, n, 3), -7.0, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        print("noalloc call: ndet %d nside %d n %d" % (ndet, nside, n), flush=True)
        rc = f(scan5.ctypes.data, d3.ctypes.data, ndet, nside, i0, n, pix.data_ptr(), w.data_ptr(), None)
        assert rc == 0, L.madam_last_error().decode()
        torch.cuda.synchronize()
        pix_ref = torch.empty_like(pix)
        w_ref = torch.empty_like(w)
        rc = L.madam_expand_pointing(scan5.ctypes.data_as(ct.c_void_p), d3.ctypes.data_as(ct.c_void_p), ndet, nside,
                                     i0, n, ct.c_void_p(pix_ref.data_ptr()), ct.c_void_p(w_ref.data_ptr()), 1)
        assert rc == 0, L.madam_last_error().decode()
        torch.cuda.synchronize()
        assert torch.equal(pix, pix_ref)
        assert torch.equal(w.view(torch.int32), w_ref.view(torch.int32))
    # sizes beyond the staging array are refused before any GPU work
    d3 = _dets(4097)
    rc = f(scan5.ctypes.data, d3.ctypes.data, 4097, 64, 0, 10, pix.data_ptr(), w.data_ptr(), None)
    assert rc == 1 and b"4096" in L.madam_last_error()
