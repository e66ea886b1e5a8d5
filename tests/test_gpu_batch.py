"""Batched and mixed-precision solves (madam_solve_batch) through the C-ABI.

Bars:
  * FP64 batch: every right-hand side equals its own single-RHS FP64 solve
    (libmadam.so's destripe path) to 1e-10 relative rms in maps, baselines
    and right-hand side, with the same iteration count (per-RHS exit tests);
    the HIP-graph replay of the PCG step is bitwise the eager sequence;
  * mixed precision (FP32 operator inside FP64 refinement): against the
    FP64 oracle, both converged to cglimit 1e-22, maps and baselines to
    1e-6 relative rms (max and rms errors printed), less the null-space
    offset (a constant traded between the baselines and the I map, which
    solvers taking different paths leave differently);
  * binned maps and yba of every right-hand side to 1e-10 (FP64 binning).
"""
import os

import numpy as np
import pytest

import oracle
from tests import madam_cases as mc
from tests.test_gpu_parity import ring, rel_rms

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


def signals_for(case, nrhs, seed=100, scale=1.0):
    """nrhs realisations on the case's pointing: its sky plus independent
    white + 1/f noise (seeded per right-hand side, different levels)."""
    ndet = len(case["dets"])
    nsamp = case["timestamps"].size
    fs = case["pars"]["fsample"]
    pix = case["pixels"].reshape(ndet, nsamp)
    w = case["pixweights"].reshape(ndet, nsamp, -1).astype(np.float64)
    sky = case["sky"]
    base = np.einsum("dsk,dsk->ds", sky[pix][:, :, :w.shape[2]], w) if w.shape[2] == 3 else sky[pix, 0]
    out = np.zeros((nrhs, ndet, nsamp))
    for k in range(nrhs):
        rng = np.random.default_rng(seed + k)
        out[k] = base
        for d in range(ndet):
            out[k, d] += mc.one_over_f(rng, nsamp, fs, 1e-4 * scale * (1 + 0.3 * k), 0.1, -1.0)
    return out


def single(case, signal):
    import libmadam_amd
    c = mc.copy_case(case)
    c["signal"] = signal.reshape(-1).copy()
    return libmadam_amd.run(*mc.call_args(c))


def batch(case, sig, precision):
    import libmadam_amd
    a = mc.call_args(mc.copy_case(case))
    ndet = len(a[1])
    nnz = a[5].size // (a[3].size * ndet)
    m = libmadam_amd.Madam(0)
    try:
        m.configure(*a[:4], a[7], a[8], a[9], a[10], a[11], nnz, 0)
        m.set_pointing(a[4], a[5])
        m.solve_batch(sig, precision=precision)
        return [m.batch_result(k) for k in range(sig.shape[0])], m.stats()
    finally:
        m.close()


@pytest.mark.parametrize("kind", ["band", "diag", "fprecond", "cg"])
def test_fp64_batch_equals_single_solves(kind):
    """Five right-hand sides (kernel width 8: three padding lanes) against five
    single solves, each with its own iteration count."""
    extra = dict(band={}, diag={}, fprecond=dict(use_fprecond=True, iter_max=40),
                 cg=dict(use_cgprecond=True))[kind]
    case = ring(prior=kind != "diag", nsamp=30000, nperiod=2, extra=extra)
    sig = signals_for(case, 5)
    res, st = batch(case, sig, 0)
    for k in range(5):
        s1 = single(case, sig[k])
        r = res[k]
        errs = dict(map=rel_rms(r["map"], s1["map"]), binmap=rel_rms(r["binmap"], s1["binmap"]),
                    aa=rel_rms(r["aa"], s1["aa"]), yba=rel_rms(r["yba"], s1["yba"]))
        print(kind, "rhs", k, "steps", r["info"]["steps"], "single", s1["niter"], "status", r["info"]["status"],
              "rr/rr0", r["info"]["rr_inner"], "single rr", s1["rr"][-3:], errs)
        if kind != "fprecond":
            # (the Fourier preconditioner cannot move per-segment offsets: these
            # solves stall, and one whose rz reaches exactly 0 in one summation
            # order and not in the other stops at another step, same results)
            assert abs(r["info"]["steps"] - s1["niter"]) <= 1
        for name, e in errs.items():
            assert e <= 1e-10, (name, e)


def test_fp64_batch_graph_replay_bitwise():
    """The PCG step captured into a HIP graph and replayed once per iteration
    (with the host's convergence reads between replays) gives bitwise the
    results of the same launches issued eagerly."""
    case = ring(prior=True, nsamp=30000, nperiod=2)
    sig = signals_for(case, 3)
    old = {k: os.environ.get(k) for k in ("MADAM_BATCH_GRAPH", "MADAM_BS_PIECE")}
    try:
        # the two solves run on separate contexts: pin the pixel-side cut,
        # which the autotuner may otherwise choose differently for each
        os.environ["MADAM_BS_PIECE"] = "256"
        os.environ["MADAM_BATCH_GRAPH"] = "1"
        g, _ = batch(case, sig, 0)
        os.environ["MADAM_BATCH_GRAPH"] = "0"
        e, _ = batch(case, sig, 0)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    for k in range(3):
        assert g[k]["info"]["steps"] == e[k]["info"]["steps"]
        for name in ("map", "aa", "binmap"):
            assert np.array_equal(g[k][name], e[k][name]), (k, name)


def offset_free(aa, m, good):
    mm = m[good].copy()
    mm[:, 0] -= mm[:, 0].mean()
    return aa - aa.mean(), mm


@pytest.mark.parametrize("tod32", [False, True])
def test_mixed_precision_against_oracle(tod32):
    """FP32 operator and vectors inside FP64 iterative refinement, four
    right-hand sides, against the FP64 oracle solve of each (both to cglimit
    1e-22).  With tod32 the signals are passed as float32 and the oracle
    solves the same float32 values."""
    case = ring(prior=True, nsamp=40000, nperiod=2, extra=dict(cglimit=1e-22, iter_max=2000))
    sig = signals_for(case, 4)
    if tod32:
        sig = sig.astype(np.float32)
    res, st = batch(case, sig, 1)
    for k in range(4):
        c = mc.copy_case(case)
        c["signal"] = sig[k].astype(np.float64).reshape(-1).copy()
        o = oracle.destripe(*mc.call_args(c))
        r = res[k]
        good = o["mask"] > 0
        ag, mg = offset_free(r["aa"], r["map"], good)
        ao, mo = offset_free(o["aa"], o["map"], good)
        e_map, e_aa = rel_rms(mg, mo), rel_rms(ag, ao)
        mx_map = np.abs(mg - mo).max() / np.abs(mo).max()
        print("rhs", k, "passes", r["info"]["passes"], "pcg steps", r["info"]["pcg_steps"], "oracle", o["niter"],
              "rr_final", r["info"]["rr_final"], "map rms %.3g max %.3g, baselines rms %.3g" % (e_map, mx_map, e_aa),
              "binmap %.3g" % rel_rms(r["binmap"], o["binmap"]))
        assert r["info"]["rr_final"] < 1e-22
        assert rel_rms(r["binmap"], o["binmap"]) <= 1e-10
        assert rel_rms(r["yba"], o["yba"]) <= 1e-10
        assert e_map <= 1e-6 and e_aa <= 1e-6


def test_mixed_precision_single_rhs_default_tolerance():
    """One right-hand side, the default cglimit 1e-12: the refinement stops
    on the FP64 residual's test, and the maps agree with the FP64 GPU solve to
    the accuracy that tolerance leaves (less the null-space offset)."""
    case = ring(prior=True, nsamp=40000, nperiod=2)
    sig = signals_for(case, 1)
    res, _ = batch(case, sig, 1)
    s1 = single(case, sig[0])
    r = res[0]
    good = s1["mask"] > 0
    ag, mg = offset_free(r["aa"], r["map"], good)
    a1, m1 = offset_free(s1["aa"], s1["map"], good)
    print("passes", r["info"]["passes"], "steps", r["info"]["pcg_steps"], "fp64 steps", s1["niter"],
          "rr_final", r["info"]["rr_final"], "map", rel_rms(mg, m1), "aa", rel_rms(ag, a1))
    assert r["info"]["rr_final"] < 1e-12
    assert rel_rms(mg, m1) <= 1e-4


@pytest.mark.parametrize("nnz", [3, 1])
def test_mixed_precision_run_length_records(nnz):
    """Six right-hand sides (kernel width 8), so that the FP32 operator reads
    run-length records: one record per run of samples of one baseline on one
    pixel, and per run of a pixel's CSR entries from one baseline, weights
    summed.  With flagged bursts, against the same mixed solve on the sample
    records (MADAM_RLE=0) and against the FP64 oracle for two right-hand
    sides, all converged to cglimit 1e-22, less the null-space offset."""
    case = ring(prior=True, nsamp=40000, nperiod=2, nnz=nnz, extra=dict(cglimit=1e-22, iter_max=2000))
    c = mc.copy_case(case)
    ndet = len(c["dets"])
    nsamp = c["timestamps"].size
    pix = c["pixels"].reshape(ndet, nsamp)
    pix[:, 1000:1010] = -1
    pix[1, 5000:5100] = -1
    sig = signals_for(c, 6)
    res, _ = batch(c, sig, 1)
    nrec = res[0]["info"]["rle_records_gather"]
    print("run-length records: baseline side", nrec, "pixel side", res[0]["info"]["rle_records_scatter"],
          "samples", ndet * nsamp)
    assert 0 < nrec < ndet * nsamp
    old = os.environ.get("MADAM_RLE")
    try:
        os.environ["MADAM_RLE"] = "0"
        res0, _ = batch(c, sig, 1)
    finally:
        if old is None:
            os.environ.pop("MADAM_RLE", None)
        else:
            os.environ["MADAM_RLE"] = old
    assert res0[0]["info"]["rle_records_gather"] == 0
    for k in range(6):
        good = res0[k]["binmap"][:, 0] != 0
        a1, m1 = offset_free(res[k]["aa"], res[k]["map"], good)
        a0, m0 = offset_free(res0[k]["aa"], res0[k]["map"], good)
        print("rhs", k, "passes", res[k]["info"]["passes"], res0[k]["info"]["passes"], "map", rel_rms(m1, m0),
              "baselines", rel_rms(a1, a0))
        assert res[k]["info"]["rr_final"] < 1e-22
        assert rel_rms(m1, m0) <= 1e-6 and rel_rms(a1, a0) <= 1e-6
        assert np.array_equal(res[k]["binmap"], res0[k]["binmap"])
    for k in (0, 5):
        ck = mc.copy_case(c)
        ck["signal"] = sig[k].reshape(-1).copy()
        o = oracle.destripe(*mc.call_args(ck))
        good = o["mask"] > 0
        ag, mg = offset_free(res[k]["aa"], res[k]["map"], good)
        ao, mo = offset_free(o["aa"], o["map"], good)
        print("rhs", k, "vs oracle: map", rel_rms(mg, mo), "baselines", rel_rms(ag, ao))
        assert rel_rms(mg, mo) <= 1e-6 and rel_rms(ag, ao) <= 1e-6


def test_wide_batch_64():
    """Forty right-hand sides, kernel width 64 (24 padding lanes): FP64 batch
    against single solves for seven of them (1e-10, same iteration counts
    within one), and the mixed solve of the same batch (run-length records)
    against the FP64 batch to 1e-6 less the null-space offset, both at
    cglimit 1e-20."""
    case = ring(prior=True, nsamp=24000, nperiod=2, extra=dict(cglimit=1e-20, iter_max=2000))
    sig = signals_for(case, 40)
    res, _ = batch(case, sig, 0)
    for k in (0, 1, 7, 8, 31, 32, 39):
        s1 = single(case, sig[k])
        r = res[k]
        errs = dict(map=rel_rms(r["map"], s1["map"]), aa=rel_rms(r["aa"], s1["aa"]), yba=rel_rms(r["yba"], s1["yba"]))
        print("rhs", k, "steps", r["info"]["steps"], "single", s1["niter"], errs)
        assert abs(r["info"]["steps"] - s1["niter"]) <= 1
        for name, e in errs.items():
            assert e <= 1e-10, (k, name, e)
    mix, _ = batch(case, sig, 1)
    assert mix[0]["info"]["rle_records_gather"] > 0
    for k in range(0, 40, 3):
        good = res[k]["binmap"][:, 0] != 0
        a1, m1 = offset_free(mix[k]["aa"], mix[k]["map"], good)
        a0, m0 = offset_free(res[k]["aa"], res[k]["map"], good)
        assert mix[k]["info"]["rr_final"] < 1e-20
        assert rel_rms(m1, m0) <= 1e-6 and rel_rms(a1, a0) <= 1e-6, (k, rel_rms(m1, m0), rel_rms(a1, a0))


@pytest.mark.parametrize("width", [20, 30])
def test_fp64_batch_wide_band(width):
    """Three right-hand sides with band widths of the wider kernel classes
    (kd 19 and 29: the carry chains stage more than 64 terms per lane group)
    against three single solves."""
    case = ring(prior=True, ndet=4, nsamp=60000, nperiod=2,
                extra=dict(precond_width_min=width, precond_width_max=width))
    sig = signals_for(case, 3, seed=300)
    res, st = batch(case, sig, 0)
    assert st["precond_kind"] == 3
    for k in range(3):
        s1 = single(case, sig[k])
        assert s1["stats"]["nband"] == width
        r = res[k]
        errs = dict(map=rel_rms(r["map"], s1["map"]), aa=rel_rms(r["aa"], s1["aa"]))
        print("width", width, "rhs", k, "steps", r["info"]["steps"], "single", s1["niter"], errs)
        assert abs(r["info"]["steps"] - s1["niter"]) <= 1
        for name, e in errs.items():
            assert e <= 1e-10, (name, e)


def test_band_chain_packing_bitwise():
    """The band preconditioner's carry chains several per wave (kernel_variant
    + 32, staged in windows of 16 chunks) and one per wave (+ 64) give the same
    bits: single FP64 solves with two chain groups per segment, and batched
    FP64 and mixed solves (right-hand sides indexed across the packed lanes)."""
    import libmadam_amd
    case = ring(prior=True, ndet=2, nsamp=500000, nperiod=2)
    g = {v: libmadam_amd.run(*mc.call_args(mc.copy_case(case)), kernel_variant=v) for v in (32, 64)}
    assert g[32]["stats"]["precond_kind"] == 3
    assert g[32]["niter"] == g[64]["niter"]
    for name in ("map", "aa", "binmap"):
        assert np.array_equal(np.asarray(g[32][name]), np.asarray(g[64][name])), name

    case = ring(prior=True, nsamp=30000, nperiod=2)
    sig = signals_for(case, 9)
    a = mc.call_args(mc.copy_case(case))
    ndet = len(a[1])
    nnz = a[5].size // (a[3].size * ndet)
    old = os.environ.get("MADAM_BS_PIECE")
    os.environ["MADAM_BS_PIECE"] = "256"   # the same pixel-side cut in both contexts
    try:
        for precision in (0, 1):
            res = {}
            for v in (32, 64):
                m = libmadam_amd.Madam(0)
                try:
                    m.configure(*a[:4], a[7], a[8], a[9], a[10], a[11], nnz, 0)
                    m.set_kernel_variant(v)
                    m.set_pointing(a[4], a[5])
                    m.solve_batch(sig, precision=precision)
                    res[v] = [m.batch_result(k) for k in range(sig.shape[0])]
                finally:
                    m.close()
            for k in range(sig.shape[0]):
                assert res[32][k]["info"]["steps"] == res[64][k]["info"]["steps"], (precision, k)
                for name in ("map", "aa"):
                    assert np.array_equal(res[32][k][name], res[64][k][name]), (precision, k, name)
    finally:
        if old is None:
            os.environ.pop("MADAM_BS_PIECE", None)
        else:
            os.environ["MADAM_BS_PIECE"] = old
