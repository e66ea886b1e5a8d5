"""The request loop (madam_server_*): jobs of different scans, sizes and
precisions submitted together, against each job solved alone.

Bars: FP64 jobs equal their solve alone to 1e-10 relative rms (maps,
baselines), mixed-precision jobs to 1e-6 less the null-space offset; hits
bit-exact.  The loop must
have batched same-scan jobs, reused resident scans and run several jobs at
once, and admission must keep the estimated device memory within the budget.
"""
import numpy as np
import pytest

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


def scans():
    """three scans: (case, scan key).  The solves converge to cglimit 1e-20:
    at the default 1e-12, two mixed-precision solves that take different
    rounding paths (a batch of two against one alone) differ by up to about
    1e-5, the spread that tolerance leaves."""
    tight = dict(cglimit=1e-20, iter_max=2000)
    a = ring(nside=8, ndet=4, nsamp=20000, prior=True, nperiod=2, seed=1, extra=tight)
    b = ring(nside=16, ndet=2, nsamp=30000, prior=True, nperiod=2, seed=2, extra=tight)
    c = ring(nside=8, ndet=4, nsamp=12000, prior=False, nperiod=1, seed=3, extra=tight)
    return [(a, 101), (b, 202), (c, 303)]


def jobs():
    """(scan index, signal seed, precision, float32 TOD)"""
    return [(0, 0, 0, False), (0, 1, 0, False), (0, 2, 0, False), (1, 3, 1, True), (1, 4, 1, True),
            (2, 5, 0, False), (2, 6, 1, False), (0, 7, 0, False)]


def signal(case, seed):
    rng = np.random.default_rng(1000 + seed)
    ndet = len(case["dets"])
    nsamp = case["timestamps"].size
    s = case["signal"].reshape(ndet, nsamp).copy()
    for d in range(ndet):
        s[d] += mc.one_over_f(rng, nsamp, case["pars"]["fsample"], 1e-4, 0.1, -1.0)
    return s


def submit(srv, case, key, sig, precision):
    return srv.submit(case["pars"], case["weights"], case["timestamps"], case["pixels"], case["pixweights"], sig,
                      case["periods"], case["npsd"], case["psdstarts"], case["psdfreqs"], case["psdvals"],
                      nnz=case["nnz"], scan_key=key, precision=precision)


def test_jobs_alone_and_in_the_mix():
    import libmadam_amd
    sc = scans()
    js = jobs()
    sigs = []
    for (si, seed, prec, f32) in js:
        s = signal(sc[si][0], seed)
        sigs.append(s.astype(np.float32) if f32 else s)
    # in the mix: everything submitted at once, three workers
    srv = libmadam_amd.Server(0, hbm_budget=8e9, max_inflight=3)
    try:
        ids = [submit(srv, sc[si][0], sc[si][1], sigs[k], prec) for k, (si, seed, prec, f32) in enumerate(js)]
        mix = []
        for jid in ids:
            srv.wait(jid)
            mix.append(srv.result(jid))
        st = srv.stats()
    finally:
        srv.close()
    print("server stats", st)
    print("batch sizes", [r["batch"] for r in mix], "resident", [r["resident"] for r in mix])
    assert st["done"] == len(js) and st["failed"] == 0
    assert st["batches"] >= 1 and st["max_inflight"] >= 2
    assert st["peak_bytes"] <= 8e9
    # alone: one job per fresh server, without a scan key
    for k, (si, seed, prec, f32) in enumerate(js):
        srv = libmadam_amd.Server(0, hbm_budget=8e9, max_inflight=1)
        try:
            jid = submit(srv, sc[si][0], 0, sigs[k], prec)
            srv.wait(jid)
            alone = srv.result(jid)
        finally:
            srv.close()
        r = mix[k]
        tol = 1e-10 if prec == 0 else 1e-6
        good = alone["hits"] > 0
        m1, m0 = r["map"][good].copy(), alone["map"][good].copy()
        a1, a0 = r["aa"].copy(), alone["aa"].copy()
        if prec == 1:
            # less the null-space offset traded between the baselines and the I map
            m1[:, 0] -= m1[:, 0].mean()
            m0[:, 0] -= m0[:, 0].mean()
            a1 -= a1.mean()
            a0 -= a0.mean()
        errs = dict(map=rel_rms(m1, m0), aa=rel_rms(a1, a0), binmap=rel_rms(r["binmap"][good], alone["binmap"][good]))
        print("job", k, "scan", si, "precision", prec, "batch", r["batch"], "resident", r["resident"], errs)
        assert np.array_equal(r["hits"], alone["hits"])
        for name, e in errs.items():
            assert e <= tol, (k, name, e)


def test_admission_by_budget():
    """A budget that holds one job at a time: every job still completes, the
    estimated memory in use never exceeds the budget, and a job larger than
    the budget is refused at submission."""
    import libmadam_amd
    case, key = scans()[0]
    sig = signal(case, 0)
    srv0 = libmadam_amd.Server(0, hbm_budget=1e12, max_inflight=1)
    # the estimate of one job, read back from a generous server's peak
    jid = submit(srv0, case, 0, sig, 0)
    srv0.wait(jid)
    one = srv0.stats()["peak_bytes"]
    srv0.close()
    srv = libmadam_amd.Server(0, hbm_budget=1.5 * one, max_inflight=4)
    try:
        ids = [submit(srv, case, 0, signal(case, s), 0) for s in range(4)]
        for jid in ids:
            srv.wait(jid)
        st = srv.stats()
        assert st["done"] == 4 and st["peak_bytes"] <= 1.5 * one and st["max_inflight"] == 1
        small = libmadam_amd.Server(0, hbm_budget=0.5 * one, max_inflight=1)
        with pytest.raises(Exception, match="budget"):
            submit(small, case, 0, sig, 0)
        small.close()
    finally:
        srv.close()


def test_device_inputs_match_host_inputs(monkeypatch):
    """Jobs whose pointing and signal are device tensors (on_device) give the
    same results as the same jobs from host arrays, bit for bit: an FP64 job
    alone, a mixed job alone and a batch of two mixed jobs on one scan (the
    pixel-side cut is pinned: the autotuned one can differ between runs)."""
    import torch
    import libmadam_amd
    monkeypatch.setenv("MADAM_BS_PIECE", "256")
    case, key = scans()[1]
    ndet, nsamp = len(case["dets"]), case["timestamps"].size
    plan = [(0, 0), (1, 1), (1, 2), (1, 3)]   # (precision, seed); the last two share a batch
    res = {}
    for dev in (False, True):
        srv = libmadam_amd.Server(0, hbm_budget=8e9, max_inflight=1)
        try:
            keep, ids = [], []
            for prec, seed in plan:
                sig = signal(case, seed)
                sig = sig.astype(np.float32) if prec == 1 else sig
                pix, pw = case["pixels"], case["pixweights"]
                if dev:
                    pix = torch.from_numpy(pix.copy()).cuda()
                    pw = torch.from_numpy(pw.copy()).cuda()
                    sig = torch.from_numpy(np.ascontiguousarray(sig)).cuda()
                    torch.cuda.synchronize()
                keep.append((pix, pw, sig))
                ids.append(srv.submit(case["pars"], case["weights"], case["timestamps"], pix, pw, sig,
                                      case["periods"], case["npsd"], case["psdstarts"], case["psdfreqs"],
                                      case["psdvals"], nnz=case["nnz"], scan_key=key if seed >= 2 else 0,
                                      precision=prec, on_device=dev))
            out = []
            for jid in ids:
                srv.wait(jid)
                out.append(srv.result(jid))
        finally:
            srv.close()
        res[dev] = out
    for k in range(len(plan)):
        h, d = res[False][k], res[True][k]
        assert np.array_equal(h["hits"], d["hits"])
        assert h["batch"] == d["batch"]
        for name in ("map", "binmap", "aa"):
            assert np.array_equal(h[name], d[name]), (k, name)
