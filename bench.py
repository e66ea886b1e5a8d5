"""Benchmark: destriped TOD samples/s per PCG iteration on one MI355X.

Workload (the FP64 destripe with noise prior): 16 detectors x 2^24 samples at
100 Hz, synthetic ring scan (60 s spin, 85 deg opening angle, spin axis
advancing 1 deg/hour), nested HEALPix Nside 512 IQU, 0.25 s baselines, noise
prior from the PSD (sigma 2e-4 K, fknee 0.05 Hz, alpha -1.5), two pointing
periods of 2^23 samples (FFT length 2^20), PCG tolerance 1e-12, FP64.

A step is one PCG iteration of that solve (scatter P^T W F p, noise-prior
FFT convolution, gather F^T W P m, band-Cholesky preconditioner, FP64 dots and
vector updates), with every input resident in HBM.  The pointing is expanded
on the GPU, the TOD is synthesised on the GPU (torch is only a buffer holder
and test-data generator here).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "destriped TOD samples/s per PCG iter + p50/p99 solve latency; % HBM roofline"
HBM_PEAK_GBS = 8000.0

SCAN = dict(fsample=100.0, spin_period=60.0, prec_period=360 * 3600.0, opening_turns=85.0 / 360.0, lon0_turns=0.0)
SIGMA, FKNEE, ALPHA = 2e-4, 0.05, -1.5


def detectors(ndet):
    return [dict(psi_turns=(0.0, 0.125, 0.25, 0.375)[j % 4], open_offset_turns=(j // 4) * 0.1 / 360.0,
                 spin_offset_turns=0.0) for j in range(ndet)]


def psd_table(fs):
    freqs = np.logspace(-5, np.log10(fs / 2), 200)
    return freqs, SIGMA ** 2 / fs * (1.0 + (freqs / FKNEE) ** ALPHA)


def params(nside, fs, base):
    return dict(base_first=base, fsample=fs, nside_map=nside, nside_cross=nside, nside_submap=16, kfilter=True,
                cglimit=1e-12, iter_max=1000, info=0)


def make_tod_gpu(torch, pix, w, nside, ndet, nsamp, fs, seed):
    """Sky (100/5/5 uK white per pixel) + white noise + 1/f shaped by FFT, on the GPU."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    npix = 12 * nside * nside
    sky = torch.randn(npix, 3, generator=g, device="cuda", dtype=torch.float64)
    sky *= torch.tensor([100e-6, 5e-6, 5e-6], device="cuda", dtype=torch.float64)
    tod = torch.empty(ndet, nsamp, device="cuda", dtype=torch.float64)
    f = torch.fft.rfftfreq(nsamp, 1.0 / fs, device="cuda", dtype=torch.float64)
    shape = torch.zeros_like(f)
    shape[1:] = torch.sqrt(1.0 + (f[1:] / FKNEE) ** ALPHA)
    for d in range(ndet):
        white = torch.randn(nsamp, generator=g, device="cuda", dtype=torch.float64) * SIGMA
        noise = torch.fft.irfft(torch.fft.rfft(white) * shape, n=nsamp)
        sig = (sky[pix[d].long()] * w[d].double()).sum(-1)
        tod[d] = sig + noise
        del white, noise, sig
    del sky
    return tod


def gpu_run(ndet, nsamp, nside, fs, base, nperiod, steps, warmup, nsolve, seed=1234):
    import torch
    import libmadam_amd

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dets = detectors(ndet)
    pix = torch.empty(ndet, nsamp, dtype=torch.int32, device="cuda")
    w = torch.empty(ndet, nsamp, 3, dtype=torch.float32, device="cuda")
    libmadam_amd.expand_pointing_device(SCAN | dict(fsample=fs), dets, nside, 0, nsamp, pix, w)
    torch.cuda.synchronize()
    tod0 = make_tod_gpu(torch, pix, w, nside, ndet, nsamp, fs, seed)
    torch.cuda.synchronize()
    freqs, psd = psd_table(fs)
    periods = (np.arange(nperiod) * (nsamp // nperiod)).astype(np.int64)
    m = libmadam_amd.Madam(0)
    m.configure(params(nside, fs, base), ["det%02d" % j for j in range(ndet)], np.full(ndet, 1.0 / SIGMA ** 2),
                np.arange(nsamp) / fs, periods, np.ones(ndet, np.int64), np.zeros(ndet), freqs,
                np.tile(psd, ndet), 3)
    t0 = time.perf_counter()
    m.set_pointing(pix, w, on_device=True)
    t_pointing = time.perf_counter() - t0
    # full solves: latency.  The first includes the one-off allocations and the
    # segment / filter / preconditioner build; later ones (same pointing, new
    # TOD) reuse those when nna is bitwise unchanged.
    lat = []
    st_first = None
    for _ in range(nsolve):
        tod = tod0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m.solve(tod, on_device=True)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        del tod
        if st_first is None:
            st_first = m.stats()
    st_solve = m.stats()
    # timed PCG iterations (continuing one solve's trajectory)
    m.set_profiling(True)
    m.iterate(warmup, restart=True)
    m.reset_timers()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.iterate(steps, restart=False)
    torch.cuda.synchronize()
    t_steps = time.perf_counter() - t0
    st = m.stats()
    m.close()
    return dict(t_steps=t_steps, lat=lat, st=st, st_solve=st_solve, st_first=st_first, t_pointing=t_pointing)


def cpu_baseline(ndet, nsamp, nside, fs, base, seed=99):
    """The FP64 CPU oracle (a port of the reference algorithm) on a bounded
    sample of the same workload, timed on this host's cores."""
    import oracle
    from tests import madam_cases as mc

    rng = np.random.default_rng(seed)
    dets = detectors(ndet)
    pix = np.zeros((ndet, nsamp), np.int32)
    w = np.zeros((ndet, nsamp, 3), np.float32)
    for d in range(ndet):
        pix[d], w[d] = oracle.expand_pointing(SCAN | dict(fsample=fs), dets[d], nside, 0, nsamp)
    npix = 12 * nside * nside
    sky = rng.standard_normal((npix, 3)) * np.array([100e-6, 5e-6, 5e-6])
    sig = np.einsum("dsk,dsk->ds", sky[pix], w.astype(np.float64))
    for d in range(ndet):
        sig[d] += mc.one_over_f(rng, nsamp, fs, SIGMA, FKNEE, ALPHA)
    freqs, psd = psd_table(fs)
    pars = dict(params(nside, fs, base), iter_max=12)
    t0 = time.perf_counter()
    r = oracle.destripe(pars, ["d%d" % d for d in range(ndet)], np.full(ndet, 1.0 / SIGMA ** 2), np.arange(nsamp) / fs,
                        pix.reshape(-1), w.reshape(-1), sig.reshape(-1), np.array([0], np.int64),
                        np.ones(ndet, np.int64), np.zeros(ndet), freqs, np.tile(psd, ndet))
    wall = time.perf_counter() - t0
    t_it = r["t_iter"] / max(r["niter"], 1)
    return dict(value=ndet * nsamp / t_it, unit="samples/s", cores=int(r["nthreads"]), kind="port",
                sample="%d det x %d samples (1 period), Nside %d, %.2f s baselines, noise prior, FP64; %d PCG "
                       "iterations timed (%.2f s per iteration), solve wall %.1f s"
                       % (ndet, nsamp, nside, base, r["niter"], t_it, wall))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nsolve", type=int, default=5)
    ap.add_argument("--ndet", type=int, default=16)
    ap.add_argument("--log2nsamp", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the run, copy 1 GiB device-to-device three times (a plain copy kernel of known "
                         "size, the reference point for the FETCH_SIZE / WRITE_SIZE counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU (--gpus 1): the solve runs on one GPU, and sharding it across "
              "several GPUs is not built", file=sys.stderr)
        sys.exit(2)
    ndet, nsamp, nside, fs, base, nperiod = a.ndet, 2 ** a.log2nsamp, 512, 100.0, 0.25, 2
    print("bench: workload %d det x 2^%d samples, Nside %d" % (ndet, a.log2nsamp, nside), file=sys.stderr, flush=True)
    r = gpu_run(ndet, nsamp, nside, fs, base, nperiod, a.steps, a.warmup, a.nsolve)
    st = r["st"]
    N = ndet * nsamp
    ms_step = r["t_steps"] / a.steps * 1e3
    value = N / (ms_step * 1e-3)
    # algorithmic bytes per launch of the two pointing-stream kernels
    npix = 12 * nside * nside
    nbtot = ndet * ((nsamp // nperiod - 1) // int(round(base * fs)) + 1) * nperiod
    nact = st["nactive"]
    rec = st.get("rec_bytes") or 16   # bytes per pointing record the projections stream (16, 12 or 4)
    bytes_gather = N * rec + npix * 24 + nbtot * 5 * 8
    bytes_scatter = nact * rec + nbtot * 8 + npix * (8 + 48 + 32)
    t_g = st["t_gather"] / max(st["n_gather"], 1)
    t_s = st["t_scatter"] / max(st["n_scatter"], 1)
    kern = {"gather_baselines": dict(avg_ms=t_g * 1e3, alg_bytes=bytes_gather, gbs=bytes_gather / t_g / 1e9),
            "scatter_pixels": dict(avg_ms=t_s * 1e3, alg_bytes=bytes_scatter, gbs=bytes_scatter / t_s / 1e9)}
    dom = max(kern, key=lambda k: kern[k]["avg_ms"])
    traffic = None
    tf = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get(dom)
        except Exception:
            traffic = None
    lat = sorted(r["lat"][1:] or r["lat"])
    line = {
        "metric": METRIC, "value": value, "unit": "samples/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "FP64 destripe with noise prior: %d det x 2^%d samples @ 100 Hz, Nside %d IQU, "
                               "0.25 s baselines, 2 periods, nof 2^20, band-Cholesky preconditioner"
                               % (ndet, a.log2nsamp, nside),
                   "ndet": ndet, "nsamp": nsamp, "nside": nside, "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": kern[dom]["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kern[dom]["gbs"] / HBM_PEAK_GBS, "traffic": traffic},
        "kernels": kern,
        "step_breakdown_ms": {k: st[k] / max(st["n_gather"], 1) * 1e3 for k in
                              ("t_scatter", "t_prior", "t_gather", "t_apply_prec", "t_step")},
        "solve_latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                            "n": len(lat), "iterations": r["st_solve"]["niter"],
                            "preconditioner_reused": bool(r["st_solve"]["prec_reused"]), "first": r["lat"][0]},
        "setup_s": {"pointing": r["t_pointing"], "bin": r["st_first"]["t_bin"], "init": r["st_first"]["t_init"],
                    "precond": r["st_first"]["t_precond"]},
    }
    if a.copy_calibration:
        import torch
        src = torch.ones(2 ** 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("bench: copy calibration, 3 x %d bytes read and written" % (src.numel() * 8), file=sys.stderr,
              flush=True)
        del src, dst
    if not a.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(2, 2 ** 22, nside, fs, base)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
