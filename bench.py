"""Benchmark: destriped TOD samples/s per PCG iteration on one MI355X.

Workloads (--workload):
  cfg3 (default)  mixed precision: 32 detectors x 2^25 samples at 100 Hz,
                  synthetic ring scan (60 s spin, 85 deg opening angle, spin
                  axis advancing 1 deg/hour), nested HEALPix Nside 1024 IQU,
                  FP32 TOD and weights, int32 pixels, noise prior from the PSD
                  (sigma 2e-4 K, fknee 0.05 Hz, alpha -1.5), 0.25 s baselines,
                  four periods of 2^23 samples (FFT length 2^20), PCG tolerance
                  1e-12.  FP32 operator inside FP64 iterative refinement; the
                  FP64 path runs on the same workload for the ratio and for the
                  per-output errors.
  cfg2            FP64 with noise prior: 16 detectors x 2^24, Nside 512, two
                  periods of 2^23, band-Cholesky preconditioner.
  cfg4            batched right-hand sides: the cfg2 pointing, 64 noise-only
                  Monte Carlo realisations in three detector subsets (16 det x
                  2^24 samples x 32 realisations, 8 x 2^23 x 16, 4 x 2^22 x 16),
                  each subset one batched mixed-precision solve.
A step is one PCG iteration (for the mixed solves, one step of the FP32
inner PCG; the FP64 refinement passes are in the solve latency), with every
input resident in HBM.  Pointing and TOD are generated on the GPU (torch is
only a buffer holder and test-data generator here).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cfg3|cfg2|cfg4] [--no-cpu]
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


def detectors(ndet, distinct=False):
    """Four polarisation angles per horn; horns offset along the opening angle.
    distinct: every detector on its own boresight (no co-pointed pairs)."""
    if distinct:
        return [dict(psi_turns=(0.0, 0.125, 0.25, 0.375)[j % 4], open_offset_turns=j * 0.05 / 360.0,
                     spin_offset_turns=j * 0.0011) for j in range(ndet)]
    return [dict(psi_turns=(0.0, 0.125, 0.25, 0.375)[j % 4], open_offset_turns=(j // 4) * 0.1 / 360.0,
                 spin_offset_turns=0.0) for j in range(ndet)]


def psd_table(fs):
    freqs = np.logspace(-5, np.log10(fs / 2), 200)
    return freqs, SIGMA ** 2 / fs * (1.0 + (freqs / FKNEE) ** ALPHA)


def params(nside, fs, base):
    return dict(base_first=base, fsample=fs, nside_map=nside, nside_cross=nside, nside_submap=16, kfilter=True,
                cglimit=1e-12, iter_max=1000, info=0)


def make_tod_gpu(torch, pix, w, nside, ndet, nsamp, fs, seed, dtype, sky=True):
    """Sky (100/5/5 uK white per pixel) + white noise + 1/f shaped by FFT, on the GPU."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    npix = 12 * nside * nside
    skym = None
    if sky:
        skym = torch.randn(npix, 3, generator=g, device="cuda", dtype=torch.float64)
        skym *= torch.tensor([100e-6, 5e-6, 5e-6], device="cuda", dtype=torch.float64)
    tod = torch.empty(ndet, nsamp, device="cuda", dtype=dtype)
    f = torch.fft.rfftfreq(nsamp, 1.0 / fs, device="cuda", dtype=torch.float64)
    shape = torch.zeros_like(f)
    shape[1:] = torch.sqrt(1.0 + (f[1:] / FKNEE) ** ALPHA)
    for d in range(ndet):
        white = torch.randn(nsamp, generator=g, device="cuda", dtype=torch.float64) * SIGMA
        sig = torch.fft.irfft(torch.fft.rfft(white) * shape, n=nsamp)
        del white
        if sky:
            sig += (skym[pix[d].long()] * w[d].double()).sum(-1)
        tod[d] = sig.to(dtype)
        del sig
    del skym
    return tod


def configure(m, nside, fs, base, ndet, nsamp, nperiod, extra=None):
    freqs, psd = psd_table(fs)
    periods = (np.arange(nperiod) * (nsamp // nperiod)).astype(np.int64)
    pars = params(nside, fs, base)
    if extra:
        pars.update(extra)
    m.configure(pars, ["det%02d" % j for j in range(ndet)], np.full(ndet, 1.0 / SIGMA ** 2), np.arange(nsamp) / fs,
                periods, np.ones(ndet, np.int64), np.zeros(ndet), freqs, np.tile(psd, ndet), 3)


def pointing(torch, libmadam_amd, ndet, nsamp, nside, fs, distinct=False, i0=0):
    pix = torch.empty(ndet, nsamp, dtype=torch.int32, device="cuda")
    w = torch.empty(ndet, nsamp, 3, dtype=torch.float32, device="cuda")
    libmadam_amd.expand_pointing_device(SCAN | dict(fsample=fs), detectors(ndet, distinct), nside, i0, nsamp, pix, w)
    torch.cuda.synchronize()
    return pix, w


def rel_err(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    d = a - b
    den = np.sqrt(np.mean(b ** 2)) or 1.0
    return dict(rms=float(np.sqrt(np.mean(d ** 2)) / den), max=float(np.abs(d).max() / (np.abs(b).max() or 1.0)))


# ------------------------------------------------------------------ algorithmic bytes of the projections


def bytes_fp64(st, ndet, nsamp, npix, nbtot, s=8):
    """the one-RHS projections (k_gather_flat, k_sr_scatter) in value size s (8: FP64, 4: the mixed
    solve's FP32 operator), per launch: one record per sample / CSR entry, per hit pixel the map
    (3 values, read once) and on the pixel side the head / ordinal bits (8 B), the inverse pixel
    matrix (6 values) and the map written (4 values), per baseline nna, p, dwb, ap r/w (5 values)
    and pw (1 value)"""
    rec = st.get("rec_bytes") or 16
    nact = st["nactive"]
    return dict(gather=ndet * nsamp * rec + npix * 3 * s + nbtot * 5 * s,
                scatter=nact * rec + nbtot * s + npix * (8 + 6 * s + 4 * s))


def bytes_batch(st, ndet, nsamp, npix_hit, nbtot, R, s, info=None):
    """the batched kernels (k_bgather_w, k_bscatter) in type size s for R right-hand sides: the
    pointing records once (16-byte run-length records when the FP32 operator uses them), per
    right-hand side its vectors and map values"""
    rec = st.get("rec_bytes") or 16
    nact = st["nactive"]
    g_rec = ndet * nsamp * rec
    s_rec = nact * rec
    if info and info.get("rle_records_gather"):
        g_rec = info["rle_records_gather"] * 16
        s_rec = info["rle_records_scatter"] * 16
    return dict(gather=g_rec + npix_hit * 3 * s * R + nbtot * (2 * s + 8 + 3 * s * R),
                scatter=s_rec + nbtot * s * R + npix_hit * (6 * s + 4 + 24 + 3 * s * R))


def roofline_from(times, nbytes, traffic_key, workload):
    """the longest of the projection kernels: algorithmic bytes per launch / mean launch time; traffic =
    its HBM bytes per launch measured on this workload (profiles/pmc_traffic.json), or null"""
    kern = {k: dict(avg_ms=times[k] * 1e3, alg_bytes=nbytes[k], gbs=nbytes[k] / times[k] / 1e9) for k in times}
    dom = max(kern, key=lambda k: kern[k]["avg_ms"])
    traffic = None
    tf = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get(workload, {}).get(traffic_key[dom])
        except Exception:
            traffic = None
    roof = {"bound": "hbm", "kernel": traffic_key[dom], "achieved": kern[dom]["gbs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": kern[dom]["gbs"] / HBM_PEAK_GBS, "traffic": traffic}
    return roof, kern


# ------------------------------------------------------------------ workloads


def run_cfg3(a):
    import torch
    import libmadam_amd
    ndet, nsamp, nside, fs, base, nperiod = 32, 2 ** 25, 1024, 100.0, 0.25, 4
    if a.small:
        ndet, nsamp, nside = 4, 2 ** 22, 256
    npix = 12 * nside * nside
    log = lambda *x: print("bench:", *x, file=sys.stderr, flush=True)
    log("cfg3 workload %d det x %d samples, Nside %d" % (ndet, nsamp, nside))
    pix, w = pointing(torch, libmadam_amd, ndet, nsamp, nside, fs, a.distinct)
    tod32 = make_tod_gpu(torch, pix, w, nside, ndet, nsamp, fs, 1234, torch.float32)
    torch.cuda.synchronize()
    m = libmadam_amd.Madam(0)
    configure(m, nside, fs, base, ndet, nsamp, nperiod)
    t0 = time.perf_counter()
    m.set_pointing(pix, w, on_device=True)
    t_pointing = time.perf_counter() - t0
    # mixed precision: full solves (latency), then the timed inner PCG steps
    lat = []
    for k in range(a.nsolve):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m.solve_batch(tod32.view(1, ndet, nsamp), precision=1, on_device=True)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        log("mixed solve %d: %.3f s" % (k, lat[-1]))
    res = m.batch_result(0)
    st = m.stats()
    m.iterate_batch(a.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.iterate_batch(a.steps)
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / a.steps
    m.set_profiling(True)
    m.reset_timers()
    m.iterate_batch(a.prof_steps)
    prof = m.batch_result(0)["info"]
    m.set_profiling(False)
    log("mixed step %.3f ms" % (t_mixed * 1e3))
    # the FP64 path on the same workload (same FP32 TOD values in FP64)
    tod64 = tod32.double()
    del tod32
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.solve(tod64, on_device=True)
    torch.cuda.synchronize()
    lat64 = time.perf_counter() - t0
    st64 = m.stats()
    ref = m.maps()
    ref.update(m.baselines())
    m.iterate(a.warmup, restart=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.iterate(a.steps, restart=False)
    torch.cuda.synchronize()
    t_fp64 = (time.perf_counter() - t0) / a.steps
    log("fp64 step %.3f ms, solve %.3f s" % (t_fp64 * 1e3, lat64))
    m.close()
    del tod64, pix, w
    torch.cuda.empty_cache()
    # per-output errors of the mixed solve against the FP64 solve (the FP64
    # path equals the CPU oracle to 1e-10 at this per-detector shape)
    good = ref["mask"] > 0
    errs = {}
    for j, name in enumerate(("I", "Q", "U")):
        mg, mr = res["map"][good, j].copy(), ref["map"][good, j].copy()
        if j == 0:   # the null-space offset traded between baselines and the I map
            mg -= mg.mean()
            mr -= mr.mean()
        errs["map_" + name] = rel_err(mg, mr)
    errs["binmap"] = rel_err(res["binmap"][good], ref["binmap"][good])
    errs["baselines"] = rel_err(res["aa"] - res["aa"].mean(), ref["aa"] - ref["aa"].mean())
    N = ndet * nsamp
    nbtot = ndet * ((nsamp // nperiod - 1) // int(round(base * fs)) + 1) * nperiod
    npix_hit = int(st["npix_hit"]) if st.get("npix_hit") else npix
    nprof = max(prof["n_prof"], 1)
    times = dict(scatter=prof["t_scatter"] / nprof, gather=prof["t_gather"] / nprof)
    # one right-hand side runs the FP32 instantiations of the FP64 path's kernels
    roof, kern = roofline_from(times, bytes_fp64(st, ndet, nsamp, npix_hit, nbtot, 4),
                               dict(scatter="k_sr_scatter<float, 3, 4, 1>", gather="k_gather_flat<float, 3, 1>"),
                               "cfg3")
    value = N / t_mixed
    line = {
        "metric": METRIC, "value": value, "unit": "samples/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": t_mixed * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 operator / fp64 refinement, dots and pixel matrices", "data": "synthetic",
        "config": {"workload": "mixed-precision destripe: %d det x 2^%d samples @ 100 Hz, Nside %d IQU, FP32 TOD, "
                               "0.25 s baselines, %d periods, noise prior (nof 2^20), band preconditioner"
                               % (ndet, int(np.log2(nsamp)), nside, nperiod),
                   "ndet": ndet, "nsamp": nsamp, "nside": nside, "parallelism": "1 GPU"},
        "roofline": roof,
        "kernels": kern,
        "step_breakdown_ms": {k: prof[k] / nprof * 1e3 for k in ("t_scatter", "t_prior", "t_gather", "t_prec",
                                                                 "t_step")},
        "fp64_same_workload": {"ms_per_step": t_fp64 * 1e3, "samples_per_s": N / t_fp64,
                               "iterations": st64["niter"], "solve_s": lat64},
        "mixed_vs_fp64_step_speedup": t_fp64 / t_mixed,
        "mixed_solve": {"solve_s": lat, "pcg_steps": res["info"]["pcg_steps"], "passes": res["info"]["passes"],
                        "rr_final": res["info"]["rr_final"], "piece": res["info"]["piece"],
                        "set_pointing_s": t_pointing, "fp64_solve_over_mixed_solve": lat64 / min(lat)},
        "errors_vs_fp64": errs,
    }
    return line


def run_cfg2(a):
    import torch
    import libmadam_amd
    ndet, nsamp, nside, fs, base, nperiod = 16, 2 ** 24, 512, 100.0, 0.25, 2
    if a.small:
        ndet, nsamp, nside = 4, 2 ** 22, 256
    print("bench: cfg2 workload %d det x %d samples, Nside %d" % (ndet, nsamp, nside), file=sys.stderr, flush=True)
    pix, w = pointing(torch, libmadam_amd, ndet, nsamp, nside, fs, a.distinct)
    tod0 = make_tod_gpu(torch, pix, w, nside, ndet, nsamp, fs, 1234, torch.float64)
    m = libmadam_amd.Madam(0)
    configure(m, nside, fs, base, ndet, nsamp, nperiod)
    t0 = time.perf_counter()
    m.set_pointing(pix, w, on_device=True)
    t_pointing = time.perf_counter() - t0
    lat = []
    st_first = None
    for _ in range(a.nsolve):
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
    m.set_profiling(True)
    m.iterate(a.warmup, restart=True)
    m.reset_timers()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.iterate(a.steps, restart=False)
    torch.cuda.synchronize()
    t_steps = time.perf_counter() - t0
    st = m.stats()
    m.close()
    N = ndet * nsamp
    ms_step = t_steps / a.steps * 1e3
    npix = 12 * nside * nside
    nbtot = ndet * ((nsamp // nperiod - 1) // int(round(base * fs)) + 1) * nperiod
    ng = max(st["n_gather"], 1)
    times = dict(gather=st["t_gather"] / ng, scatter=st["t_scatter"] / ng)
    npix_hit = int(st["npix_hit"]) if st.get("npix_hit") else npix
    roof, kern = roofline_from(times, bytes_fp64(st, ndet, nsamp, npix_hit, nbtot),
                               dict(gather="k_gather_flat<double, 3, 1>", scatter="k_sr_scatter<double, 3, 4, 1>"),
                               "cfg2")
    lat_s = sorted(lat[1:] or lat)
    return {
        "metric": METRIC, "value": N / (ms_step * 1e-3), "unit": "samples/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "FP64 destripe with noise prior: %d det x 2^%d samples @ 100 Hz, Nside %d IQU, "
                               "0.25 s baselines, 2 periods, nof 2^20, band-Cholesky preconditioner%s"
                               % (ndet, int(np.log2(nsamp)), nside, ", distinct boresights" if a.distinct else ""),
                   "ndet": ndet, "nsamp": nsamp, "nside": nside, "parallelism": "1 GPU"},
        "roofline": roof, "kernels": kern,
        "step_breakdown_ms": {k: st[k] / ng * 1e3 for k in ("t_scatter", "t_prior", "t_gather", "t_apply_prec",
                                                            "t_step")},
        "solve_s": {"repeated": lat_s, "first": lat[0], "iterations": st_solve["niter"],
                    "preconditioner_reused": bool(st_solve["prec_reused"])},
        "setup_s": {"pointing": t_pointing, "bin": st_first["t_bin"], "init": st_first["t_init"],
                    "precond": st_first["t_precond"]},
    }


def run_cfg4(a):
    """64 Monte Carlo noise realisations on the cfg2 pointing in three detector
    subsets, each a batched mixed-precision solve; the FP64 single-RHS step of
    the full set is the reference rate."""
    import torch
    import libmadam_amd
    ndet0, nsamp0, nside, fs, base, nperiod = 16, 2 ** 24, 512, 100.0, 0.25, 2
    groups = [(16, 2 ** 24, 32), (8, 2 ** 23, 16), (4, 2 ** 22, 16)]
    if a.small:
        ndet0, nsamp0, nside = 4, 2 ** 21, 128
        groups = [(4, 2 ** 21, 8), (2, 2 ** 20, 4), (1, 2 ** 20, 4)]
    log = lambda *x: print("bench:", *x, file=sys.stderr, flush=True)
    pix0, w0 = pointing(torch, libmadam_amd, ndet0, nsamp0, nside, fs, a.distinct)
    out = []
    tot_rate, tot_work, tot_time, fp64_rate = 0.0, 0.0, 0.0, None
    for gi, (nd, ns, R) in enumerate(groups):
        pix = pix0[:nd, :ns].contiguous()
        w = w0[:nd, :ns].contiguous()
        m = libmadam_amd.Madam(0)
        configure(m, nside, fs, base, nd, ns, nperiod)
        m.set_pointing(pix, w, on_device=True)
        sig = torch.empty(R, nd, ns, dtype=torch.float32, device="cuda")
        for r in range(R):
            sig[r] = make_tod_gpu(torch, pix, w, nside, nd, ns, fs, 1000 + 97 * gi + r, torch.float32, sky=False)
        torch.cuda.synchronize()
        log("group %d: %d realisations generated" % (gi, R))
        t0 = time.perf_counter()
        m.solve_batch(sig, precision=1, on_device=True)
        torch.cuda.synchronize()
        t_solve = time.perf_counter() - t0
        log("group %d: batched solve done in %.2f s" % (gi, t_solve))
        info = m.batch_result(0)["info"]
        m.iterate_batch(a.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m.iterate_batch(a.steps)
        torch.cuda.synchronize()
        t_step = (time.perf_counter() - t0) / a.steps
        m.set_profiling(True)
        m.reset_timers()
        m.iterate_batch(a.prof_steps)
        prof = m.batch_result(0)["info"]
        m.set_profiling(False)
        nprof = max(prof["n_prof"], 1)
        brk = {k: prof[k] / nprof * 1e3 for k in ("t_scatter", "t_prior", "t_gather", "t_prec", "t_step")}
        if gi == 0:
            # the roofline of the largest subset's projections
            st = m.stats()
            nbt = nd * ((ns // nperiod - 1) // int(round(base * fs)) + 1) * nperiod
            nph = int(st["npix_hit"]) if st.get("npix_hit") else 12 * nside * nside
            rl = 0 if prof.get("rle_records_gather") else 1   # REC16 for run-length records, else REC12
            roof4, kern4 = roofline_from(dict(scatter=prof["t_scatter"] / nprof, gather=prof["t_gather"] / nprof),
                                         bytes_batch(st, nd, ns, nph, nbt, R, 4, prof),
                                         dict(scatter="k_bscatter<float, %d, %d, 3>" % (R, rl),
                                              gather="k_bgather_w<float, %d, %d, 3, %d>" % (R, rl, min(16 * R // 64, 8))),
                                         "cfg4")
            kern4["rle_records"] = dict(gather=prof.get("rle_records_gather"), scatter=prof.get("rle_records_scatter"))
        log("group %d: %d det x %d samples x %d RHS: step %.3f ms, solve %.2f s, breakdown %s"
            % (gi, nd, ns, R, t_step * 1e3, t_solve, brk))
        if gi == 0:
            # the FP64 single-RHS path on the first realisation, same pointing
            tod64 = sig[0].double().contiguous()
            m.solve(tod64, on_device=True)
            m.iterate(a.warmup, restart=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            m.iterate(a.steps, restart=False)
            torch.cuda.synchronize()
            t64 = (time.perf_counter() - t0) / a.steps
            fp64_rate = nd * ns / t64
            del tod64
        m.close()
        del sig, pix, w
        torch.cuda.empty_cache()
        out.append(dict(ndet=nd, nsamp=ns, nrhs=R, ms_per_step=t_step * 1e3, solve_s=t_solve,
                        sample_rhs_per_s=nd * ns * R / t_step, pcg_steps=info["pcg_steps"], passes=info["passes"],
                        step_breakdown_ms=brk, piece=info["piece"]))
        tot_work += nd * ns * R
        tot_time += t_step
        if a.groups and gi + 1 >= a.groups:
            break
    value = tot_work / tot_time
    return {
        "metric": METRIC, "value": value, "unit": "sample-RHS/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": tot_time * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 operator / fp64 refinement, dots and pixel matrices", "data": "synthetic",
        "config": {"workload": "batched RHS: 64 noise-only realisations on the %d det x 2^%d pointing (Nside %d) "
                               "in subsets %s (det, samples, RHS), mixed precision"
                               % (ndet0, int(np.log2(nsamp0)), nside, groups),
                   "parallelism": "1 GPU"},
        "roofline": roof4, "kernels": kern4,
        "groups": out, "fp64_single_rhs_samples_per_s": fp64_rate,
        "batched_over_fp64_single": value / fp64_rate if fp64_rate else None,
    }


def run_cfg5(a):
    """Serving stream: Poisson arrivals of destripe jobs through the request
    loop (madam_server_*).  Per job: Nside uniform in {128, 256, 512, 1024},
    detectors in {1, 2, 4, 8, 16, 32}, samples per detector log-uniform in
    [2^lo, 2^hi] at 100 Hz (2^18..2^25 by default), baseline length uniform in
    0.25..1 s, noise prior on half of the jobs, FP32 TOD, mixed precision; 30 %
    of the jobs reuse a scan already seen (same pointing and parameters, a new
    signal).  A feeder thread generates each job's pointing and signal on the
    GPU at its arrival time (on a stream of its own) and submits the device
    buffers; a job's inputs are freed when its result has been read.  The
    feeder holds back while the inputs of the jobs in the system exceed
    --input-cap GB (the time held back is reported)."""
    import threading
    import torch
    import libmadam_amd
    rng = np.random.default_rng(2024)
    njobs, lo, hi, rate = a.jobs, a.log2min, a.log2max, a.rate
    log = lambda *x: print("bench:", *x, file=sys.stderr, flush=True)
    fs = 100.0
    scans = []
    specs = []
    for k in range(njobs):
        if scans and rng.random() < 0.3:
            si = int(rng.integers(len(scans)))
        else:
            nside = int(rng.choice([128, 256, 512, 1024]))
            ndet = int(rng.choice([1, 2, 4, 8, 16, 32]))
            nsamp = int(2 ** rng.uniform(lo, hi)) // 64 * 64
            base = float(rng.uniform(0.25, 1.0))
            prior = bool(rng.random() < 0.5)
            scans.append(dict(nside=nside, ndet=ndet, nsamp=nsamp, base=base, prior=prior, key=len(scans) + 1))
            si = len(scans) - 1
        specs.append(si)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, njobs))
    total_samples = sum(scans[si]["ndet"] * scans[si]["nsamp"] for si in specs)
    log("cfg5: %d jobs over %d scans, samples/det 2^%g..2^%g, %.3g samples in all, rate %g/s"
        % (njobs, len(scans), lo, hi, total_samples, rate))
    freqs, psd = psd_table(fs)
    torch.cuda.init()   # torch's runtime first, then the library's
    srv = libmadam_amd.Server(0, hbm_budget=200e9, max_inflight=a.inflight)
    cap = a.input_cap * 1e9
    lock = threading.Lock()
    ids = [None] * njobs
    keep = {}
    state = dict(pending=0.0, held=0.0, failed=None)
    stop = threading.Event()
    gen_stream = torch.cuda.Stream()
    g = torch.Generator(device="cuda")
    g.manual_seed(2024)
    t0 = time.perf_counter()

    def feeder():
        with torch.cuda.stream(gen_stream):
            for k, si in enumerate(specs):
                if stop.is_set():
                    return
                dt = t0 + arrivals[k] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                sc = scans[si]
                nbytes = 20.0 * sc["ndet"] * sc["nsamp"]
                th0 = time.perf_counter()
                while not stop.is_set():
                    with lock:
                        if state["pending"] + nbytes <= cap or state["pending"] == 0:
                            state["pending"] += nbytes
                            break
                    time.sleep(0.002)
                state["held"] += time.perf_counter() - th0
                nd, ns = sc["ndet"], sc["nsamp"]
                pix = torch.empty(nd, ns, dtype=torch.int32, device="cuda")
                w = torch.empty(nd, ns, 3, dtype=torch.float32, device="cuda")
                libmadam_amd.expand_pointing_device(SCAN | dict(fsample=fs), detectors(nd), sc["nside"], 0, ns, pix, w)
                # white noise plus a random walk (1/f-like), in place
                sig = torch.randn(nd, ns, generator=g, device="cuda").mul_(2e-4)
                sig.add_(torch.cumsum(torch.randn(nd, ns, generator=g, device="cuda").mul_(2e-5), dim=1))
                gen_stream.synchronize()
                pars = dict(params(sc["nside"], fs, sc["base"]), kfilter=sc["prior"], nside_submap=min(16, sc["nside"]))
                jid = srv.submit(pars, np.full(nd, 1.0 / SIGMA ** 2), np.arange(ns) / fs, pix.view(-1), w.view(-1),
                                 sig, np.array([0], np.int64), np.ones(nd, np.int64), np.zeros(nd), freqs,
                                 np.tile(psd, nd), nnz=3, scan_key=sc["key"], precision=1, on_device=True)
                with lock:
                    keep[jid] = (pix, w, sig, nbytes)
                ids[k] = jid
                if (k + 1) % 100 == 0:
                    log("%d jobs submitted at %.1f s" % (k + 1, time.perf_counter() - t0))

    th = threading.Thread(target=feeder)
    th.start()
    # results are collected (and the jobs' inputs released) as the jobs finish
    lat, samples, steps, batch = [0.0] * njobs, 0, 0, [0] * njobs
    open_jobs, nxt, ndone = [], 0, 0
    try:
        while ndone < njobs:
            while nxt < njobs and ids[nxt] is not None:
                open_jobs.append(nxt)
                nxt += 1
            progressed = False
            for k in list(open_jobs):
                jid = ids[k]
                if not srv.wait(jid, 0.0):
                    continue
                r = srv.result(jid, maps=False)
                lat[k] = r["latency"]
                sc = scans[specs[k]]
                samples += sc["ndet"] * sc["nsamp"]
                steps += r["steps"]
                batch[k] = r["batch"]
                srv.release(jid)
                with lock:
                    state["pending"] -= keep.pop(jid)[3]
                open_jobs.remove(k)
                ndone += 1
                progressed = True
                if ndone % 200 == 0:
                    log("%d jobs done at %.1f s" % (ndone, time.perf_counter() - t0))
            if not progressed:
                time.sleep(0.002)
    except Exception:
        log("a job failed; server stats %s" % srv.stats())
        stop.set()
        th.join()
        raise
    th.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat = np.array(lat)
    return {
        "metric": METRIC, "value": samples / wall, "unit": "samples/s (solved jobs' samples per second of stream)",
        "n_gpus": 1, "steps": njobs, "warmup": 0, "ms_per_step": wall / njobs * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 operator / fp64 refinement, dots and pixel matrices", "data": "synthetic",
        "config": {"workload": "serving stream: %d Poisson jobs at %g/s, Nside {128..1024}, det {1..32}, samples/det "
                               "log-uniform 2^%g..2^%g at 100 Hz, baselines 0.25..1 s, prior on half, FP32 TOD, 30%% "
                               "resident-scan reuse, 200 GB HBM budget, %d workers; inputs generated on the GPU at "
                               "arrival" % (njobs, rate, lo, hi, a.inflight),
                   "parallelism": "1 GPU"},
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "mean": float(lat.mean()), "max": float(lat.max()), "n": int(lat.size)},
        "jobs_per_s": njobs / wall, "wall_s": wall, "samples": samples, "pcg_steps": steps,
        "feeder_held_s": state["held"], "server": st, "mean_batch": float(np.mean(batch)),
    }


# ------------------------------------------------------------------ CPU baseline


def cpu_baseline(ndet, nsamp, nside, fs, base, nperiod, niter=3, seed=99):
    """The FP64 CPU oracle (a port of the reference algorithm) on a sample of
    the timed workload (ndet detectors of one horn, the workload's Nside,
    baselines, periods and noise prior), timed on this host's cores; raises if
    the sample has nothing to solve."""
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
    pars = dict(params(nside, fs, base), iter_max=niter)
    periods = (np.arange(nperiod) * (nsamp // nperiod)).astype(np.int64)
    t0 = time.perf_counter()
    r = oracle.destripe(pars, ["d%d" % d for d in range(ndet)], np.full(ndet, 1.0 / SIGMA ** 2),
                        np.arange(nsamp) / fs, pix.reshape(-1), w.reshape(-1), sig.reshape(-1), periods,
                        np.ones(ndet, np.int64), np.zeros(ndet), freqs, np.tile(psd, ndet))
    wall = time.perf_counter() - t0
    if r["niter"] < 1:
        raise RuntimeError("CPU baseline sample ran no PCG iteration")
    t_it = r["t_iter"] / r["niter"]
    return dict(value=ndet * nsamp / t_it, unit="samples/s", cores=int(r["nthreads"]), kind="port",
                sample="%d det x 2^%d samples in %d periods (nof %d), Nside %d, %.2f s baselines, noise prior, band "
                       "preconditioner, FP64; %d PCG iterations timed (%.2f s per iteration), solve wall %.1f s"
                       % (ndet, int(np.log2(nsamp)), nperiod, r["nof"], nside, base, r["niter"], t_it, wall))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nsolve", type=int, default=2)
    ap.add_argument("--prof-steps", type=int, default=5)
    ap.add_argument("--workload", default="cfg3", choices=["cfg3", "cfg2", "cfg4", "cfg5"])
    ap.add_argument("--jobs", type=int, default=2000)
    ap.add_argument("--log2min", type=float, default=18.0)
    ap.add_argument("--log2max", type=float, default=25.0)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--input-cap", type=float, default=40.0, help="cfg5: GB of job inputs held on the GPU at most")
    ap.add_argument("--inflight", type=int, default=6)
    ap.add_argument("--groups", type=int, default=0, help="cfg4: run only the first N detector subsets")
    ap.add_argument("--distinct", action="store_true", help="every detector on its own boresight")
    ap.add_argument("--small", action="store_true", help="a reduced shape for quick checks (not a bench number)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the run, copy 1 GiB device-to-device three times (a plain copy kernel of known "
                         "size, the reference point for the FETCH_SIZE / WRITE_SIZE counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU (--gpus 1): the solve runs on one GPU, and sharding it across "
              "several GPUs is not built", file=sys.stderr)
        sys.exit(2)
    run = dict(cfg3=run_cfg3, cfg2=run_cfg2, cfg4=run_cfg4, cfg5=run_cfg5)[a.workload]
    line = run(a)
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
    if not a.no_cpu and a.workload != "cfg5":
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        # one horn of four polarisation angles: a single detector leaves every
        # pixel's 3x3 matrix singular, so there would be nothing to solve
        if a.workload == "cfg3" and not a.small:
            line["cpu_baseline"] = cpu_baseline(4, 2 ** 23, 1024, 100.0, 0.25, 4)
        elif not a.small:
            line["cpu_baseline"] = cpu_baseline(4, 2 ** 22, 512, 100.0, 0.25, 2)
        else:
            line["cpu_baseline"] = cpu_baseline(4, 2 ** 18, 256, 100.0, 0.25, 2)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
