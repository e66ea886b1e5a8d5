// Host orchestration and the C-ABI of libmadam.so.
//
// Reference entry points: destripe() (src/madam.h:2, src/smadam.F90:64-590),
// destripe_with_cache() (src/smadam.F90:593-918) and clear_caches()
// (src/smadam.F90:921-951).  The host side parses the parameter string
// (src/inputparam.f90:24-371, defaults src/commonparam.f90:200-305), sets up
// detectors (src/inputparam.f90:73-163), intervals and baselines
// (src/read_data.f90:203-323, src/parameter_control.f90:27-409, 752-977),
// and drives the HIP kernels; every per-sample, per-pixel and per-baseline
// step runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "madam_context.hpp"
#include "madam_fits.hpp"


namespace madam {

static thread_local std::string g_err;

void set_last_error(const char *msg) { g_err = msg ? msg : ""; }


[[noreturn]] void fail(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw Error{buf};
}

// ------------------------------------------------------------------ parameters


static std::string strip(const std::string &s)
{
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    return s.substr(a, b - a + 1);
}

static std::string token(const std::string &v)
{
    std::string s = strip(v);
    size_t e = s.find_first_of(" ,\t");
    return e == std::string::npos ? s : s.substr(0, e);
}

static bool parse_int(const std::string &v, int &out)
{
    std::string t = token(v);
    if (t.empty()) return false;
    char *end = nullptr;
    long x = strtol(t.c_str(), &end, 10);
    if (*end) return false;
    out = (int)x;
    return true;
}

static bool parse_real(const std::string &v, double &out)
{
    std::string t = token(v);
    if (t.empty()) return false;
    for (auto &c : t)
        if (c == 'd' || c == 'D') c = 'e';
    char *end = nullptr;
    double x = strtod(t.c_str(), &end);
    if (*end) return false;
    out = x;
    return true;
}

static bool parse_bool(const std::string &v, int &out)
{
    std::string t = token(v);
    size_t i = (!t.empty() && t[0] == '.') ? 1 : 0;
    if (i >= t.size()) return false;
    if (t[i] == 'T' || t[i] == 't') { out = 1; return true; }
    if (t[i] == 'F' || t[i] == 'f') { out = 0; return true; }
    return false;
}

static bool is_all(const std::string &s)
{
    return s.empty() || s == "all" || s == "All" || s == "ALL";
}

// src/inputparam.f90:169-371
static void read_line(Params &p, std::string line)
{
    line = strip(line);
    if (line.empty() || line[0] == '#') return;
    size_t eq = line.find('=');
    if (eq == std::string::npos) fail("Syntax error in parameter definition: %s", line.c_str());
    std::string key = strip(line.substr(0, eq)), value = strip(line.substr(eq + 1));
    size_t h = value.find('#');
    if (h != std::string::npos) value = strip(value.substr(0, h));
    if (key.compare(0, 8, "detector") == 0) key = "detector";
    if (value.empty() && key.compare(0, 4, "file") != 0) return;
    bool ok = true;
    int ig;
    double rg;
    if (key == "info") ok = parse_int(value, p.info);
    else if (key == "nthreads") ok = parse_int(value, p.nthreads);
    else if (key == "nsubchunk") ok = parse_int(value, p.nsubchunk);
    else if (key == "isubchunk") ok = parse_int(value, p.isubchunk);
    else if (key == "time_unit") { if (value[0] != 'p' && value[0] != 'P') ok = parse_real(value, rg); }
    else if (key == "base_first" || key == "nshort") ok = parse_real(value, p.dnshort);
    else if (key == "nside_map") ok = parse_int(value, p.nside_map);
    else if (key == "nside_cross") ok = parse_int(value, p.nside_cross);
    else if (key == "nside_submap") ok = parse_int(value, p.nside_submap);
    else if (key == "good_baseline_fraction") ok = parse_real(value, p.good_baseline_fraction);
    else if (key == "concatenate_messages" || key == "allreduce" || key == "reassign_submaps" ||
             key == "flag_by_horn" || key == "checknan" || key == "sync_output" || key == "unaligned_fft" ||
             key == "write_base" || key == "binary_output") ok = parse_bool(value, ig);
    else if (key == "write_leakmatrix") ok = parse_bool(value, p.write_leakmatrix);
    else if (key == "write_cut") ok = parse_bool(value, p.write_cut);
    else if (key == "skip_existing") ok = parse_bool(value, p.skip_existing);
    else if (key == "force_pol") ok = parse_bool(value, p.force_pol);
    else if (key == "write_map") ok = parse_bool(value, p.write_map);
    else if (key == "write_matrix") ok = parse_bool(value, p.write_matrix);
    else if (key == "write_wcov") ok = parse_bool(value, p.write_wcov);
    else if (key == "bin_subsets") ok = parse_bool(value, p.bin_subsets);
    else if (key == "pixmode_map") ok = parse_int(value, p.pixmode_map);
    else if (key == "pixmode_cross") ok = parse_int(value, p.pixmode_cross);
    else if (key == "pixlim_map") ok = parse_real(value, p.pixlim_map);
    else if (key == "pixlim_cross") ok = parse_real(value, p.pixlim_cross);
    else if (key == "incomplete_matrices") ok = parse_bool(value, p.incomplete_matrices);
    else if (key == "allow_decoupling") ok = parse_bool(value, p.allow_decoupling);
    else if (key == "kfirst") ok = parse_bool(value, p.kfirst);
    else if (key == "basis_func") {
        std::string v4 = value.substr(0, 4);
        if (v4 != "poly" && v4 != "four" && v4 != "cheb" && v4 != "lege") fail("Unknown function basis : %s", value.c_str());
    } else if (key == "basis_order") {
        ok = parse_int(value, p.basis_order);
        if (ok && p.basis_order < 0) fail("basis order must be nonnegative");
    } else if (key == "iter_min") ok = parse_int(value, p.iter_min);
    else if (key == "iter_max") ok = parse_int(value, p.iter_max);
    else if (key == "cglimit") ok = parse_real(value, p.cglimit);
    else if (key == "fsample") ok = parse_real(value, p.fsample);
    else if (key == "mode_detweight") ok = parse_int(value, p.mode_detweight);
    else if (key == "temperature_only") ok = parse_bool(value, p.temperature_only);
    else if (key == "noise_weights_from_psd") ok = parse_bool(value, p.noise_weights_from_psd);
    else if (key == "radiometers") ok = parse_bool(value, p.radiometers);
    else if (key == "psdlen" || key == "psd_down" || key == "psd_downsample" || key == "nwrite_binary")
        ok = parse_int(value, ig);
    else if (key == "kfilter") ok = parse_bool(value, p.kfilter);
    else if (key == "diagfilter") ok = parse_real(value, p.diagfilter);
    else if (key == "precond_width_min") ok = parse_int(value, p.precond_width_min);
    else if (key == "precond_width_max") ok = parse_int(value, p.precond_width_max);
    else if (key == "use_fprecond") ok = parse_bool(value, p.use_fprecond);
    else if (key == "use_cgprecond") ok = parse_bool(value, p.use_cgprecond);
    else if (key == "rm_monopole") ok = parse_bool(value, p.rm_monopole);
    else if (key == "path_output") p.path_output = value;
    else if (key == "file_root") p.file_root = value;
    else if (key == "unit_tod") p.unit_tod = value;
    else if (key == "file_gap_out" || key == "file_mc" || key == "file_gap") {
    } else if (key == "write_binmap") ok = parse_bool(value, p.write_binmap);
    else if (key == "write_hits") ok = parse_bool(value, p.write_hits);
    else if (key == "write_mask") ok = parse_bool(value, p.write_mask);
    else if (key == "write_tod") ok = parse_bool(value, p.write_tod);
    else if (key == "file_inmask") p.has_inmask = !value.empty();
    else if (key == "file_spectrum") p.has_spectrum = !value.empty();
    else if (key == "file_covmat") p.has_covmat = !value.empty();
    else if (key == "detset" || key == "detset_nopol") {
        // parse_detset (src/inputparam.f90:375-434): an empty or "all" member names the full set
        size_t c = value.find(':');
        if (c == std::string::npos) fail("Failed to parse: %s for valid detector set name", value.c_str());
        Params::DetSet ds;
        ds.name = strip(value.substr(0, c));
        if (ds.name.empty()) fail("Empty detset name");
        ds.nopol = key == "detset_nopol";
        std::string rest = strip(value.substr(c + 1));
        bool full = false;
        for (;;) {
            size_t comma = rest.find(',');
            std::string d = strip(comma == std::string::npos ? rest : rest.substr(0, comma));
            if (d.empty() || is_all(d)) {
                full = true;
                break;
            }
            ds.dets.push_back(d);
            if (comma == std::string::npos) break;
            rest = strip(rest.substr(comma + 1));
        }
        if (full) p.detset0 = ds.name;
        else p.detsets.push_back(ds);
        p.ndetset = (int)p.detsets.size();
    } else if (key == "survey") {
        // parse_survey (src/inputparam.f90:438-487): "name : start - stop, start - stop" or "name : all"
        size_t c = value.find(':');
        if (c == std::string::npos) fail("Failed to parse: %s for valid survey name", value.c_str());
        Params::Survey sv;
        sv.name = strip(value.substr(0, c));
        if (sv.name.empty()) fail("Empty survey name");
        std::string rest = strip(value.substr(c + 1));
        bool full = false;
        for (;;) {
            size_t dash = rest.find('-');
            if (dash == std::string::npos) {
                if (rest.empty() || is_all(rest)) {
                    full = true;
                    break;
                }
                fail("Failed to parse: %s for valid survey span marker (-)", rest.c_str());
            }
            double a = 0, b = 0;
            if (!parse_real(strip(rest.substr(0, dash)), a)) fail("Failed to parse: %s for valid survey start", rest.c_str());
            rest = strip(rest.substr(dash + 1));
            size_t comma = rest.find(',');
            if (!parse_real(strip(comma == std::string::npos ? rest : rest.substr(0, comma)), b))
                fail("Failed to parse: %s for valid survey stop", rest.c_str());
            sv.starts.push_back(a);
            sv.stops.push_back(b);
            if (comma == std::string::npos) break;
            rest = strip(rest.substr(comma + 1));
        }
        if (full) p.survey0 = sv.name;
        else p.surveys.push_back(sv);
        p.nsurvey = (int)p.surveys.size();
    } else if (key == "mcmode") ok = parse_bool(value, p.mcmode);
    else fail("Unknown parameter: %s", key.c_str());
    if (!ok) fail("Failed to parse %s from %s", key.c_str(), value.c_str());
}

// src/inputparam.f90:24-67, 500-505
static Params read_parameters(const char *parstring)
{
    Params p;
    std::string s(parstring ? parstring : "");
    size_t pos = 0;
    while (pos < s.size()) {
        size_t e = s.find(';', pos);
        std::string line = s.substr(pos, e == std::string::npos ? std::string::npos : e - pos);
        read_line(p, line);
        if (e == std::string::npos) break;
        pos = e + 1;
    }
    p.dnshort = (double)llround(p.dnshort * p.fsample);
    if (p.kfirst && p.dnshort < 0) p.dnshort = F32(1e10);
    return p;
}

// ------------------------------------------------------------------ detectors and plan


static void interpolate_psd(const std::vector<double> &freq, const double *psd, const double *tf, int64_t n_out,
                            double *tp)
{
    // src/noise_routines.f90:130-190 (host copy for the plateau and noise weights)
    int64_t n_in = (int64_t)freq.size();
    std::vector<double> lf(n_in), lp(n_in);
    int64_t startbin = 0;
    while (startbin < n_in && freq[startbin] == 0) startbin++;
    for (int64_t i = 0; i < n_in; ++i) { lf[i] = std::log(freq[i]); lp[i] = std::log(psd[i]); }
    int64_t j = startbin;
    for (int64_t i = 0; i < n_out; ++i) {
        if (tf[i] == 0) { tp[i] = 0; continue; }
        double flog = std::log(tf[i]);
        if (j < n_in - 2) {
            while (lf[j + 1] < flog) {
                j++;
                if (j == n_in - 2) break;
            }
        }
        if (psd[j] > 0 && psd[j + 1] > 0) {
            double r = (flog - lf[j]) / (lf[j + 1] - lf[j]);
            tp[i] = std::exp(lp[j] * (1.0 - r) + lp[j + 1] * r);
        } else {
            double r = (tf[i] - freq[j]) / (freq[j + 1] - freq[j]);
            tp[i] = psd[j] * (1.0 - r) + psd[j + 1] * r;
        }
    }
}





static int psd_index_det(const Detector &d, double t)
{
    // src/noise_routines.f90:211-236
    if (!d.has_psd && d.npsd == 1) return 0;
    int ipsd = 0;
    for (;;) {
        if (ipsd == d.npsd - 1) return d.psdstarts[ipsd] > t ? -1 : ipsd;
        if (d.psdstarts[ipsd + 1] > t) return ipsd;
        ipsd++;
    }
}

static int psd_index(const Context &c, int idet, double t)
{
    // src/noise_routines.f90:193-208
    for (int i = 0; i < c.npsdtot; ++i)
        if (c.psddet[i] == idet && c.psdstart[i] <= t && c.psdstop[i] >= t) return i;
    return -1;
}

void ensure_stream(Context &c)
{
    if (!c.stream) {
        MADAM_CHECK(hipSetDevice(c.device));
        MADAM_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    }
}


// ------------------------------------------------------------------ configure

static void configure(Context &c, const char *parstring, long ndet, const char *detstring, const double *detweights,
                      long nsamp, long nnz, const double *timestamps, long nperiod, const long *periods,
                      const long *npsd, long npsdtot, const double *psdstarts, long npsdbin, const double *psdfreqs,
                      long npsdval, const double *psdvals, int pass)
{
    c.timestamps.assign(timestamps, timestamps + (nsamp > 0 ? nsamp : 0));
    c.det_names.clear();
    {
        std::string ds = detstring ? detstring : "";
        size_t a = 0;
        while (a <= ds.size() && (long)c.det_names.size() < ndet) {
            size_t b = ds.find(';', a);   // names separated by ';' (dets2detstring)
            if (b == std::string::npos) b = ds.size();
            c.det_names.push_back(strip(ds.substr(a, b - a)));
            a = b + 1;
        }
        while ((long)c.det_names.size() < ndet) c.det_names.push_back("det" + std::to_string(c.det_names.size()));
    }
    (void)npsdval;
    ensure_stream(c);
    Params p = read_parameters(parstring);
    // src/smadam.F90:138-144
    if (p.mcmode && p.nsubchunk > 1) fail("MCMode is not compatible with nsubchunk > 1.");
    if (p.temperature_only && nnz != 1) fail("temperature_only=T but pointing weights are polarized");
    if (ndet > 10000) fail("ndet > NDETMAX");
    if (nnz != 1 && nnz != 3) fail("nnz=%ld is not supported (nmap must be 1 or 3)", nnz);
    if (p.has_inmask || p.has_spectrum || p.has_covmat) fail("file_inmask / file_spectrum / file_covmat are not supported");
    if ((uint64_t)ndet * (uint64_t)nsamp >= (1ULL << 32)) fail("ndet * nsamp must be below 2^32");
    c.ndet = (int)ndet;
    c.nsamp = nsamp;
    c.nnz = (int)nnz;
    c.nmap = (int)nnz;
    c.pass = pass;
    c.psdfreqs.assign(psdfreqs, psdfreqs + npsdbin);
    // read_detectors
    c.det.assign(ndet, Detector());
    long ipsd = 0;
    for (long d = 0; d < ndet; ++d) {
        Detector &D = c.det[d];
        int nint = (int)npsd[d];
        D.npsd = nint == 0 ? 1 : nint;
        double w = detweights[d];
        D.weights.assign(D.npsd, w);
        D.sigmas.assign(D.npsd, w > 0 ? 1.0 / std::sqrt(w) : 0.0);
        D.plateaus.resize(D.npsd);
        for (int j = 0; j < D.npsd; ++j) D.plateaus[j] = D.sigmas[j] * D.sigmas[j] / p.fsample;
        D.psdstarts.assign(D.npsd, 0.0);
        if (nint > 0) {
            D.has_psd = true;
            D.psds.resize((size_t)nint * npsdbin);
            for (int j = 0; j < nint; ++j) {
                D.psdstarts[j] = psdstarts[ipsd];
                std::copy(psdvals + ipsd * npsdbin, psdvals + (ipsd + 1) * npsdbin, D.psds.begin() + (size_t)j * npsdbin);
                ipsd++;
            }
        }
    }
    (void)npsdtot;
    // intervals (check_files) and parameters (init_parameters)
    c.nival = (int)nperiod;
    c.ivs.assign(nperiod + 1, 0);
    std::vector<int64_t> len(nperiod);
    for (long i = 0; i < nperiod; ++i) {
        len[i] = i < nperiod - 1 ? periods[i + 1] - periods[i] : nsamp - periods[i];
        c.ivs[i + 1] = c.ivs[i] + len[i];
    }
    if (nperiod < 1) fail("ERROR: intervals not known.");
    if (!p.radiometers)
        for (auto &D : c.det) {
            if (!D.has_psd) continue;
            for (int j = 0; j < D.npsd; ++j) {
                double *psd = D.psds.data() + (size_t)j * npsdbin;
                long im = 0;
                for (long i = 1; i < npsdbin; ++i) if (psd[i] > psd[im]) im = i;
                for (long i = 0; i < im; ++i) psd[i] = psd[im];
            }
        }
    if (p.noise_weights_from_psd) {
        for (auto &D : c.det) {
            if (!D.has_psd) fail("noise_weights_from_psd needs PSDs for every detector");
            for (int j = 0; j < D.npsd; ++j) {
                const double *psd = D.psds.data() + (size_t)j * npsdbin;
                if (p.radiometers) {
                    double fr = p.fsample / 2, pl;
                    interpolate_psd(c.psdfreqs, psd, &fr, 1, &pl);
                    D.sigmas[j] = std::sqrt(pl * p.fsample);
                    D.plateaus[j] = pl;
                } else {
                    int nbin = 10000;
                    double fstep = p.fsample * .5 / nbin;
                    std::vector<double> fr(nbin), dd(nbin);
                    for (int i = 0; i < nbin; ++i) fr[i] = fstep * (i + 1);
                    double fbase = 1 / (p.dnshort / p.fsample);
                    int ibase = (int)(fbase / fstep), ilower = (int)(1 / fstep), iupper = (int)(10 / fstep);
                    interpolate_psd(c.psdfreqs, psd, fr.data(), nbin, dd.data());
                    double pl = dd[ilower - 1];
                    for (int i = ilower; i < iupper && i < nbin; ++i) pl = std::min(pl, dd[i]);
                    double rms = ibase * pl;
                    for (int i = ibase; i < nbin; ++i) rms += dd[i];
                    D.sigmas[j] = std::sqrt(rms / nbin * p.fsample);
                    D.plateaus[j] = pl;
                    D.weights[j] = 1 / (pl * p.fsample);
                }
            }
        }
    }
    if (p.nsubchunk < 2) { p.nsubchunk = 0; p.isubchunk = 0; }
    if (p.isubchunk > p.nsubchunk) fail("ERROR: isubchunk > nsubchunk");
    if (pass < 0 || pass > p.nsubchunk) fail("subchunk pass %d out of range", pass);
    if (p.good_baseline_fraction < 0 || p.good_baseline_fraction > 1) fail("ERROR: good_baseline_fraction must be in 0..1");
    if (p.nside_cross < 0) p.nside_cross = p.nside_map;
    p.nside_submap = std::min(p.nside_submap, std::min(p.nside_map, p.nside_cross));   // parameter_control.f90:253
    for (int ns : {p.nside_map, p.nside_submap, p.nside_cross})
        if (ns < 1 || (ns & (ns - 1))) fail("nside must be a power of two");
    if (p.nside_map < p.nside_cross) fail("nside_cross > nside_map not supported by Madam/TOAST");
    if (p.pixlim_cross < 0) p.pixlim_cross = p.pixlim_map;
    if (p.pixmode_map < 0 || p.pixmode_map > 2) fail("Pixmode_map must be in range [0-2]");
    if (p.pixmode_cross < 0 || p.pixmode_cross > 4) fail("Pixmode_cross must be in range [0-4]");
    if (p.basis_order != 0) fail("basis_order > 0 is not supported");
    if (c.nmap == 1) p.temperature_only = 1;
    for (auto &D : c.det)
        for (int j = 0; j < D.npsd; ++j) {
            if (p.mode_detweight == 0) D.weights[j] = D.sigmas[j] == 0 ? 0 : 1.0 / D.plateaus[j] / p.fsample;
            else if (p.mode_detweight == 1) D.weights[j] = 1;
            else fail("mode_detweight=%d is not supported", p.mode_detweight);
        }
    c.nside_max = std::max(p.nside_map, p.nside_cross);
    c.npix_map = 12LL * p.nside_map * p.nside_map;
    c.npix_cross = 12LL * p.nside_cross * p.nside_cross;
    c.ndeg = c.npix_map / c.npix_cross;
    if (c.ndeg < 1 || (c.ndeg & (c.ndeg - 1)) != 0)
        fail("nside_cross must not exceed nside_map, and both must be powers of two");
    p.nshort = (int)(p.dnshort + .5);
    if (!p.kfirst) {
        p.dnshort = 1e6;
        p.nshort = (int)(p.dnshort + .5);
        p.precond_width_min = 0;
        p.precond_width_max = 0;
        p.kfilter = 0;
    } else if (p.precond_width_max < p.precond_width_min) p.precond_width_max = p.precond_width_min;
    c.noba_pp.resize(nperiod);
    c.kstart_pp.resize(nperiod);
    c.nbase = 0;
    c.noba_pp_max = 0;
    for (long i = 0; i < nperiod; ++i) {
        c.noba_pp[i] = (int64_t)((len[i] - 1) / p.dnshort + 1);
        c.kstart_pp[i] = c.nbase;
        c.nbase += c.noba_pp[i];
        c.noba_pp_max = std::max(c.noba_pp_max, c.noba_pp[i]);
    }
    // start_timeloop: baseline lengths (src/parameter_control.f90:787-818)
    c.bstart.assign(c.nbase + 1, 0);
    c.btime.assign(c.nbase, 0.0);
    {
        std::vector<int64_t> bl(c.nbase);
        int64_t m = 0;
        for (long i = 0; i < nperiod; ++i) {
            double dn = 0.0;
            int64_t n = 0;
            for (int64_t k = 0; k < c.noba_pp[i] - 1; ++k) {
                int64_t n0 = n;
                dn = dn + p.dnshort;
                n = (int64_t)(dn + .5);
                bl[m++] = n - n0;
            }
            bl[m++] = len[i] - n;
            if (p.kfirst && (bl[m - 1] < 0 || bl[m - 1] > (int64_t)p.dnshort + 1))
                fail("ERROR in start_timeloop: Lengths do not match.");
        }
        for (int64_t k = 0; k < c.nbase; ++k) {
            c.bstart[k + 1] = c.bstart[k] + bl[k];
            c.btime[k] = (c.bstart[k] >= 0 && c.bstart[k] < nsamp) ? timestamps[c.bstart[k]] : 0.0;
        }
    }
    c.nbtot = (int64_t)c.ndet * c.nbase;
    // per (det, interval) PSD indices and weights
    c.dw_ival.assign((size_t)ndet * nperiod, 0.0);
    c.ipsd_det_ival.assign((size_t)ndet * nperiod, -1);
    for (long d = 0; d < ndet; ++d)
        for (long i = 0; i < nperiod; ++i) {
            int ip = c.noba_pp[i] > 0 ? psd_index_det(c.det[d], c.btime[c.kstart_pp[i]]) : -1;
            c.ipsd_det_ival[d * nperiod + i] = ip;
            c.dw_ival[d * nperiod + i] = ip < 0 ? 0.0 : c.det[d].weights[ip];
        }
    // filter tables (init_filter / build_filter bookkeeping)
    c.npsdtot = 0;
    c.psddet.clear();
    c.psdstart.clear();
    c.psdstop.clear();
    c.psd_w.clear();
    c.nof = 0;
    c.p = p;
    if (p.kfilter) {
        int64_t nof_min = 2 * c.noba_pp_max;
        if (p.precond_width_max > nof_min - 1) nof_min = p.precond_width_max + 1;
        int64_t nof = 1;
        while (nof < nof_min) nof *= 2;
        if (nof < 4) nof = 4;
        if (nof > (1LL << 22)) fail("FFT length %lld exceeds the supported 2^22", (long long)nof);
        c.nof = nof;
        for (long d = 0; d < ndet; ++d) {
            if (!c.det[d].has_psd) fail("All detectors must have at least one PSD to build the filter");
            for (int j = 0; j < c.det[d].npsd; ++j) {
                c.psddet.push_back((int)d);
                c.psdstart.push_back(c.det[d].psdstarts[j]);
                c.psdstop.push_back(j < c.det[d].npsd - 1 ? c.det[d].psdstarts[j + 1] : F32(1e30));
                c.psd_w.push_back(c.det[d].weights[j]);
            }
        }
        c.npsdtot = (int)c.psddet.size();
    }
    c.ipsd_glob_ival.assign((size_t)ndet * nperiod, -1);
    if (p.kfilter)
        for (long d = 0; d < ndet; ++d)
            for (long i = 0; i < nperiod; ++i)
                if (c.noba_pp[i] > 0) c.ipsd_glob_ival[d * nperiod + i] = psd_index(c, (int)d, c.btime[c.kstart_pp[i]]);
    // device plan arrays
    hipStream_t s = c.stream;
    c.d_ivs.alloc(c.ivs.size());
    h2d(c.d_ivs.p, c.ivs.data(), c.ivs.size(), s);
    c.d_bstart.alloc(c.bstart.size());
    h2d(c.d_bstart.p, c.bstart.data(), c.bstart.size(), s);
    c.d_kstart_pp.alloc(nperiod);
    h2d(c.d_kstart_pp.p, c.kstart_pp.data(), nperiod, s);
    c.d_noba_pp.alloc(nperiod);
    h2d(c.d_noba_pp.p, c.noba_pp.data(), nperiod, s);
    c.d_dw_ival.alloc(c.dw_ival.size());
    h2d(c.d_dw_ival.p, c.dw_ival.data(), c.dw_ival.size(), s);
    {
        std::vector<double> dwb(c.nbtot), sdwb(c.nbtot), wbinv(c.nbtot);
        std::vector<int8_t> has(c.nbtot);
        for (int d = 0; d < c.ndet; ++d)
            for (int i = 0; i < c.nival; ++i) {
                double w = c.dw_ival[(size_t)d * c.nival + i];
                int ip = c.ipsd_det_ival[(size_t)d * c.nival + i];
                for (int64_t k = c.kstart_pp[i]; k < c.kstart_pp[i] + c.noba_pp[i]; ++k) {
                    int64_t b = (int64_t)d * c.nbase + k;
                    dwb[b] = w;
                    sdwb[b] = std::sqrt(w);
                    has[b] = ip >= 0;
                    wbinv[b] = ip >= 0 ? 1. / c.det[d].weights[ip] : 1.0;
                }
            }
        c.d_dwb.alloc(c.nbtot);
        c.d_sdwb.alloc(c.nbtot);
        c.d_wbinv.alloc(c.nbtot);
        c.d_haspsd.alloc(c.nbtot);
        h2d(c.d_dwb.p, dwb.data(), c.nbtot, s);
        h2d(c.d_sdwb.p, sdwb.data(), c.nbtot, s);
        h2d(c.d_wbinv.p, wbinv.data(), c.nbtot, s);
        h2d(c.d_haspsd.p, has.data(), c.nbtot, s);
        MADAM_CHECK(hipStreamSynchronize(s));
    }
    c.configured = true;
    c.pointed = false;
    c.solved = false;
    c.prec_cached = false;
}

// ------------------------------------------------------------------ noise filter on the device

static void build_filters(Context &c)
{
    hipStream_t s = c.stream;
    const Params &p = c.p;
    int64_t nh = c.nof / 2 + 1;
    int64_t n_in = (int64_t)c.psdfreqs.size();
    c.d_fcov.alloc((size_t)nh * c.npsdtot);
    c.d_tab.alloc(4 * n_in);
    DBuf<int> bad;
    bad.alloc(1);
    MADAM_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    int64_t startbin = 0;
    while (startbin < n_in && c.psdfreqs[startbin] == 0) startbin++;
    std::vector<double> tab(4 * n_in);
    int ip = 0;
    for (int d = 0; d < c.ndet; ++d) {
        Detector &D = c.det[d];
        for (int j = 0; j < D.npsd; ++j, ++ip) {
            const double *psd = D.psds.data() + (size_t)j * n_in;
            double sigma = D.sigmas[j];
            double plateau;
            if (p.noise_weights_from_psd) plateau = D.plateaus[j];
            else {
                if (p.radiometers) {
                    double fr = p.fsample / 2;
                    interpolate_psd(c.psdfreqs, psd, &fr, 1, &plateau);
                } else {
                    std::vector<double> fr(1000), dd(1000);
                    for (int i = 0; i < 1000; ++i) fr[i] = 1 + (double)(i * 10) / 1000;
                    interpolate_psd(c.psdfreqs, psd, fr.data(), 1000, dd.data());
                    plateau = *std::min_element(dd.begin(), dd.end());
                }
                D.sigmas[j] = std::sqrt(plateau * p.fsample);
            }
            for (int64_t i = 0; i < n_in; ++i) {
                tab[i] = c.psdfreqs[i];
                tab[n_in + i] = psd[i];
                tab[2 * n_in + i] = std::log(c.psdfreqs[i]);
                tab[3 * n_in + i] = std::log(psd[i]);
            }
            h2d(c.d_tab.p, tab.data(), tab.size(), s);
            launch_build_filter(s, c.nof, p.fsample, p.nshort, p.radiometers, sigma, plateau, n_in, startbin, c.d_tab.p,
                                c.d_fcov.p + (size_t)ip * nh, bad.p);
            MADAM_CHECK(hipGetLastError());
            MADAM_CHECK(hipStreamSynchronize(s));   // tab is reused
        }
    }
    int hbad = 0;
    d2h(&hbad, bad.p, 1, s);
    MADAM_CHECK(hipStreamSynchronize(s));
    if (hbad) fail("non-positive PSD values are not supported by the noise filter");
    if (!c.fft_ready || c.fft.N != c.nof / 2) {
        if (c.fft_ready) fft_plan_free(c.fft);
        fft_plan_init(c.fft, c.nof, s);
        c.fft_ready = true;
    }
    c.d_fcov_rows.alloc((size_t)nh * c.npsdtot);
    fft_spectrum_rows(s, c.fft, c.d_fcov.p, c.npsdtot, nh, c.d_fcov_rows.p);
}

// ------------------------------------------------------------------ pointing

// dummy_is_flag: pixel 12 * nside_max**2, which a previous call leaves in the
// caller's array for its flagged samples, flags the sample again (the cached
// path, src/smadam.F90:713-715)
static void set_pointing(Context &c, int32_t *pix, float *w, int on_device, bool dummy_is_flag = false)
{
    if (!c.configured) fail("madam_set_pointing called before configure");
    hipStream_t s = c.stream;
    double t0 = now();
    int64_t n = (int64_t)c.ndet * c.nsamp;
    if (on_device) {
        c.d_pix = pix;
        c.d_w = w;
    } else {
        c.d_pix_own.alloc(n);
        c.d_w_own.alloc((size_t)n * c.nnz);
        h2d(c.d_pix_own.p, pix, n, s);
        h2d(c.d_w_own.p, w, (size_t)n * c.nnz, s);
        c.d_pix = c.d_pix_own.p;
        c.d_w = c.d_w_own.p;
    }
    if (dummy_is_flag) launch_dummy_to_flag(s, c.d_pix, n, (int32_t)(12LL * c.nside_max * c.nside_max));
    // pixel range check (src/smadam.F90:158-171)
    if (n > 0) {
        DBuf<int> mm;
        mm.alloc(2);
        int init[2] = {INT32_MAX, INT32_MIN};
        h2d(mm.p, init, 2, s);
        launch_minmax(s, c.d_pix, n, mm.p, mm.p + 1);
        int hm[2];
        d2h(hm, mm.p, 2, s);
        MADAM_CHECK(hipStreamSynchronize(s));
        if (hm[0] < -1) fail("Too small pixel number %d", hm[0]);
        if ((int64_t)hm[1] > 12LL * c.nside_max * c.nside_max - 1) fail("Too large pixel number %d", hm[1]);
    }
    // check_files flagging (src/read_data.f90:241-272)
    DBuf<unsigned long long> cnt;
    cnt.alloc((size_t)c.ndet * c.nival + 1);
    MADAM_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long) * ((size_t)c.ndet * c.nival + 1), s));
    launch_flag_and_count(s, c.d_pix, c.d_w, c.nnz, c.ndet, c.nsamp, c.d_ivs.p, c.nival, cnt.p,
                          cnt.p + (size_t)c.ndet * c.nival);
    std::vector<unsigned long long> valid((size_t)c.ndet * c.nival + 1);
    d2h(valid.data(), cnt.p, valid.size(), s);
    MADAM_CHECK(hipStreamSynchronize(s));
    std::vector<int8_t> flag((size_t)c.ndet * c.nival, 0);
    bool anyflag = false;
    for (int d = 0; d < c.ndet; ++d)
        for (int i = 0; i < c.nival; ++i) {
            unsigned long long ng = valid[(size_t)d * c.nival + i];
            int64_t len = c.ivs[i + 1] - c.ivs[i];
            if (ng == 0 || len <= 0) continue;
            double frac = (double)ng / len;
            if (frac < F32(1e-3)) { flag[(size_t)d * c.nival + i] = 1; anyflag = true; }
        }
    if (anyflag) {
        DBuf<int8_t> df;
        df.alloc(flag.size());
        h2d(df.p, flag.data(), flag.size(), s);
        launch_flag_intervals(s, c.d_pix, c.nsamp, c.d_ivs.p, c.nival, c.ndet, df.p);
        MADAM_CHECK(hipStreamSynchronize(s));
    }
    c.pix_host_dirty = true;
    // pack and build the CSR by map pixel
    // every record array carries REC_PAD zero records past its end: the batched
    // projections read whole chunks of records without clamping their index
    c.d_pixw.alloc(n + REC_PAD);
    MADAM_CHECK(hipMemsetAsync(c.d_pixw.p + n, 0, sizeof(PixW) * REC_PAD, s));
    // compact records for the projections: temperature-only projections use no
    // weight (src/smadam_routines.f90:1038-1070), polarised ones drop w0 when
    // it is 1 for every active sample; the packing pass writes them, and the
    // CSR sort keys, beside the full records
    int rl = c.nmap == 1 ? REC4 : REC12;
    c.d_pixc.alloc((size_t)(n + REC_PAD) * rec_bytes(rl));
    MADAM_CHECK(hipMemsetAsync(c.d_pixc.p + (size_t)n * rec_bytes(rl), 0, (size_t)REC_PAD * rec_bytes(rl), s));
    CsrSort cs;
    csr_alloc(cs, n, c.npix_map);
    {
        DBuf<unsigned long long> cnt;
        cnt.alloc(1);
        MADAM_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), s));
        launch_pack(s, c.d_pix, c.d_w, c.nnz, c.nmap, c.ndet, c.nsamp, c.d_ivs.p, c.nival, c.d_dw_ival.p,
                    c.p.nsubchunk, c.pass, c.d_pixw.p, cs.k[0].p, (uint32_t)c.npix_map, c.d_pixc.p, cnt.p);
        MADAM_CHECK(hipGetLastError());
        unsigned long long h = 0;
        d2h(&h, cnt.p, 1, s);
        MADAM_CHECK(hipStreamSynchronize(s));
        if (c.nmap == 3 && h != 0) rl = REC16;
    }
    c.rec_layout = rl;
    c.stats[ST_REC_BYTES] = rec_bytes(rl);
    if (rl == REC16) c.d_pixc.release();
    // the CSR by map pixel: a stable sort of the samples by pixel, the hit
    // counts from its runs, the offsets from the counts
    csr_sort(s, cs);
    c.d_hits.alloc(c.npix_map);
    csr_hits(s, cs, c.d_hits.p);
    c.d_off.alloc(c.npix_map + 1);
    int64_t nchunks = (c.npix_map + 1 + 4095) / 4096;
    c.d_scan_tmp.alloc(nchunks);
    launch_exclusive_scan_i32_to_i64(s, c.d_hits.p, c.npix_map, c.d_off.p, c.d_scan_tmp.p, nchunks);
    int64_t total = 0;
    d2h(&total, c.d_off.p + c.npix_map, 1, s);
    MADAM_CHECK(hipStreamSynchronize(s));
    c.nactive = total;
    // the fill reads only the sorted sample indices
    cs.k[0].release();
    cs.k[1].release();
    cs.v[cs.cur ^ 1].release();
    c.d_ent.alloc(total + REC_PAD);
    MADAM_CHECK(hipMemsetAsync(c.d_ent.p + total, 0, sizeof(c.d_ent.p[0]) * REC_PAD, s));
    c.d_src.alloc(total > 0 ? total : 1);
    c.d_bord.alloc(std::max<int64_t>(c.nbtot, 1));
    c.d_bslot.alloc(std::max<int64_t>(c.nbtot, 1));
    baseline_slots(s, c.d_pixw.p, c.ndet, c.nsamp, c.d_bstart.p, c.nbase, c.npix_map, c.d_bord.p, c.d_bslot.p);
    if (rl != REC16) {
        c.d_entc.alloc((size_t)(total + REC_PAD) * rec_bytes(rl));
        MADAM_CHECK(hipMemsetAsync(c.d_entc.p + (size_t)total * rec_bytes(rl), 0, (size_t)REC_PAD * rec_bytes(rl), s));
    } else {
        c.d_entc.release();
    }
    csr_fill(s, cs, c.d_pixw.p, c.nsamp, c.d_bstart.p, c.nbase, total, c.d_bslot.p, rl, c.d_ent.p, c.d_src.p,
             rl != REC16 ? (void *)c.d_entc.p : nullptr);
    MADAM_CHECK(hipGetLastError());
    MADAM_CHECK(hipStreamSynchronize(s));
    c.stats[ST_T_POINTING] = now() - t0;
    c.stats[ST_NACTIVE] = (double)total;
    c.pointed = true;
    c.tiles_built = false;
    c.prec_cached = false;
    c.stats[ST_POINTING_REUSED] = 0;
    batch_invalidate(c);
}

// ------------------------------------------------------------------ solve

static void segtiles_alloc(Context::SegTiles &t, int64_t nent)
{
    int64_t nt = segred_ntiles(nent);
    t.nent = nent;
    t.heads.alloc(segred_heads_bytes(nent));
    t.k0.alloc(nt > 0 ? nt : 1);
    t.fix_end.alloc(nt > 0 ? nt : 1);
    t.fix_ord.alloc(nt > 0 ? nt : 1);
}

// segment tables of the two projections; they depend on the pointing only
static void build_segtiles(Context &c)
{
    hipStream_t s = c.stream;
    // pixel side: non-empty cross pixels of the CSR
    {
        int64_t npc = c.npix_cross;
        DBuf<int32_t> flag;
        DBuf<int64_t> pos, tmp;
        DBuf<int64_t> &segstart = c.d_segstart_s;
        flag.alloc(npc);
        pos.alloc(npc + 1);
        int64_t nchunks = (npc + 1 + 4095) / 4096;
        tmp.alloc(nchunks);
        c.d_nzq.alloc(npc);
        segstart.alloc(npc + 1);
        segred_pixel_segments(s, c.d_off.p, npc, c.ndeg, flag.p, pos.p, tmp.p, nchunks, c.d_nzq.p, segstart.p);
        int64_t last_pos = 0;
        int32_t last_flag = 0;
        d2h(&last_pos, pos.p + npc - 1, 1, s);
        d2h(&last_flag, flag.p + npc - 1, 1, s);
        MADAM_CHECK(hipStreamSynchronize(s));
        int64_t nseg = last_pos + last_flag;
        c.nzpix = nseg;
        c.stats[ST_NPIX_HIT] = (double)nseg;
        segtiles_alloc(c.sr_s, c.nactive);
        segred_build(s, segstart.p, nseg, c.nactive, c.sr_s.heads.p, c.sr_s.k0.p, c.sr_s.fix_end.p,
                     c.sr_s.fix_ord.p);
        MADAM_CHECK(hipStreamSynchronize(s));
        c.sr_s.ok = true;
    }
    // baseline side: every baseline is a non-empty sample range covering the stream
    {
        bool ok = c.bstart.front() == 0 && c.bstart.back() == c.nsamp;
        for (int64_t k = 0; k < c.nbase && ok; ++k) ok = c.bstart[k + 1] > c.bstart[k];
        c.sr_g.ok = false;
        if (ok) {
            int64_t nent = (int64_t)c.ndet * c.nsamp;
            DBuf<int64_t> segstart;
            segstart.alloc(c.nbtot + 1);
            segred_baseline_segments(s, c.d_bstart.p, c.nbase, c.ndet, c.nsamp, segstart.p);
            segtiles_alloc(c.sr_g, nent);
            segred_build(s, segstart.p, c.nbtot, nent, c.sr_g.heads.p, c.sr_g.k0.p, c.sr_g.fix_end.p,
                         c.sr_g.fix_ord.p);
            MADAM_CHECK(hipStreamSynchronize(s));
            c.sr_g.ok = true;
        }
    }
    int64_t nt = std::max(segred_ntiles(c.nactive), segred_ntiles((int64_t)c.ndet * c.nsamp));
    c.d_sr_part.alloc(2 * (nt > 0 ? nt : 1));
}


int group_for(const Context &c)
{
    int64_t L = (int64_t)c.p.dnshort;
    int g = 4;
    while (g < 64 && g < L) g *= 2;
    return g;
}

static void apply_prior(Context &c, const double *x, double *y)
{
    // cinvmul: y = x outside segments, convolution inside (src/noise_routines.f90:745-913);
    // intervals without a PSD are left as y = x (convolve_interval returns early)
    if (!c.prior_covers_all) launch_copy(c.stream, y, x, c.nbtot);
    if (c.segconv.empty()) return;
    fft_convolve_segments(c.stream, c.fft, c.d_segconv.p, (int64_t)c.segconv.size(), x, y, c.d_fcov.p,
                          c.d_fcov_rows.p, c.nof / 2 + 1, c.d_fftwork.p, c.fft_batch, (int64_t)c.segconv.size(),
                          c.segconv.data());
}

// apply_CG_preconditioner (src/noise_routines.f90:1305-1409) on every segment
// listed in c.segcg; x is written into z on those segments
void cg_apply(Context &c, const double *r, double *z)
{
    int64_t ns = (int64_t)c.segcg.size();
    if (ns == 0) return;
    hipStream_t s = c.stream;
    int64_t nb = c.nbtot;
    for (auto *b : {&c.d_cg_resid, &c.d_cg_prec, &c.d_cg_zr, &c.d_cg_p, &c.d_cg_ap}) b->alloc(nb);
    c.d_cg_part.alloc((size_t)ns * 32);
    c.d_cg_sc.alloc((size_t)ns * 4);
    c.d_cg_act.alloc(ns);
    c.d_segact.alloc(ns);
    std::vector<int> act(ns, 1);
    h2d(c.d_cg_act.p, act.data(), ns, s);
    launch_cg_init(s, c.d_segcg.p, ns, c.d_cg_act.p, c.d_cg_ic0.p, r, c.d_nna.p, z, c.d_cg_resid.p, c.d_cg_prec.p,
                   c.d_cg_zr.p, c.d_cg_p.p);
    double *rr_d = c.d_cg_sc.p, *rz_d = c.d_cg_sc.p + ns, *an_d = c.d_cg_sc.p + 2 * ns, *ab_d = c.d_cg_sc.p + 3 * ns;
    launch_cg_dot(s, c.d_segcg.p, ns, c.d_cg_act.p, c.d_cg_resid.p, c.d_cg_resid.p, c.d_cg_part.p, rr_d);
    launch_cg_dot(s, c.d_segcg.p, ns, c.d_cg_act.p, c.d_cg_resid.p, c.d_cg_zr.p, c.d_cg_part.p, rz_d);
    std::vector<double> rr(ns), rr0(ns), rz(ns), an(ns), ab(ns);
    d2h(rr.data(), rr_d, ns, s);
    d2h(rz.data(), rz_d, ns, s);
    MADAM_CHECK(hipStreamSynchronize(s));
    for (int64_t i = 0; i < ns; ++i) {
        rr0[i] = rr[i];
        if (rr[i] == 0) act[i] = 0;   // x = 0
    }
    const double lim = F32(1e-6);
    for (int iter = 1;; ++iter) {
        std::vector<Segment> live;
        for (int64_t i = 0; i < ns; ++i)
            if (act[i]) live.push_back(c.segcg[i]);
        if (live.empty()) break;
        h2d(c.d_cg_act.p, act.data(), ns, s);
        h2d(c.d_segact.p, live.data(), live.size(), s);
        // ap = convolve(p) + nna p
        fft_convolve_segments(s, c.fft, c.d_segact.p, (int64_t)live.size(), c.d_cg_p.p, c.d_cg_ap.p, c.d_fcov.p,
                              c.d_fcov_rows.p, c.nof / 2 + 1, c.d_fftwork.p, c.fft_batch, c.fft_nseg_max,
                              live.data());
        launch_cg_nna(s, c.d_segcg.p, ns, c.d_cg_act.p, c.d_nna.p, c.d_cg_p.p, c.d_cg_ap.p);
        launch_cg_dot(s, c.d_segcg.p, ns, c.d_cg_act.p, c.d_cg_p.p, c.d_cg_ap.p, c.d_cg_part.p, an_d);
        d2h(an.data(), an_d, ns, s);
        MADAM_CHECK(hipStreamSynchronize(s));
        for (int64_t i = 0; i < ns; ++i) ab[i] = act[i] ? rz[i] / an[i] : 0.0;
        h2d(ab_d, ab.data(), ns, s);
        launch_cg_axpy(s, c.d_segcg.p, ns, c.d_cg_act.p, ab_d, z, c.d_cg_p.p, c.d_cg_resid.p, c.d_cg_ap.p);
        launch_cg_dot(s, c.d_segcg.p, ns, c.d_cg_act.p, c.d_cg_resid.p, c.d_cg_resid.p, c.d_cg_part.p, rr_d);
        d2h(rr.data(), rr_d, ns, s);
        MADAM_CHECK(hipStreamSynchronize(s));
        bool any = false;
        for (int64_t i = 0; i < ns; ++i) {
            if (!act[i]) continue;
            if (rr[i] / rr0[i] < lim || iter == 100) act[i] = 0;
            else any = true;
        }
        if (!any) break;
        h2d(c.d_cg_act.p, act.data(), ns, s);
        launch_cg_prec(s, c.d_segcg.p, ns, c.d_cg_act.p, c.d_cg_prec.p, c.d_cg_resid.p, c.d_cg_zr.p);
        std::vector<double> rzo = rz;
        launch_cg_dot(s, c.d_segcg.p, ns, c.d_cg_act.p, c.d_cg_resid.p, c.d_cg_zr.p, c.d_cg_part.p, rz_d);
        d2h(rz.data(), rz_d, ns, s);
        MADAM_CHECK(hipStreamSynchronize(s));
        for (int64_t i = 0; i < ns; ++i) ab[i] = act[i] ? rz[i] / rzo[i] : 0.0;
        h2d(ab_d, ab.data(), ns, s);
        launch_cg_p(s, c.d_segcg.p, ns, c.d_cg_act.p, ab_d, c.d_cg_zr.p, c.d_cg_p.p);
    }
    MADAM_CHECK(hipStreamSynchronize(s));
}

static void apply_precond(Context &c, const double *r, double *z)
{
    hipStream_t s = c.stream;
    switch (c.precond_kind) {
    case 0: launch_copy(s, z, r, c.nbtot); break;
    case 1: launch_diag_precond(s, r, c.d_prec.p, z, c.nbtot); break;
    case 2:
        launch_copy(s, z, r, c.nbtot);
        if (!c.segconv.empty())
            fft_convolve_segments(s, c.fft, c.d_segconv.p, (int64_t)c.segconv.size(), r, z, c.d_fprec.p,
                                  c.d_fprec_rows.p, c.nof / 2 + 1, c.d_fftwork.p, c.fft_batch,
                                  (int64_t)c.segconv.size(), c.segconv.data());
        break;
    default:
        if (!c.prec_covers_all) launch_copy(s, z, r, c.nbtot);
        c.band.chain_pack = band_chain_pack(c.kernel_variant);
        band_apply(s, c.band, r, z);
        cg_apply(c, r, z);
        break;
    }
}

static void build_segments(Context &c)
{
    hipStream_t s = c.stream;
    int64_t cap = (int64_t)c.ndet * c.nival * 4 + 1024;
    for (;;) {
        c.d_seg.alloc(cap);
        c.d_cnt.alloc(1);
        MADAM_CHECK(hipMemsetAsync(c.d_cnt.p, 0, sizeof(unsigned long long), s));
        int64_t gapmin = (int64_t)(60 * c.p.fsample / c.p.dnshort + 1);
        launch_find_segments(s, c.d_nna.p, c.ndet, c.nbase, c.d_kstart_pp.p, c.d_noba_pp.p, c.nival, gapmin, c.d_seg.p,
                             c.d_cnt.p, cap);
        unsigned long long ns = 0;
        d2h(&ns, c.d_cnt.p, 1, s);
        MADAM_CHECK(hipStreamSynchronize(s));
        if ((int64_t)ns > cap) { cap = (int64_t)ns; continue; }
        c.seg.resize(ns);
        d2h(c.seg.data(), c.d_seg.p, ns, s);
        MADAM_CHECK(hipStreamSynchronize(s));
        break;
    }
    std::sort(c.seg.begin(), c.seg.end(), [](const Segment &a, const Segment &b) { return a.b0 < b.b0; });
    for (auto &g : c.seg) {
        size_t di = (size_t)g.det * c.nival + g.ival;
        int64_t k0 = g.b0 - (int64_t)g.det * c.nbase;
        // convolve_interval looks the PSD up at the interval's first baseline
        g.ipsd = c.p.kfilter ? c.ipsd_glob_ival[di] : -1;
        int ipd = psd_index_det(c.det[g.det], c.btime[k0]);
        g.skip = (ipd < 0 || c.det[g.det].weights[ipd] == 0) ? 1 : 0;
        if (c.p.kfilter && psd_index(c, g.det, c.btime[k0]) < 0) g.skip = 1;
    }
    if (!c.seg.empty()) h2d(c.d_seg.p, c.seg.data(), c.seg.size(), s);
    // the preconditioner looks the PSD up at the trimmed segment's first
    // baseline (src/noise_routines.f90:1063, 1229), the convolution at the
    // interval's (src/noise_routines.f90:830)
    c.segp = c.seg;
    for (auto &g : c.segp) g.ipsd = c.p.kfilter ? psd_index(c, g.det, c.btime[g.b0 - (int64_t)g.det * c.nbase]) : -1;
    if (!c.segp.empty()) {
        c.d_segp.alloc(c.segp.size());
        h2d(c.d_segp.p, c.segp.data(), c.segp.size(), s);
    }
    c.segconv.clear();
    c.prior_covers_all = false;
    for (auto &g : c.seg)
        if (g.ipsd >= 0) c.segconv.push_back(g);
    {
        int64_t rows = 0;
        for (const Segment &g : c.segconv) rows += g.noba;
        c.prior_covers_all = rows == c.nbtot;   // segments are disjoint
    }
    if (!c.segconv.empty()) {
        c.d_segconv.alloc(c.segconv.size());
        h2d(c.d_segconv.p, c.segconv.data(), c.segconv.size(), s);
    }
    c.d_rmask.alloc(c.nbtot);
    MADAM_CHECK(hipMemsetAsync(c.d_rmask.p, 0, c.nbtot, s));
    launch_set_segments_mask(s, c.d_seg.p, (int64_t)c.seg.size(), c.d_rmask.p);
    MADAM_CHECK(hipStreamSynchronize(s));
    c.stats[ST_NSEG] = (double)c.seg.size();
}

// widest band (nband = kd + 1) of the band-Cholesky kernels' kd classes
static constexpr int BAND_WIDTH_GPU = 50;

static void build_preconditioner(Context &c)
{
    hipStream_t s = c.stream;
    const Params &p = c.p;
    if (c.band_ready) { band_free(c.band); c.band_ready = false; }
    c.prec_covers_all = false;
    if (p.precond_width_max < 1) { c.precond_kind = 0; return; }
    if (!p.kfilter || p.precond_width_max == 1) {
        c.precond_kind = 1;
        c.d_prec.alloc(c.nbtot);
        launch_prec_diag_build(s, c.d_nna.p, c.d_wbinv.p, c.d_haspsd.p, c.d_prec.p, c.nbtot);
        return;
    }
    int64_t nh = c.nof / 2 + 1;
    c.d_fprec.alloc((size_t)nh * c.npsdtot);
    {
        DBuf<double> wp;
        wp.alloc(c.npsdtot);
        h2d(wp.p, c.psd_w.data(), c.npsdtot, s);
        launch_fprecond(s, c.d_fcov.p, nh, wp.p, p.nshort, c.d_fprec.p, c.npsdtot);
        c.d_fprec_rows.alloc((size_t)nh * c.npsdtot);
        fft_spectrum_rows(s, c.fft, c.d_fprec.p, c.npsdtot, nh, c.d_fprec_rows.p);
        MADAM_CHECK(hipStreamSynchronize(s));
    }
    if (p.use_fprecond) { c.precond_kind = 2; return; }
    c.precond_kind = 3;
    // invcov = c2r(fcov) / nof per PSD
    c.d_invcov.alloc((size_t)c.nof * c.npsdtot);
    fft_c2r_rows(s, c.fft, c.d_fcov.p, c.npsdtot, c.d_invcov.p, nullptr);
    std::vector<double> ic((size_t)c.nof * c.npsdtot);
    d2h(ic.data(), c.d_invcov.p, ic.size(), s);
    MADAM_CHECK(hipStreamSynchronize(s));
    // band width per PSD and segment (src/noise_routines.f90:1071-1090)
    const double plim = F32(1 - 1e-5f);
    std::vector<int> nband(c.segp.size(), 0);
    std::vector<double> ptot(c.npsdtot, 0.0);
    for (int ip = 0; ip < c.npsdtot; ++ip)
        for (int64_t i = 0; i < c.nof / 2; ++i) ptot[ip] += ic[(size_t)ip * c.nof + i] * ic[(size_t)ip * c.nof + i];
    std::vector<int> tries(c.segp.size(), 1);
    std::vector<char> use_cg(c.segp.size(), 0);
    // Segments whose band would be wider than the factor kernels take
    // (BAND_WIDTH_GPU baselines) use the inner-CG preconditioner of the same
    // segment instead: another symmetric positive definite preconditioner, so
    // the PCG converges to the same solution, in a different number of steps.
    int64_t nwide = 0;
    for (size_t is = 0; is < c.segp.size(); ++is) {
        const Segment &g = c.segp[is];
        if (g.skip || g.ipsd < 0) continue;
        if (p.use_cgprecond) { use_cg[is] = 1; continue; }
        const double *v = ic.data() + (size_t)g.ipsd * c.nof;
        int tr = 1, nb = 0, ok = 0;
        for (;;) {
            nb = (int)std::min<int64_t>(g.noba, (int64_t)tr * p.precond_width_min);
            nb = std::min(nb, p.precond_width_max);
            double pp = 0;
            for (int i = 0; i < nb; ++i) pp += v[i] * v[i];
            if (pp > plim * ptot[g.ipsd]) { ok = 1; break; }
            if (nb == g.noba || nb == p.precond_width_max) break;
            if (tr == 1000) break;
            tr++;
        }
        if (!ok) { use_cg[is] = 1; continue; }   // src/noise_routines.f90:1113-1120
        if (nb > BAND_WIDTH_GPU) { use_cg[is] = 1; nwide++; continue; }
        nband[is] = nb;
        tries[is] = tr;
    }
    for (;;) {
        std::vector<int> info(c.segp.size(), 0);
        if (band_build(s, c.band, c.segp.data(), c.d_segp.p, (int64_t)c.segp.size(), nband.data(), c.d_invcov.p, c.nof,
                       c.d_nna.p, info.data(), !(c.kernel_variant & 16)) != 0)
            fail("band preconditioner wider than %d baselines reached the factor kernels", BAND_WIDTH_GPU);
        c.band_ready = true;
        bool retry = false;
        for (size_t is = 0; is < c.segp.size(); ++is) {
            if (info[is] == 0 || nband[is] == 0) continue;
            const Segment &g = c.segp[is];
            if (tries[is] == 1000 || nband[is] == g.noba || nband[is] == p.precond_width_max) {
                use_cg[is] = 1;   // Cholesky failed at the widest band: inner CG
                nband[is] = 0;
                continue;
            }
            tries[is]++;
            nband[is] = std::min(nband[is] + p.precond_width_min, p.precond_width_max);
            if (nband[is] > BAND_WIDTH_GPU) {
                use_cg[is] = 1;
                nband[is] = 0;
                nwide++;
                continue;
            }
            retry = true;
        }
        if (!retry) break;
        band_free(c.band);
        c.band_ready = false;
    }
    int nbmax = 0;
    for (int v : nband) nbmax = std::max(nbmax, v);
    c.stats[ST_NBAND] = nbmax;
    c.stats[ST_BAND_CHUNKS] = (double)c.band.nchunk;
    c.stats[ST_BAND_CHUNKS_UNIQUE] = (double)c.band.nchunk_unique;
    c.stats[ST_BAND_BLOCKS] = (double)c.band.nblock;
    c.stats[ST_BAND_BLOCKS_UNIQUE] = (double)c.band.nblock_unique;
    c.stats[ST_BAND_ROWS_FACTORED] = (double)c.band.nrows_factored;
    c.segcg.clear();
    c.ic0cg.clear();
    for (size_t is = 0; is < c.segp.size(); ++is)
        if (use_cg[is]) {
            c.segcg.push_back(c.segp[is]);
            c.ic0cg.push_back(ic[(size_t)c.segp[is].ipsd * c.nof]);
        }
    if (!c.segcg.empty()) {
        c.d_segcg.alloc(c.segcg.size());
        h2d(c.d_segcg.p, c.segcg.data(), c.segcg.size(), s);
        c.d_cg_ic0.alloc(c.ic0cg.size());
        h2d(c.d_cg_ic0.p, c.ic0cg.data(), c.ic0cg.size(), s);
        MADAM_CHECK(hipStreamSynchronize(s));
    }
    c.stats[ST_NCG] = (double)c.segcg.size();
    c.stats[ST_NWIDE] = (double)nwide;
    {
        int64_t rows = c.band.nrows;
        for (const Segment &g : c.segcg) rows += g.noba;
        c.prec_covers_all = rows == c.nbtot;
    }
}

// r = yba, z = M r, p = z (src/smadam_routines.f90:732-748)
static void pcg_init(Context &c)
{
    hipStream_t s = c.stream;
    int64_t nb = c.nbtot;
    for (auto *b : {&c.d_r, &c.d_ro, &c.d_z, &c.d_p, &c.d_pw, &c.d_ap}) b->alloc(nb);
    c.d_scal.alloc(8);
    int nblk = dots_nblocks(nb);
    c.d_part.alloc((size_t)nblk * 3);
    launch_fill(s, c.d_aa.p, 0.0, nb);
    launch_copy(s, c.d_r.p, c.d_yba.p, nb);
    apply_precond(c, c.d_r.p, c.d_z.p);
    launch_copy(s, c.d_p.p, c.d_z.p, nb);
    launch_pw_slots(s, c.d_p.p, c.d_bord.p, c.d_dwb_s.p, c.d_pw.p, nb);
    launch_dots(s, 1, c.d_r.p, c.d_z.p, c.d_r.p, nullptr, c.d_rmask.p, nb, c.d_part.p, nblk);
    launch_finalize(s, c.d_part.p, nblk, 3, c.d_scal.p + 2);
    double hs[8];
    d2h(hs, c.d_scal.p, 8, s);
    MADAM_CHECK(hipStreamSynchronize(s));
    c.stats[ST_RZINIT] = hs[2];
    c.stats[ST_RRINIT] = hs[4];
    h2d(c.d_scal.p, &hs[2], 1, s);   // scal[0] = rz
    MADAM_CHECK(hipStreamSynchronize(s));
    c.rr_hist.clear();
    c.pcg_steps_done = 0;
}

// PCG iterations (src/smadam_routines.f90:785-924).  With converge = false the
// exit tests are skipped and exactly nmax iterations run (benchmark mode).
static int pcg_steps(Context &c, int nmax, bool converge)
{
    hipStream_t s = c.stream;
    const Params &p = c.p;
    int64_t nb = c.nbtot;
    int nblk = dots_nblocks(nb);
    int grp = group_for(c);
    double rrinit = c.stats[ST_RRINIT];
    double hs[8];
    Timer t_sc, t_ga, t_pr, t_pc, t_it;
    int prof = c.profile;
    int done = 0;
    for (;;) {
        if (done >= nmax) break;
        done++;
        c.pcg_steps_done++;
        if (prof) { t_it.start(s); t_sc.start(s); }
        if ((c.kernel_variant & 15) == 0 || (c.kernel_variant & 15) == 3)   // segmented tiles (default)
            launch_scatter_seg(s, c.nmap, c.nactive, c.rec_layout == REC16 ? (const void *)c.d_ent.p : c.d_entc.p,
                               c.rec_layout, c.d_pw.p, c.sr_s.heads.p, c.sr_s.k0.p,
                               c.sr_s.fix_end.p, c.sr_s.fix_ord.p, c.d_nzq.p, c.nzpix, c.d_ccinv_c.p, c.d_m4x.p,
                               c.d_sr_part.p);
        else if ((c.kernel_variant & 15) == 2)
            launch_scatter_tiles(s, c.nmap, c.npix_cross, c.ndeg, c.d_off.p, c.nactive, c.d_ent.p, c.d_pw.p,
                                 c.d_ccinv6x.p, c.d_m4x.p, c.d_tileinfo.p, c.d_tilescratch.p);
        else
            launch_scatter_pixels(s, c.nmap, c.npix_cross, c.ndeg, c.d_off.p, c.d_ent.p, c.d_pw.p, c.d_ccinv6x.p,
                                  c.d_m4x.p);
        if (prof) { t_sc.stop(s); t_pr.start(s); }
        if (p.kfilter) apply_prior(c, c.d_p.p, c.d_ap.p);
        else if (p.diagfilter != 0) launch_axpby_nna(s, c.d_ap.p, c.d_nna.p, c.d_p.p, p.diagfilter, nb);
        else launch_fill(s, c.d_ap.p, 0.0, nb);
        if (prof) { t_pr.stop(s); t_ga.start(s); }
        if ((c.kernel_variant & 15) == 3 && c.sr_g.ok)
            launch_gather_seg(s, c.nmap, (int64_t)c.ndet * c.nsamp, c.d_pixw.p, c.d_m4x.p, c.ndeg, c.sr_g.heads.p,
                              c.sr_g.k0.p, c.sr_g.fix_end.p, c.sr_g.fix_ord.p, c.d_dwb.p, c.d_nna.p, c.d_p.p,
                              c.d_ap.p, c.d_sr_part.p);
        else if ((c.kernel_variant & 15) != 1)
            launch_gather_flat(s, c.nmap, c.ndet, c.nsamp, c.nbase, c.d_bstart.p,
                               c.rec_layout == REC16 ? (const void *)c.d_pixw.p : c.d_pixc.p, c.rec_layout,
                               c.d_m4x.p, c.ndeg, c.d_dwb.p, c.d_nna.p, c.d_p.p, c.d_ap.p);
        else
            launch_gather_baselines(s, c.nmap, c.ndet, c.nsamp, c.nbase, c.d_bstart.p, c.d_pixw.p, c.d_m4x.p, c.ndeg,
                                    c.d_dwb.p, c.d_nna.p, c.d_p.p, c.d_ap.p, grp);
        if (prof) t_ga.stop(s);
        launch_dots(s, 0, c.d_p.p, c.d_ap.p, nullptr, nullptr, c.d_rmask.p, nb, c.d_part.p, nblk);
        launch_finalize(s, c.d_part.p, nblk, 1, c.d_scal.p + 1);
        launch_pcg_update_r(s, c.d_scal.p, c.d_aa.p, c.d_r.p, c.d_ro.p, c.d_p.p, c.d_ap.p, nb);
        if (prof) t_pc.start(s);
        apply_precond(c, c.d_r.p, c.d_z.p);
        if (prof) t_pc.stop(s);
        launch_dots(s, 1, c.d_r.p, c.d_z.p, c.d_ro.p, nullptr, c.d_rmask.p, nb, c.d_part.p, nblk);
        launch_finalize_scalars(s, c.d_part.p, nblk, c.d_scal.p);
        bool last = false;
        if (converge) {
            d2h(hs, c.d_scal.p, 8, s);
            MADAM_CHECK(hipStreamSynchronize(s));
            double rz = hs[0], rr = hs[4];
            c.rr_hist.push_back(rr / rrinit);
            if (rr / rrinit > 1e3) { c.stats[ST_NITER] = c.pcg_steps_done; fail("CG is diverging"); }
            if (rr / rrinit < p.cglimit && c.pcg_steps_done > p.iter_min) last = true;
            if (rz == 0) last = true;
        }
        if (!last) {
            // (FP64: the slot-order gather of p measured faster than writing
            // pw scattered from the p update, which the FP32 path does)
            launch_pcg_update_p(s, c.d_scal.p, c.d_p.p, c.d_z.p, nb);
            launch_pw_slots(s, c.d_p.p, c.d_bord.p, c.d_dwb_s.p, c.d_pw.p, nb);
        }
        if (prof) {
            t_it.stop(s);
            c.stats[ST_T_SCATTER] += t_sc.ms() * 1e-3;
            c.stats[ST_T_PRIOR] += t_pr.ms() * 1e-3;
            c.stats[ST_T_GATHER] += t_ga.ms() * 1e-3;
            c.stats[ST_T_APPLY_PREC] += t_pc.ms() * 1e-3;
            c.stats[ST_T_STEP] += t_it.ms() * 1e-3;
            c.stats[ST_N_SCATTER] += 1;
            c.stats[ST_N_GATHER] += 1;
        }
        if (last) break;
    }
    MADAM_CHECK(hipStreamSynchronize(s));
    c.stats[ST_NITER] = c.pcg_steps_done;
    return done;
}

// y = A x for the destriping operator of the last solve: FᵀZF x (pixel side,
// ccmultiply, baseline side) plus the noise prior (or diagfilter) term, as one
// PCG step applies it (src/smadam_routines.f90:808-872).  pw is scratch of nbtot.
static void apply_operator(Context &c, const double *x, double *y, double *pw)
{
    hipStream_t s = c.stream;
    const Params &p = c.p;
    int64_t nb = c.nbtot;
    launch_pw_slots(s, x, c.d_bord.p, c.d_dwb_s.p, pw, nb);
    launch_scatter_seg(s, c.nmap, c.nactive, c.rec_layout == REC16 ? (const void *)c.d_ent.p : c.d_entc.p,
                       c.rec_layout, pw, c.sr_s.heads.p, c.sr_s.k0.p, c.sr_s.fix_end.p, c.sr_s.fix_ord.p,
                       c.d_nzq.p, c.nzpix, c.d_ccinv_c.p, c.d_m4x.p, c.d_sr_part.p);
    if (p.kfilter) apply_prior(c, x, y);
    else if (p.diagfilter != 0) launch_axpby_nna(s, y, c.d_nna.p, x, p.diagfilter, nb);
    else launch_fill(s, y, 0.0, nb);
    launch_gather_flat(s, c.nmap, c.ndet, c.nsamp, c.nbase, c.d_bstart.p,
                       c.rec_layout == REC16 ? (const void *)c.d_pixw.p : c.d_pixc.p, c.rec_layout, c.d_m4x.p,
                       c.ndeg, c.d_dwb.p, c.d_nna.p, x, y);
}

// one operator of the last solve on a host vector of nbtot baselines:
// which 0 = noise prior C_a^-1 x (cinvmul), 1 = preconditioner M^-1 x,
// 2 = the destriping operator A x
static void apply_host(Context &c, int which, const double *hx, double *hy)
{
    if (!c.solved || !c.p.kfirst || !c.tiles_built) fail("madam_apply needs a completed madam_solve with destriping");
    if (which == 0 && !c.p.kfilter) fail("madam_apply: no noise prior (kfilter = F)");
    if (which < 0 || which > 2) fail("madam_apply: unknown operator %d", which);
    hipStream_t s = c.stream;
    int64_t nb = c.nbtot;
    DBuf<double> x, y, pw;
    x.alloc(std::max<int64_t>(nb, 1));
    y.alloc(std::max<int64_t>(nb, 1));
    pw.alloc(std::max<int64_t>(nb, 1));
    h2d(x.p, hx, nb, s);
    if (which == 0) apply_prior(c, x.p, y.p);
    else if (which == 1) apply_precond(c, x.p, y.p);
    else apply_operator(c, x.p, y.p, pw.p);
    MADAM_CHECK(hipGetLastError());
    d2h(hy, y.p, nb, s);
    MADAM_CHECK(hipStreamSynchronize(s));
}

static void run_subsets(Context &c);

template <typename TT>
void solve_setup(Context &c, const TT *tod)
{
    if (!c.pointed) fail("madam_solve called before the pointing was set");
    hipStream_t s = c.stream;
    const Params &p = c.p;
    Timer tm;
    for (int i : {ST_T_SCATTER, ST_T_PRIOR, ST_T_GATHER, ST_T_APPLY_PREC, ST_T_STEP, ST_N_SCATTER, ST_N_GATHER})
        c.stats[i] = 0;
    // binning
    tm.start(s);
    c.d_cc6.alloc(c.npix_map * 6);
    c.d_rhs4.alloc(c.npix_map * 4);
    c.d_dwb_s.alloc(std::max<int64_t>(c.nbtot, 1));
    c.d_sdwb_s.alloc(std::max<int64_t>(c.nbtot, 1));
    launch_gather_order(s, c.d_dwb.p, c.d_bord.p, c.d_dwb_s.p, c.nbtot);
    launch_gather_order(s, c.d_sdwb.p, c.d_bord.p, c.d_sdwb_s.p, c.nbtot);
    launch_bin(s, c.nmap, c.npix_map, c.d_off.p, c.d_ent.p, c.d_src.p, tod, c.d_dwb_s.p, c.d_sdwb_s.p, c.d_cc6.p,
               c.d_rhs4.p);
    double *cc6x = c.d_cc6.p, *rhs4x = c.d_rhs4.p;
    if (c.ndeg > 1) {
        c.d_cc6x.alloc(c.npix_cross * 6);
        c.d_rhs4x.alloc(c.npix_cross * 4);
        launch_degrade(s, c.nmap, c.npix_cross, c.ndeg, c.d_cc6.p, c.d_rhs4.p, c.d_cc6x.p, c.d_rhs4x.p);
        cc6x = c.d_cc6x.p;
        rhs4x = c.d_rhs4x.p;
    }
    tm.stop(s);
    c.stats[ST_T_BIN] = tm.ms() * 1e-3;
    c.d_aa.alloc(c.nbtot);
    launch_fill(s, c.d_aa.p, 0.0, c.nbtot);
    c.stats[ST_NITER] = 0;
    c.rr_hist.clear();
    if (p.kfirst) {
        tm.start(s);
        c.d_ccinv6x.alloc(c.npix_cross * 6);
        launch_invert(s, c.nmap, c.npix_cross, cc6x, c.d_ccinv6x.p, nullptr, nullptr, p.pixmode_cross, p.pixlim_cross,
                      p.incomplete_matrices, p.allow_decoupling, nullptr);
        c.d_m4x.alloc(c.npix_cross * 4);
        launch_ccmultiply(s, c.nmap, c.npix_cross, c.d_ccinv6x.p, rhs4x, c.d_m4x.p);
        if (!c.tiles_built) {   // tables of the projections: they depend on the pointing only
            c.d_tilescratch.alloc(scatter_tiles_scratch_bytes(c.nactive));
            c.d_tileinfo.alloc(scatter_tiles_info_bytes(c.nactive));
            launch_tile_info(s, c.npix_cross, c.ndeg, c.d_off.p, c.nactive, c.d_tileinfo.p);
            build_segtiles(c);
            c.tiles_built = true;
        }
        c.d_ccinv_c.alloc((size_t)std::max<int64_t>(c.nzpix, 1) * 6);
        segred_compact_cc(s, c.d_nzq.p, c.nzpix, c.d_ccinv6x.p, c.d_ccinv_c.p);
        c.d_yba.alloc(c.nbtot);
        c.d_nna.alloc(c.nbtot);
        int grp = group_for(c);
        launch_init_a(s, c.nmap, c.ndet, c.nsamp, c.nbase, c.d_bstart.p, c.d_pixw.p, tod, c.d_m4x.p, c.ndeg,
                      c.d_dwb.p, c.d_yba.p, c.d_nna.p, grp);
        MADAM_CHECK(hipGetLastError());
        bool reuse = false;
        if (c.prec_cached && c.prec_cached_variant == c.kernel_variant) {
            c.d_flag.alloc(1);
            MADAM_CHECK(hipMemsetAsync(c.d_flag.p, 0, sizeof(int), s));
            launch_bits_differ(s, c.d_nna.p, c.d_nna_key.p, c.nbtot, c.d_flag.p);
            int differ = 1;
            d2h(&differ, c.d_flag.p, 1, s);
            MADAM_CHECK(hipStreamSynchronize(s));
            reuse = differ == 0;
        }
        if (!reuse) build_segments(c);
        c.stats[ST_PREC_REUSED] = reuse ? 1 : 0;
        tm.stop(s);
        c.stats[ST_T_INIT] = tm.ms() * 1e-3;
        double tp = now();
        if (!reuse && p.kfilter) {
            build_filters(c);
            // batches whose work buffer (16 B per complex point) stays within
            // 128 MiB, so the row and inverse column passes re-read it from the
            // 256 MB last-level cache (32 segments of 2^19 points: 2 x 16 is
            // 6 % faster than 1 x 32; smaller batches leave the GPU underfilled)
            // the inner-CG preconditioner convolves its segments with the same buffers
            int64_t nsc = std::max<int64_t>((int64_t)std::max(c.segconv.size(), c.segp.size()), 1);
            int64_t cap = std::max<int64_t>((128ll << 20) / (16 * c.fft.N), 1);
            cap = std::min<int64_t>(cap, 64);
            int64_t nbatch = (nsc + cap - 1) / cap;
            c.fft_batch = (nsc + nbatch - 1) / nbatch;
            // batch buffer, ipsd ints and the per-segment partial sums
            c.d_fftwork.alloc(fft_convolve_work_doubles(c.fft, c.fft_batch, nsc));
            c.fft_nseg_max = nsc;
        }
        if (!reuse) {
            build_preconditioner(c);
            c.d_nna_key.alloc(std::max<int64_t>(c.nbtot, 1));
            MADAM_CHECK(hipMemcpyAsync(c.d_nna_key.p, c.d_nna.p, sizeof(double) * c.nbtot, hipMemcpyDeviceToDevice, s));
            c.prec_cached = true;
            c.prec_cached_variant = c.kernel_variant;
        }
        MADAM_CHECK(hipStreamSynchronize(s));
        c.stats[ST_T_PRECOND] = now() - tp;
        c.stats[ST_PRECOND_KIND] = c.precond_kind;
        c.stats[ST_NOF] = (double)c.nof;
    }
}

template void solve_setup<double>(Context &, const double *);
template void solve_setup<float>(Context &, const float *);

static void solve(Context &c, double *signal, int on_device)
{
    if (!c.pointed) fail("madam_solve called before the pointing was set");
    hipStream_t s = c.stream;
    const Params &p = c.p;
    double tstart = now();
    int64_t n = (int64_t)c.ndet * c.nsamp;
    if (on_device) c.d_tod = signal;
    else {
        c.d_tod_own.alloc(n);
        h2d(c.d_tod_own.p, signal, n, s);
        c.d_tod = c.d_tod_own.p;
    }
    solve_setup(c, (const double *)c.d_tod);
    Timer tm;
    if (p.kfirst) {
        // PCG (src/smadam_routines.f90:677-1483)
        double ti = now();
        pcg_init(c);
        if (!(c.stats[ST_RZINIT] < F32(1.e-30)) && !std::isnan(c.stats[ST_RZINIT])) pcg_steps(c, p.iter_max, true);
        c.stats[ST_T_ITER] = now() - ti;
        c.stats[ST_RR_FINAL] = c.rr_hist.empty() ? 0 : c.rr_hist.back();
    }
    // final maps
    tm.start(s);
    c.d_rhsd4.alloc(c.npix_map * 4);
    MADAM_CHECK(hipMemcpyAsync(c.d_rhsd4.p, c.d_rhs4.p, sizeof(double) * c.npix_map * 4, hipMemcpyDeviceToDevice, s));
    if (p.kfirst) {
        c.d_aa_s.alloc(std::max<int64_t>(c.nbtot, 1));
        launch_gather_order(s, c.d_aa.p, c.d_bord.p, c.d_aa_s.p, c.nbtot);
        launch_subtract_baselines(s, c.nmap, c.npix_map, c.d_off.p, c.d_ent.p, c.d_aa_s.p, c.d_dwb_s.p, c.d_rhsd4.p);
    }
    c.d_ccinv6.alloc(c.npix_map * 6);
    c.d_mask.alloc(c.npix_map);
    c.d_crit.alloc(c.npix_map);
    MADAM_CHECK(hipMemsetAsync(c.d_crit.p, 0, sizeof(double) * c.npix_map, s));
    launch_invert(s, c.nmap, c.npix_map, c.d_cc6.p, c.d_ccinv6.p, c.d_crit.p, c.d_mask.p, p.pixmode_map, p.pixlim_map,
                  p.incomplete_matrices, p.allow_decoupling, nullptr);
    c.d_map4.alloc(c.npix_map * 4);
    c.d_binmap4.alloc(c.npix_map * 4);
    launch_ccmultiply(s, c.nmap, c.npix_map, c.d_ccinv6.p, c.d_rhs4.p, c.d_binmap4.p);
    if (p.kfirst) launch_ccmultiply(s, c.nmap, c.npix_map, c.d_ccinv6.p, c.d_rhsd4.p, c.d_map4.p);
    else MADAM_CHECK(hipMemsetAsync(c.d_map4.p, 0, sizeof(double) * c.npix_map * 4, s));
    MADAM_CHECK(hipStreamSynchronize(s));
    c.solved = true;
    // subset maps (src/smadam.F90:430, 954-1257) see the TOD before the final clean
    if (c.subsets_enabled) run_subsets(c);
    // clean_tod (src/smadam_routines.f90:1545-1585), final full-data pass only
    bool clean = c.pass == 0 && p.kfirst && (p.write_tod || p.ndetset > 0 || p.nsurvey > 0);
    if (clean) {
        launch_clean_tod(s, c.ndet, c.nsamp, c.nbase, c.d_bstart.p, c.d_aa.p, c.d_tod);
        if (!on_device) d2h(signal, c.d_tod, n, s);
    }
    tm.stop(s);
    MADAM_CHECK(hipGetLastError());
    c.stats[ST_T_FINAL] = tm.ms() * 1e-3;
    MADAM_CHECK(hipStreamSynchronize(s));
    c.stats[ST_T_TOTAL] = now() - tstart;
    c.solved = true;
}

// restore_pixels for host arrays: flagged samples -> 12 * nside_max**2
static void write_back_pixels(Context &c, int32_t *host_pix)
{
    if (!c.pix_host_dirty) return;
    int64_t n = (int64_t)c.ndet * c.nsamp;
    launch_restore_pixels(c.stream, c.d_pix, n, (int32_t)(12LL * c.nside_max * c.nside_max));
    d2h(host_pix, c.d_pix, n, c.stream);
    MADAM_CHECK(hipStreamSynchronize(c.stream));
    c.pix_host_dirty = false;
}

static void get_maps(Context &c, double *map, double *binmap, int32_t *hits, double *cc, int32_t *mask, double *crit)
{
    if (!c.solved) fail("no solution available");
    hipStream_t s = c.stream;
    int64_t np = c.npix_map;
    int nm = c.nmap;
    std::vector<double> m4(np * 4), b4(np * 4), ci(np * 6);
    std::vector<int32_t> mk(np);
    d2h(m4.data(), c.d_map4.p, np * 4, s);
    d2h(b4.data(), c.d_binmap4.p, np * 4, s);
    d2h(ci.data(), c.d_ccinv6.p, np * 6, s);
    d2h(mk.data(), c.d_mask.p, np, s);
    if (hits) d2h(hits, c.d_hits.p, np, s);
    if (crit) d2h(crit, c.d_crit.p, np, s);
    MADAM_CHECK(hipStreamSynchronize(s));
    if (c.p.rm_monopole) {
        // makemaps monopole removal (src/map_routines.f90:341-353)
        int64_t cover = 0;
        for (int64_t q = 0; q < np; ++q) cover += mk[q] > 0;
        if (cover > 0) {
            for (auto *v : {&m4, &b4}) {
                double mono = 0;
                for (int64_t q = 0; q < np; ++q) mono += (*v)[q * 4];
                mono = mono / cover;
                for (int64_t q = 0; q < np; ++q) if (mk[q] > 0) (*v)[q * 4] -= mono;
            }
        }
    }
    for (int64_t q = 0; q < np; ++q)
        for (int a = 0; a < nm; ++a) {
            if (map) map[q * nm + a] = m4[q * 4 + a];
            if (binmap) binmap[q * nm + a] = b4[q * 4 + a];
        }
    if (cc) {
        for (int64_t q = 0; q < np; ++q) {
            const double *u = ci.data() + q * 6;
            if (nm == 1) cc[q] = u[0];
            else {
                double full[9] = {u[0], u[1], u[2], u[1], u[3], u[4], u[2], u[4], u[5]};
                for (int a = 0; a < 9; ++a) cc[q * 9 + a] = full[a];
            }
        }
    }
    if (mask) std::copy(mk.begin(), mk.end(), mask);
}

// ------------------------------------------------------------------ FITS products
// src/output.f90:114-1175 (write_map, write_binmap, write_mask, write_hits,
// write_matrix), file naming src/parameter_control.f90:47-72, src/output.f90:86-106
// (addpath), 1329-1366 (check_outfile), src/smadam.F90:1285-1314 (add_subchunk_id).

static const float HEALPIX_UNDEF = -1.6375e30f;

struct OutNames {
    std::string map, binmap, hit, mask, matrix, wcov;
    std::string leak;   // file_leakmatrix: "<leak>_<detector>.fits" per detector
};

static bool file_exists(const std::string &f)
{
    FILE *x = fopen(f.c_str(), "rb");
    if (x) fclose(x);
    return x != nullptr;
}

static std::string check_outfile(const std::string &fn)
{
    if (fn.empty() || fn[0] == '!' || !file_exists(fn)) return fn;
    size_t n = fn.rfind('.');
    std::string base = n == std::string::npos ? fn : fn.substr(0, n);
    std::string ending = n == std::string::npos ? "" : fn.substr(n);
    for (int k = 1; k < 1000; ++k) {
        char s[16];
        snprintf(s, sizeof s, "_%03d", k);
        std::string f = base + s + ending;
        if (!file_exists(f)) return f;
    }
    return "";
}

static std::string add_subchunk_id(const std::string &fn, int isub, int nsub)
{
    if (fn.empty() || isub == 0) return fn;
    char s[32];
    if (nsub < 10) snprintf(s, sizeof s, "_sub%1dof%1d", isub, nsub);
    else snprintf(s, sizeof s, "_sub%02dof%02d", isub, nsub);
    size_t i = fn.rfind('.');
    if (i == std::string::npos) return fn + s;
    return fn.substr(0, i) + s + fn.substr(i);
}

static std::string add_path(std::string path, const std::string &fn)
{
    if (path.empty() || fn.empty()) return fn;
    if (path.back() != '/') path += '/';
    if (fn[0] == '!') return "!" + path + fn.substr(1);
    return path + fn;
}

static OutNames output_names(const Params &p, const std::string &path, int isub)
{
    std::string sub = p.file_root;
    if (!p.detset0.empty()) sub += "_" + p.detset0;
    if (!p.survey0.empty()) sub += "_" + p.survey0;
    OutNames n;
    auto mk = [&](bool on, const char *suffix) {
        return on ? add_path(path, add_subchunk_id(sub + suffix, isub, p.nsubchunk)) : std::string();
    };
    n.map = mk(p.write_map, "_map.fits");
    n.binmap = mk(p.write_binmap, "_bmap.fits");
    n.hit = mk(p.write_hits, "_hmap.fits");
    n.matrix = mk(p.write_matrix, "_wcov_inv.fits");
    n.wcov = mk(p.write_wcov, "_wcov.fits");
    n.mask = mk(p.write_mask, "_mask.fits");
    n.leak = mk(p.write_leakmatrix, "_leakmatrix");
    return n;
}

// host copy of one set of products (the solve's, or one subset's)
struct HostProducts {
    int nmap = 3;
    bool temponly = false, kfirst = true;
    std::vector<double> map, binmap, cc6, ci6, crit;   // maps [npix][nmap]
    std::vector<int32_t> hits, mask;
};

static void fetch_products(Context &c, HostProducts &h)
{
    int64_t np = c.npix_map;
    h.nmap = c.nmap;
    h.temponly = c.p.temperature_only != 0;
    h.kfirst = c.p.kfirst != 0;
    h.map.resize(np * c.nmap);
    h.binmap.resize(np * c.nmap);
    h.hits.resize(np);
    h.mask.resize(np);
    h.crit.resize(np);
    get_maps(c, h.map.data(), h.binmap.data(), h.hits.data(), nullptr, h.mask.data(), h.crit.data());
    h.cc6.resize(np * 6);
    h.ci6.resize(np * 6);
    d2h(h.cc6.data(), c.d_cc6.p, np * 6, c.stream);
    d2h(h.ci6.data(), c.d_ccinv6.p, np * 6, c.stream);
    MADAM_CHECK(hipStreamSynchronize(c.stream));
}

// header keyword groups (src/output.f90:1376-1590)
static void hdr_code(std::vector<FitsKey> &k)
{
    k.push_back(FitsKey::note("----------------------------------"));
    k.push_back(FitsKey::note("         CODE             "));
    k.push_back(FitsKey::note("----------------------------------"));
    k.push_back(FitsKey::s("CREATOR", "Madam", "Software creating FITS file"));
    k.push_back(FitsKey::s("VERSION", "3.7", "Software version"));
    k.push_back(FitsKey::note("Serial version"));
    k.push_back(FitsKey::s("PARFILE", "", "Parameter file"));
}

static void hdr_simulation(const Context &c, int isub, const HostProducts &h, std::vector<FitsKey> &k)
{
    const Params &p = c.p;
    double to_d = 1.0 / p.fsample / 24.0 / 3600.0;
    k.push_back(FitsKey::note("-------------------------------"));
    k.push_back(FitsKey::note("          Simulation           "));
    k.push_back(FitsKey::note("-------------------------------"));
    k.push_back(FitsKey::r("FSAMPLE", F32(p.fsample), "Sampling frequency (Hz)"));
    k.push_back(FitsKey::r("TMISSION", F32(c.nsamp * to_d), "Mission duration/days"));
    k.push_back(FitsKey::r("TSTART", 0.0, "Mission start/days"));
    k.push_back(FitsKey::i("IMISSION", c.nsamp, "Mission duration/samples"));
    k.push_back(FitsKey::i("ISTART", 0, "Mission start/samples"));
    k.push_back(FitsKey::i("NSUBCHNK", p.nsubchunk, "Number of subchunks"));
    k.push_back(FitsKey::i("ISUBCHNK", isub, "Index of subchunk"));
    k.push_back(FitsKey::note("---------------------------------"));
    k.push_back(FitsKey::note("         Detector info           "));
    k.push_back(FitsKey::note("---------------------------------"));
    k.push_back(FitsKey::i("NODET", c.ndet, "Number of detectors"));
    k.push_back(FitsKey::i("MDETW", p.mode_detweight, "Detector weighting mode"));
    for (int i = 0; i < c.ndet; ++i)
        k.push_back(FitsKey::s("DETNAM" + std::to_string(i + 1), c.det_names[i], "Detector name"));
    k.push_back(FitsKey::b("TEMPONLY", h.temponly, "Temperature map only"));
    k.push_back(FitsKey::note("----------------------------"));
    k.push_back(FitsKey::note("          Pointing          "));
    k.push_back(FitsKey::note("----------------------------"));
}

static void hdr_destriping(const Context &c, const HostProducts &h, std::vector<FitsKey> &k)
{
    const Params &p = c.p;
    k.push_back(FitsKey::note("---------------------------"));
    k.push_back(FitsKey::note("     Data distribution     "));
    k.push_back(FitsKey::note("---------------------------"));
    k.push_back(FitsKey::note("Standard mode"));
    k.push_back(FitsKey::i("NSIDESUB", p.nside_submap, "Submap resolution"));
    k.push_back(FitsKey::note("-------------------------------------------"));
    k.push_back(FitsKey::note("        Destriping scheme                  "));
    k.push_back(FitsKey::note("-------------------------------------------"));
    k.push_back(FitsKey::b("KFIRST", h.kfirst, "First destriping phase on/off"));
    if (h.kfirst) {
        k.push_back(FitsKey::i("NSHORT", p.nshort, "Baseline length "));
        k.push_back(FitsKey::b("KFILTER", p.kfilter != 0, "Noise filter on/off"));
        k.push_back(FitsKey::s("BASIS", "LEGENDRE", "Destriping function basis"));
        k.push_back(FitsKey::i("BORDER", p.basis_order, "Destriping function order"));
    }
    k.push_back(FitsKey::i("NSIDECR", p.nside_cross, "Crossing point resolution"));
    k.push_back(FitsKey::i("PIXMCR", p.pixmode_cross, "Pixel removal criterion"));
    k.push_back(FitsKey::r("PIXLCR", F32(p.pixlim_cross), "Pixel removal limit"));
    k.push_back(FitsKey::r("CGLIMIT", F32(p.cglimit), "Conjugate gradient convergence limit"));
    k.push_back(FitsKey::i("ITERMAX", p.iter_max, "Maximum number of iterations"));
    if (p.kfilter) {
        k.push_back(FitsKey::note("---------------------------------------"));
        k.push_back(FitsKey::note("           Noise filter                "));
        k.push_back(FitsKey::note("---------------------------------------"));
        k.push_back(FitsKey::r("FILTTIM", (double)c.noba_pp_max * p.nshort / p.fsample, "Noise filter length/seconds"));
        for (int i = 0; i < c.ndet; ++i) {
            k.push_back(FitsKey::note(c.det_names[i]));
            k.push_back(FitsKey::r("SIGMA" + std::to_string(i + 1), F32(c.det[i].sigmas[0]), "White noise std"));
            k.push_back(FitsKey::r("PLATEAU" + std::to_string(i + 1), F32(c.det[i].plateaus[0]),
                                   "White noise plateau"));
        }
    }
}

static void hdr_healpix(const Context &c, const HostProducts &h, std::vector<FitsKey> &k)
{
    const Params &p = c.p;
    k.push_back(FitsKey::note("----------------------------------"));
    k.push_back(FitsKey::note("   Healpix Map Specific Keywords  "));
    k.push_back(FitsKey::note("----------------------------------"));
    k.push_back(FitsKey::s("PIXTYPE", "HEALPIX", "HEALPIX Pixelisation"));
    k.push_back(FitsKey::s("ORDERING", "NESTED", "Pixel ordering scheme"));
    k.push_back(FitsKey::i("NSIDE", p.nside_map, "Resolution parameter for HEALPIX"));
    k.push_back(FitsKey::i("FIRSTPIX", 0, "First pixel # (0 based)"));
    k.push_back(FitsKey::i("LASTPIX", 12LL * p.nside_map * p.nside_map - 1, "Last pixel # (0 based)"));
    k.push_back(FitsKey::b("POLAR", h.nmap != 1, "Polarisation included (True/False)"));
    k.push_back(FitsKey::s("POLCCONV", "COSMO", "Coord. convention for polarisation (COSMO/IAU)"));
    if (p.write_cut) {
        k.push_back(FitsKey::note("Cut sky data"));
        k.push_back(FitsKey::s("OBJECT", "PARTIAL"));
        k.push_back(FitsKey::s("INDXSCHM", "EXPLICIT", " Indexing : IMPLICIT or EXPLICIT"));
        k.push_back(FitsKey::i("GRAIN", 1, " Grain of pixel indexing"));
    } else {
        k.push_back(FitsKey::note("Full sky data"));
        k.push_back(FitsKey::s("OBJECT", "FULLSKY"));
        k.push_back(FitsKey::s("INDXSCHM", "IMPLICIT", " Indexing : IMPLICIT or EXPLICIT"));
        k.push_back(FitsKey::i("GRAIN", 0, " Grain of pixel indexing"));
    }
    k.push_back(FitsKey::note("GRAIN=0 : no indexing of pixel data                         (IMPLICIT)"));
    k.push_back(FitsKey::note("GRAIN=1 : 1 pixel index -> 1 pixel data                     (EXPLICIT)"));
    k.push_back(FitsKey::note("GRAIN>1 : 1 pixel index -> data of GRAIN consecutive pixels (EXPLICIT)"));
}

static void hdr_sky(const Context &c, int64_t skycover, std::vector<FitsKey> &k)
{
    const Params &p = c.p;
    k.push_back(FitsKey::note("----------------------------------"));
    k.push_back(FitsKey::note("         Sky coverage             "));
    k.push_back(FitsKey::note("----------------------------------"));
    k.push_back(FitsKey::i("PIXCOV", skycover, "Recovered pixels"));
    k.push_back(FitsKey::r("SKYCOV", F32(skycover / (12.0 * p.nside_map * p.nside_map)), "Sky coverage"));
    k.push_back(FitsKey::i("PIXMMAP", p.pixmode_map, "Pixel removal criterion"));
    k.push_back(FitsKey::r("PIXLMAP", F32(p.pixlim_map), "Pixel removal limit"));
}

// One product file.  sel: pixels written in cut-sky form (write_cut), vals:
// per-column data over all pixels (float or double by type, int32 for 'J').
struct ProductCol {
    std::string name, unit;
    char type;
    std::vector<float> f;
    std::vector<double> d;
    std::vector<int32_t> j;
};

static std::string write_product(const Context &c, const std::string &fname, const std::vector<ProductCol> &cols,
                                 const std::vector<char> &sel, int64_t repeat_full, const std::vector<FitsKey> &keys)
{
    std::string fn = check_outfile(fname);
    if (fn.empty()) return fn;
    bool overwrite = fn[0] == '!';
    if (overwrite) fn = fn.substr(1);
    int64_t np = c.npix_map;
    std::vector<FitsColumn> fc;
    std::vector<int64_t> idx;
    std::vector<ProductCol> packed;
    int64_t nrows;
    if (c.p.write_cut) {
        for (int64_t q = 0; q < np; ++q)
            if (sel[q]) idx.push_back(q);
        packed = cols;
        for (auto &pc : packed) {
            int64_t n = 0;
            for (int64_t q : idx) {
                if (pc.type == 'E') pc.f[n] = pc.f[q];
                else if (pc.type == 'D') pc.d[n] = pc.d[q];
                else pc.j[n] = pc.j[q];
                ++n;
            }
        }
        FitsColumn ic;
        ic.name = "Pixel index";
        ic.type = 'K';
        ic.repcount = 1;
        ic.data = idx.data();
        fc.push_back(ic);
        nrows = (int64_t)idx.size();
    } else {
        packed = cols;
        nrows = np / repeat_full;
    }
    for (auto &pc : packed) {
        FitsColumn x;
        x.name = pc.name;
        x.unit = pc.unit;
        x.type = pc.type;
        x.repcount = c.p.write_cut ? 1 : (int)repeat_full;
        x.data = pc.type == 'E' ? (const void *)pc.f.data() : pc.type == 'D' ? (const void *)pc.d.data()
                                                                              : (const void *)pc.j.data();
        fc.push_back(x);
    }
    std::string err = fits_write_bintable(fn, fc, nrows, keys, c.p.write_cut ? "CUT SKY MAP" : "FULL SKY MAP");
    if (!err.empty()) fail("FITS output failed: %s", err.c_str());
    (void)overwrite;
    return fn;
}

static void write_product_set(Context &c, const OutNames &n, int isub, const HostProducts &h)
{
    const Params &p = c.p;
    int64_t np = c.npix_map;
    int nm = h.nmap;
    int64_t nosubpix = (int64_t)(p.nside_map / p.nside_submap) * (p.nside_map / p.nside_submap);
    std::vector<char> good(np), hit(np), ccnz(np);
    int64_t skycover = 0;
    for (int64_t q = 0; q < np; ++q) {
        good[q] = h.mask[q] > 0;
        hit[q] = h.hits[q] > 0;
        ccnz[q] = h.cc6[q * 6] != 0;
        skycover += good[q];
    }
    static const char *names3[3] = {"TEMPERATURE", "Q-POLARISATION", "U-POLARISATION"};
    auto map_cols = [&](const std::vector<double> &m) {
        std::vector<ProductCol> cols(nm);
        for (int a = 0; a < nm; ++a) {
            cols[a].name = names3[a];
            cols[a].unit = p.unit_tod;
            cols[a].type = 'E';
            cols[a].f.resize(np);
            for (int64_t q = 0; q < np; ++q) cols[a].f[q] = good[q] ? (float)m[q * nm + a] : HEALPIX_UNDEF;
        }
        return cols;
    };
    auto matrix_cols = [&](const std::vector<double> &m6) {
        // cc(i, j), i <= j, from the 6-element symmetric storage (11, 12, 13, 22, 23, 33)
        std::vector<ProductCol> cols;
        static const int sl[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
        for (int a = 0; a < nm; ++a)
            for (int b = a; b < nm; ++b) {
                ProductCol pc;
                pc.name = "cc(" + std::to_string(a + 1) + "," + std::to_string(b + 1) + ")";
                pc.type = 'D';
                pc.d.resize(np);
                for (int64_t q = 0; q < np; ++q) pc.d[q] = m6[q * 6 + sl[a][b]];
                cols.push_back(pc);
            }
        return cols;
    };
    auto base_keys = [&](bool destriping) {
        std::vector<FitsKey> k;
        hdr_code(k);
        hdr_simulation(c, isub, h, k);
        if (destriping) hdr_destriping(c, h, k);
        return k;
    };
    if (!n.matrix.empty()) {
        auto k = base_keys(false);
        k.push_back(FitsKey::s("objtype", "madam.matrix", "Object type"));
        hdr_healpix(c, h, k);
        std::string fw = write_product(c, n.matrix, matrix_cols(h.cc6), ccnz, nosubpix, k);
        printf(" Pixel matrix written in %s\n", fw.c_str());
    }
    if (!n.wcov.empty()) {
        auto k = base_keys(false);
        k.push_back(FitsKey::s("objtype", "madam.matrix", "Object type"));
        hdr_healpix(c, h, k);
        std::string fw = write_product(c, n.wcov, matrix_cols(h.ci6), good, nosubpix, k);
        printf(" Pixel matrix written in %s\n", fw.c_str());
    }
    if (!n.map.empty()) {
        auto k = base_keys(true);
        k.push_back(FitsKey::s("objtype", "madam.map", "Object type"));
        hdr_healpix(c, h, k);
        hdr_sky(c, skycover, k);
        k.push_back(FitsKey::b("MONOPOLE", p.rm_monopole != 0, "Monopole removed (T map)"));
        std::string fw = write_product(c, n.map, map_cols(h.map), good, nosubpix, k);
        printf(" Map written in %s\n", fw.c_str());
    }
    if (!n.binmap.empty()) {
        auto k = base_keys(false);
        k.push_back(FitsKey::s("objtype", "madam.binmap", "Object type"));
        hdr_healpix(c, h, k);
        hdr_sky(c, skycover, k);
        k.push_back(FitsKey::b("MONOPOLE", p.rm_monopole != 0, "Monopole removed (T map)"));
        std::string fw = write_product(c, n.binmap, map_cols(h.binmap), good, nosubpix, k);
        printf(" Binned map written in %s\n", fw.c_str());
    }
    if (!n.mask.empty()) {
        auto k = base_keys(true);
        k.push_back(FitsKey::s("objtype", "madam.mask", "Object type"));
        hdr_healpix(c, h, k);
        hdr_sky(c, skycover, k);
        std::vector<ProductCol> cols(2);
        cols[0].name = "Mask";
        cols[0].type = 'J';
        cols[0].j = h.mask;
        cols[1].name = "criterion";
        cols[1].type = 'E';
        cols[1].f.resize(np);
        for (int64_t q = 0; q < np; ++q) cols[1].f[q] = (float)h.crit[q];
        if (p.write_cut) cols.erase(cols.begin());   // the mask is redundant in cut-sky form
        std::string fw = write_product(c, n.mask, cols, good, nosubpix, k);
        printf(" Mask written in %s\n", fw.c_str());
    }
    if (!n.hit.empty()) {
        auto k = base_keys(false);
        k.push_back(FitsKey::s("objtype", "madam.hits", "Object type"));
        hdr_healpix(c, h, k);
        std::vector<ProductCol> cols(1);
        cols[0].name = "Total hits";
        cols[0].type = 'J';
        cols[0].j = h.hits;
        int64_t rep = std::min<int64_t>(std::max<int64_t>(12 * 1024, nosubpix), np);
        std::string fw = write_product(c, n.hit, cols, hit, rep, k);
        printf(" Hit count written in %s\n", fw.c_str());
    }
}

// leakmatrix + write_leakmatrix (src/smadam_routines.f90:38-105, src/output.f90:1181-1320):
// per detector, gamma = C^-1 C_det with C_det the pixel matrix of that detector's
// samples (restricted to sflag's time spans), C^-1 the products' inverse matrices
static void write_leakmatrices(Context &c, const std::string &base, const std::vector<char> *sflag,
                               const std::vector<char> &dallowed, int isub, const HostProducts &h)
{
    if (base.empty()) return;
    const Params &p = c.p;
    hipStream_t s = c.stream;
    int64_t n = (int64_t)c.ndet * c.nsamp, np = c.npix_map;
    int nm = h.nmap;
    DBuf<double> cc6, rhs4;
    DBuf<int32_t> hits;
    DBuf<uint8_t> smask;
    cc6.alloc(np * 6);
    rhs4.alloc(np * 4);
    hits.alloc(np);
    smask.alloc(n);
    std::vector<uint8_t> hm(n);
    std::vector<double> cd(np * 6);
    std::vector<char> good(np);
    for (int64_t q = 0; q < np; ++q) good[q] = h.mask[q] > 0;
    static const int sl[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    for (int d = 0; d < c.ndet; ++d) {
        if (!dallowed[d]) continue;
        for (int dd = 0; dd < c.ndet; ++dd)
            for (int64_t i = 0; i < c.nsamp; ++i)
                hm[(size_t)dd * c.nsamp + i] = dd == d && (!sflag || (*sflag)[i]);
        h2d(smask.p, hm.data(), n, s);
        launch_bin_subset(s, nm, np, c.d_off.p, c.d_ent.p, c.d_src.p, smask.p, c.d_tod, c.d_dwb_s.p, c.d_sdwb_s.p,
                          cc6.p, rhs4.p, hits.p);
        MADAM_CHECK(hipGetLastError());
        d2h(cd.data(), cc6.p, np * 6, s);
        MADAM_CHECK(hipStreamSynchronize(s));
        std::vector<ProductCol> cols(nm * nm);
        for (int a = 0; a < nm; ++a)
            for (int b = 0; b < nm; ++b) {
                ProductCol &pc = cols[a * nm + b];
                pc.name = "gamma(" + std::to_string(a + 1) + "," + std::to_string(b + 1) + ")";
                pc.type = 'D';
                pc.d.resize(np);
                for (int64_t q = 0; q < np; ++q) {
                    double v = 0;
                    for (int m = 0; m < nm; ++m) v += h.ci6[q * 6 + sl[a][m]] * cd[q * 6 + sl[m][b]];
                    pc.d[q] = v;
                }
            }
        std::vector<FitsKey> k;
        hdr_code(k);
        hdr_simulation(c, isub, h, k);
        k.push_back(FitsKey::s("objtype", "madam.leakmatrix", "Object type"));
        hdr_healpix(c, h, k);
        int64_t nosubpix = (int64_t)(p.nside_map / p.nside_submap) * (p.nside_map / p.nside_submap);
        std::string fw = write_product(c, base + "_" + c.det_names[d] + ".fits", cols, good, nosubpix, k);
        printf(" Leakage matrix written in %s\n", fw.c_str());
    }
}

static void write_products(Context &c, const OutNames &n, int isub)
{
    if (n.map.empty() && n.binmap.empty() && n.hit.empty() && n.mask.empty() && n.matrix.empty() && n.wcov.empty() &&
        n.leak.empty())
        return;
    HostProducts h;
    fetch_products(c, h);
    write_product_set(c, n, isub, h);
    write_leakmatrices(c, n.leak, nullptr, std::vector<char>(c.ndet, 1), isub, h);
}

// run_subsets (src/smadam.F90:954-1257): binned (and, with write_map, destriped)
// maps of every detector set x survey combination, with their hits, pixel
// matrices and masks on the first pass.
static void run_subsets(Context &c)
{
    const Params &p = c.p;
    if (!p.bin_subsets || (p.detsets.empty() && p.surveys.empty())) return;
    hipStream_t s = c.stream;
    int64_t n = (int64_t)c.ndet * c.nsamp, np = c.npix_map;
    bool do_map = p.write_map != 0, do_binmap = p.write_binmap != 0;
    int npass = (do_binmap && do_map) ? 2 : 1;
    if (do_map) do_binmap = true;
    DBuf<double> cc6, rhs4, ci6, crit, b4, tclean;
    DBuf<int32_t> hits, mask;
    DBuf<uint8_t> smask;
    cc6.alloc(np * 6);
    rhs4.alloc(np * 4);
    ci6.alloc(np * 6);
    crit.alloc(np);
    b4.alloc(np * 4);
    hits.alloc(np);
    mask.alloc(np);
    smask.alloc(n);
    std::vector<uint8_t> hm(n);
    std::string root = add_path(c.subsets_path, p.file_root);
    for (int pass = 1; pass <= npass; ++pass) {
        const double *tod = c.d_tod;
        if (pass == npass && do_map && p.kfirst) {
            // binned maps of the destriped TOD: tod - F a on a copy
            tclean.alloc(n);
            MADAM_CHECK(hipMemcpyAsync(tclean.p, c.d_tod, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            launch_clean_tod(s, c.ndet, c.nsamp, c.nbase, c.d_bstart.p, c.d_aa.p, tclean.p);
            tod = tclean.p;
        }
        for (int isv = 0; isv <= (int)p.surveys.size(); ++isv) {
            std::vector<char> sflag(c.nsamp, 1);
            std::string svname = p.survey0;
            if (isv > 0) {
                const auto &sv = p.surveys[isv - 1];
                svname = sv.name;
                int64_t nh = 0;
                for (int64_t i = 0; i < c.nsamp; ++i) {
                    char f = 0;
                    for (size_t j = 0; j < sv.starts.size(); ++j)
                        if (sv.starts[j] <= c.timestamps[i] && c.timestamps[i] <= sv.stops[j]) f = 1;
                    sflag[i] = f;
                    nh += f;
                }
                if (nh == 0) {
                    if (p.info > 0) printf(" %s : no overlap\n", svname.c_str());
                    continue;
                }
            }
            for (int ids = 0; ids <= (int)p.detsets.size(); ++ids) {
                if (ids == 0 && isv == 0) continue;   // the full data set is the main product
                std::vector<char> dflag(c.ndet, 1);
                std::string dsname = p.detset0;
                bool temponly = p.temperature_only != 0;
                if (ids > 0) {
                    const auto &ds = p.detsets[ids - 1];
                    dsname = ds.name;
                    int nd = 0;
                    for (int d = 0; d < c.ndet; ++d) {
                        char f = 0;
                        for (const auto &nm : ds.dets)
                            if (nm == c.det_names[d]) f = 1;
                        dflag[d] = f;
                        nd += f;
                    }
                    temponly = ds.nopol && !p.force_pol;
                    if (nd == 0) {
                        if (p.info > 0) printf(" %s : no overlap\n", dsname.c_str());
                        continue;
                    }
                }
                int nm = temponly ? 1 : c.nmap;
                std::string sub = root;
                if (!dsname.empty()) sub += "_" + dsname;
                if (!svname.empty()) sub += "_" + svname;
                auto nmf = [&](bool on, const char *suffix) {
                    std::string f = on ? add_subchunk_id(sub + suffix, c.pass, p.nsubchunk) : std::string();
                    if (!f.empty() && file_exists(f[0] == '!' ? f.substr(1) : f)) f.clear();   // existing: skipped
                    return f;
                };
                OutNames on;
                on.binmap = (pass == npass && do_map) ? nmf(true, "_map.fits") : nmf(do_binmap, "_bmap.fits");
                std::string leak;
                if (pass == 1) {
                    leak = p.write_leakmatrix ? add_subchunk_id(sub + "_leakmatrix.fits", c.pass, p.nsubchunk) : "";
                    on.hit = nmf(p.write_hits, "_hmap.fits");
                    on.matrix = nmf(p.write_matrix, "_wcov_inv.fits");
                    on.wcov = nmf(p.write_wcov, "_wcov.fits");
                    on.mask = nmf(p.write_mask, "_mask.fits");
                }
                if (on.binmap.empty() && on.hit.empty() && on.matrix.empty() && on.wcov.empty() && on.mask.empty() &&
                    leak.empty())
                    continue;
                for (int d = 0; d < c.ndet; ++d)
                    for (int64_t i = 0; i < c.nsamp; ++i) hm[(size_t)d * c.nsamp + i] = dflag[d] && sflag[i];
                h2d(smask.p, hm.data(), n, s);
                launch_bin_subset(s, nm, np, c.d_off.p, c.d_ent.p, c.d_src.p, smask.p, tod, c.d_dwb_s.p, c.d_sdwb_s.p,
                                  cc6.p, rhs4.p, hits.p);
                MADAM_CHECK(hipMemsetAsync(crit.p, 0, sizeof(double) * np, s));
                launch_invert(s, nm, np, cc6.p, ci6.p, crit.p, mask.p, p.pixmode_map, p.pixlim_map,
                              p.incomplete_matrices, p.allow_decoupling, nullptr);
                launch_ccmultiply(s, nm, np, ci6.p, rhs4.p, b4.p);
                MADAM_CHECK(hipGetLastError());
                HostProducts h;
                h.nmap = nm;
                h.temponly = temponly;
                h.kfirst = false;
                std::vector<double> m4(np * 4);
                h.cc6.resize(np * 6);
                h.ci6.resize(np * 6);
                h.crit.resize(np);
                h.hits.resize(np);
                h.mask.resize(np);
                d2h(m4.data(), b4.p, np * 4, s);
                d2h(h.cc6.data(), cc6.p, np * 6, s);
                d2h(h.ci6.data(), ci6.p, np * 6, s);
                d2h(h.crit.data(), crit.p, np, s);
                d2h(h.hits.data(), hits.p, np, s);
                d2h(h.mask.data(), mask.p, np, s);
                MADAM_CHECK(hipStreamSynchronize(s));
                if (p.rm_monopole) {   // makemaps (src/map_routines.f90:341-353)
                    int64_t cover = 0;
                    double mono = 0;
                    for (int64_t q = 0; q < np; ++q)
                        if (h.mask[q] > 0) { cover++; mono += m4[q * 4]; }
                    if (cover > 0)
                        for (int64_t q = 0; q < np; ++q)
                            if (h.mask[q] > 0) m4[q * 4] -= mono / cover;
                }
                h.binmap.resize(np * nm);
                for (int64_t q = 0; q < np; ++q)
                    for (int a = 0; a < nm; ++a) h.binmap[q * nm + a] = m4[q * 4 + a];
                write_product_set(c, on, c.pass, h);
                write_leakmatrices(c, leak, isv > 0 ? &sflag : nullptr, dflag, c.pass, h);
            }
        }
    }
}

// The caller's pix and weights are bitwise the ones the context's set-up was
// built from, pix as the reference leaves it in the caller's array (flagged
// samples as 12 * nside_max**2, which the device copy also holds once they
// have been written back).
static bool same_host_pointing(Context &c, const int32_t *pix, const float *w)
{
    if (!c.pointed || c.pix_host_dirty || !c.d_pix_own.p || c.d_pix != c.d_pix_own.p || c.d_w != c.d_w_own.p)
        return false;
    hipStream_t s = c.stream;
    int64_t n = (int64_t)c.ndet * c.nsamp;
    DBuf<int32_t> tp;
    DBuf<float> tw;
    DBuf<int> flag;
    tp.alloc(std::max<int64_t>(n, 1));
    tw.alloc(std::max<int64_t>(n * c.nnz, 1));
    flag.alloc(1);
    MADAM_CHECK(hipMemsetAsync(flag.p, 0, sizeof(int), s));
    h2d(tp.p, pix, n, s);
    h2d(tw.p, w, (size_t)n * c.nnz, s);
    launch_words_differ(s, (const uint32_t *)tp.p, (const uint32_t *)c.d_pix_own.p, n, flag.p);
    launch_words_differ(s, (const uint32_t *)tw.p, (const uint32_t *)c.d_w_own.p, n * c.nnz, flag.p);
    int differ = 1;
    d2h(&differ, flag.p, 1, s);
    MADAM_CHECK(hipStreamSynchronize(s));
    return differ == 0;
}

static Context *g_cache = nullptr;   // destripe(mcmode=T) keeps its context here

}  // namespace madam

using namespace madam;

// ============================================================================ C ABI

#define MADAM_API extern "C" __attribute__((visibility("default")))

template <typename F>
static int guarded(F &&f)
{
    try {
        f();
        return 0;
    } catch (const madam::Error &e) {
        madam::g_err = e.msg;
    } catch (const madam::HipError &e) {
        char buf[512];
        snprintf(buf, sizeof buf, "HIP error %d (%s) at %s:%d: %s", (int)e.err, hipGetErrorString(e.err), e.file, e.line,
                 e.what);
        madam::g_err = buf;
    } catch (const std::exception &e) {
        madam::g_err = e.what();
    }
    return 1;
}

static int g_last_status = 0;

// The reference aborts the MPI job on any error (src/mpi_wrappers.f90:120-134).
// The same happens here unless MADAM_RAISE is set in the environment, in which
// case the error is kept for madam_last_status() / madam_last_error().
static void abort_like_reference(const char *where)
{
    g_last_status = 1;
    fprintf(stderr, "libmadam: %s: %s\n", where, madam::g_err.c_str());
    fflush(stderr);
    if (getenv("MADAM_RAISE")) return;
    exit(1);
}

MADAM_API int madam_last_status(void) { return g_last_status; }

MADAM_API const char *madam_last_error(void) { return madam::g_err.c_str(); }

MADAM_API void *madam_create(int device)
{
    Context *c = nullptr;
    int rc = guarded([&] {
        c = new Context();
        c->device = device;
        MADAM_CHECK(hipSetDevice(device));
        ensure_stream(*c);
    });
    if (rc) { delete c; return nullptr; }
    return c;
}

MADAM_API void madam_destroy(void *ctx) { delete (Context *)ctx; }

MADAM_API int madam_set_profiling(void *ctx, int on)
{
    return guarded([&] { ((Context *)ctx)->profile = on; });
}

MADAM_API int madam_configure(void *ctx, const char *parstring, long ndet, const char *detstring,
                              const double *detweights, long nsamp, long nnz, const double *timestamps, long nperiod,
                              const long *periods, const long *npsd, long npsdtot, const double *psdstarts,
                              long npsdbin, const double *psdfreqs, long npsdval, const double *psdvals,
                              int subchunk_pass)
{
    return guarded([&] {
        Context &c = *(Context *)ctx;
        MADAM_CHECK(hipSetDevice(c.device));
        configure(c, parstring, ndet, detstring, detweights, nsamp, nnz, timestamps, nperiod, periods, npsd, npsdtot,
                  psdstarts, npsdbin, psdfreqs, npsdval, psdvals, subchunk_pass);
    });
}

MADAM_API int madam_set_pointing(void *ctx, int32_t *pix, float *pixweights, int on_device)
{
    return guarded([&] { set_pointing(*(Context *)ctx, pix, pixweights, on_device); });
}

MADAM_API int madam_solve(void *ctx, double *signal, int on_device)
{
    return guarded([&] { solve(*(Context *)ctx, signal, on_device); });
}

MADAM_API int madam_iterate(void *ctx, int niter, int restart)
{
    return guarded([&] {
        Context &c = *(Context *)ctx;
        if (!c.solved || !c.p.kfirst) fail("madam_iterate needs a completed madam_solve with destriping");
        if (restart) pcg_init(c);
        pcg_steps(c, niter, false);
    });
}

MADAM_API int madam_solve_batch(void *ctx, int nrhs, const void *signals, int tod_f32, int on_device, int precision)
{
    return guarded([&] {
        Context &c = *(Context *)ctx;
        MADAM_CHECK(hipSetDevice(c.device));
        solve_batch(c, nrhs, signals, tod_f32, on_device, precision);
    });
}

MADAM_API int madam_iterate_batch(void *ctx, int niter)
{
    return guarded([&] { batch_iterate(*(Context *)ctx, niter); });
}

MADAM_API int madam_get_batch_maps(void *ctx, int irhs, double *map, double *binmap)
{
    return guarded([&] { batch_get_maps(*(Context *)ctx, irhs, map, binmap); });
}

MADAM_API int madam_get_batch_baselines(void *ctx, int irhs, double *aa, double *yba)
{
    return guarded([&] { batch_get_baselines(*(Context *)ctx, irhs, aa, yba); });
}

MADAM_API int madam_get_batch_info(void *ctx, int irhs, double *out, int n)
{
    return guarded([&] { batch_get_info(*(Context *)ctx, irhs, out, n); });
}

MADAM_API int madam_get_filter(void *ctx, int ipsd, double *fcov)
{
    return guarded([&] {
        Context &c = *(Context *)ctx;
        if (!c.solved || !c.p.kfilter || !c.d_fcov.p) fail("madam_get_filter needs a completed madam_solve with kfilter = T");
        if (ipsd < 0 || ipsd >= c.npsdtot) fail("madam_get_filter: PSD index %d outside 0..%d", ipsd, c.npsdtot - 1);
        int64_t nh = c.nof / 2 + 1;
        d2h(fcov, c.d_fcov.p + (size_t)ipsd * nh, nh, c.stream);
        MADAM_CHECK(hipStreamSynchronize(c.stream));
    });
}

MADAM_API int madam_apply(void *ctx, int which, const double *x, double *y)
{
    return guarded([&] { apply_host(*(Context *)ctx, which, x, y); });
}

MADAM_API int madam_set_kernel_variant(void *ctx, int v)
{
    return guarded([&] {
        ((Context *)ctx)->kernel_variant = v;
        ((Context *)ctx)->prec_cached = false;
    });
}

MADAM_API int madam_reset_timers(void *ctx)
{
    return guarded([&] {
        Context &c = *(Context *)ctx;
        for (int i : {ST_T_SCATTER, ST_T_PRIOR, ST_T_GATHER, ST_T_APPLY_PREC, ST_T_STEP, ST_N_SCATTER, ST_N_GATHER})
            c.stats[i] = 0;
        batch_reset_timers(c);
    });
}

MADAM_API int madam_restore_pixels(void *ctx, int32_t *host_pix)
{
    return guarded([&] { write_back_pixels(*(Context *)ctx, host_pix); });
}

MADAM_API int madam_get_sizes(void *ctx, long *out6)
{
    return guarded([&] {
        Context &c = *(Context *)ctx;
        out6[0] = c.nmap;
        out6[1] = (long)c.npix_map;
        out6[2] = (long)c.nbase;
        out6[3] = c.ndet;
        out6[4] = (long)c.stats[ST_NITER];
        out6[5] = c.p.nside_map;
    });
}

MADAM_API int madam_get_maps(void *ctx, double *map, double *binmap, int32_t *hits, double *cc, int32_t *mask,
                             double *crit)
{
    return guarded([&] { get_maps(*(Context *)ctx, map, binmap, hits, cc, mask, crit); });
}

MADAM_API int madam_get_hits(void *ctx, int32_t *hits)
{
    return guarded([&] {
        Context &c = *(Context *)ctx;
        if (!c.pointed) fail("madam_get_hits needs the pointing");
        d2h(hits, c.d_hits.p, c.npix_map, c.stream);
        MADAM_CHECK(hipStreamSynchronize(c.stream));
    });
}

MADAM_API int madam_get_csr(void *ctx, int64_t *off, uint32_t *src, int32_t *bord)
{
    return guarded([&] {
        Context &c = *(Context *)ctx;
        if (!c.pointed) fail("madam_get_csr needs the pointing");
        if (off) d2h(off, c.d_off.p, c.npix_map + 1, c.stream);
        if (src) d2h(src, c.d_src.p, c.nactive, c.stream);
        if (bord) d2h(bord, c.d_bord.p, c.nbtot, c.stream);
        MADAM_CHECK(hipStreamSynchronize(c.stream));
    });
}

MADAM_API int madam_get_baselines(void *ctx, double *aa, double *yba, double *nna)
{
    return guarded([&] {
        Context &c = *(Context *)ctx;
        if (!c.solved) fail("no solution available");
        int64_t nb = c.nbtot;
        if (aa) d2h(aa, c.d_aa.p, nb, c.stream);
        if (c.p.kfirst) {
            if (yba) d2h(yba, c.d_yba.p, nb, c.stream);
            if (nna) d2h(nna, c.d_nna.p, nb, c.stream);
        } else {
            if (yba) std::fill(yba, yba + nb, 0.0);
            if (nna) std::fill(nna, nna + nb, 0.0);
        }
        MADAM_CHECK(hipStreamSynchronize(c.stream));
    });
}

MADAM_API int madam_get_stats(void *ctx, double *out, int n)
{
    return guarded([&] {
        Context &c = *(Context *)ctx;
        for (int i = 0; i < n && i < ST_COUNT; ++i) out[i] = c.stats[i];
    });
}

MADAM_API int madam_get_rr_history(void *ctx, double *out, int n)
{
    return guarded([&] {
        Context &c = *(Context *)ctx;
        for (int i = 0; i < n && i < (int)c.rr_hist.size(); ++i) out[i] = c.rr_hist[i];
    });
}

MADAM_API int madam_expand_pointing(const double *scan5, const double *dets3, int ndet, long nside, long i0, long n,
                                    int32_t *pix, float *weights, int on_device)
{
    return guarded([&] {
        ScanParams sp{scan5[0], scan5[1], scan5[2], scan5[3], scan5[4]};
        hipStream_t s = nullptr;
        MADAM_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        DBuf<ScanDet> dd;
        dd.alloc(ndet);
        h2d(dd.p, (const ScanDet *)dets3, ndet, s);
        int32_t *dp = pix;
        float *dw = weights;
        DBuf<int32_t> op;
        DBuf<float> ow;
        if (!on_device) {
            op.alloc((size_t)ndet * n);
            ow.alloc((size_t)ndet * n * 3);
            dp = op.p;
            dw = ow.p;
        }
        launch_expand_pointing(s, sp, dd.p, ndet, nside, i0, n, n, dp, dw);
        MADAM_CHECK(hipGetLastError());
        if (!on_device) {
            d2h(pix, dp, (size_t)ndet * n, s);
            d2h(weights, dw, (size_t)ndet * n * 3, s);
        }
        MADAM_CHECK(hipStreamSynchronize(s));
        (void)hipStreamDestroy(s);
    });
}

// ---------------------------------------------------------------- reference ABI

MADAM_API void destripe(int comm, char *parstring, long ndet, char *detstring, double *detweights, long nsamp,
                        long nnz, void *timestamps, void *pix, void *pixweights, void *signal, long nperiod,
                        long *periods, long *npsd, long npsdtot, double *psdstarts, long npsdbin, double *psdfreqs,
                        long npsdval, double *psdvals)
{
    (void)comm;   // single rank: the communicator is recorded, not used
    g_last_status = 0;
    if (g_cache) {
        g_err = "Destripe called while caches are not empty.";
        abort_like_reference("destripe");
        return;
    }
    Context *c = nullptr;
    int rc = guarded([&] {
        c = new Context();
        MADAM_CHECK(hipGetDevice(&c->device));
        configure(*c, parstring, ndet, detstring, detweights, nsamp, nnz, (const double *)timestamps, nperiod, periods,
                  npsd, npsdtot, psdstarts, npsdbin, psdfreqs, npsdval, psdvals, 0);
        // the subchunk loop (src/smadam.F90:193-201): passes isubchunk + 1 .. nsubchunk, then the full data (0)
        const Params p0 = c->p;
        for (int counter = p0.isubchunk; counter <= p0.nsubchunk; ++counter) {
            int isub = (counter + 1) % (p0.nsubchunk + 1);
            OutNames names = output_names(p0, p0.path_output, isub);
            if (p0.skip_existing && !names.map.empty() && file_exists(names.map)) continue;
            if (isub == 0) {
                c->subsets_enabled = true;
                c->subsets_path = p0.path_output;
                set_pointing(*c, (int32_t *)pix, (float *)pixweights, 0);
                solve(*c, (double *)signal, 0);
                write_products(*c, names, 0);
            } else {
                Context sub;
                sub.device = c->device;
                configure(sub, parstring, ndet, detstring, detweights, nsamp, nnz, (const double *)timestamps,
                          nperiod, periods, npsd, npsdtot, psdstarts, npsdbin, psdfreqs, npsdval, psdvals, isub);
                sub.subsets_enabled = true;
                sub.subsets_path = p0.path_output;
                set_pointing(sub, (int32_t *)pix, (float *)pixweights, 0);
                solve(sub, (double *)signal, 0);
                write_products(sub, names, isub);
            }
        }
        if (c->solved) write_back_pixels(*c, (int32_t *)pix);
    });
    if (rc) {
        delete c;
        abort_like_reference("destripe");
        return;
    }
    if (c->p.mcmode) g_cache = c;
    else delete c;
}

MADAM_API void destripe_with_cache(int comm, long ndet, long nsamp, long nnz, void *timestamps, void *pix,
                                   void *pixweights, void *signal, char *outpath)
{
    (void)comm;
    (void)timestamps;
    g_last_status = 0;
    int rc = guarded([&] {
        if (!g_cache) fail("destripe_with_cache called without cached data (run destripe with mcmode=T first)");
        Context &c = *g_cache;
        if (ndet != c.ndet || nsamp != c.nsamp || nnz != c.nnz) fail("destripe_with_cache: data shape differs from the cached run");
        // src/smadam.F90:643-656, 797-803: map and binned map into outpath when given
        std::string path = (outpath && outpath[0]) ? strip(outpath) : c.p.path_output;
        c.subsets_enabled = true;
        c.subsets_path = path;
        // Monte Carlo mode: the same pointing with a new signal reuses the
        // whole set-up (pixel grouping, projection tables and, through the
        // bitwise nna check in solve, the filters and preconditioner)
        bool same = same_host_pointing(c, (const int32_t *)pix, (const float *)pixweights);
        if (!same) set_pointing(c, (int32_t *)pix, (float *)pixweights, 0, true);
        solve(c, (double *)signal, 0);
        c.stats[ST_POINTING_REUSED] = same ? 1 : 0;
        write_back_pixels(c, (int32_t *)pix);
        OutNames names = output_names(c.p, path, 0);
        names.hit.clear();
        names.mask.clear();
        names.matrix.clear();
        names.wcov.clear();
        write_products(c, names, 0);
    });
    if (rc) abort_like_reference("destripe_with_cache");
}

MADAM_API void clear_caches(void)
{
    if (!g_cache) {
        printf("WARNING: clear_caches called while caches are already empty\n");
        return;
    }
    delete g_cache;
    g_cache = nullptr;
}

MADAM_API int madam_cache_handle(void **out)
{
    *out = g_cache;
    return 0;
}
