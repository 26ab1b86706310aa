// Map-file writer: the reference's value text (main.cpp:3481-3519), the
// extrema summary (main.cpp:3530-3547) and the completion line
// (main.cpp:3565).  Values are formatted on host threads in slices and each
// map file is written in one pass.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

#include "ag_internal.h"

namespace ag {

// Writes "%.3f\n" of the float v, byte-identical to glibc printf: v*1000 is
// exact in double (24 + 10 significant bits), and rint() rounds its ties to
// even as printf does for exactly representable halves.
static int format_fixed3(float v, char *out) {
    const double x = (double)v * 1000.0;
    if (!std::isfinite(x) || std::fabs(x) >= 9.0e15 || (v == 0.f && std::signbit(v)))
        return snprintf(out, 64, "%.3f\n", (double)v);
    long long k = (long long)std::rint(x);
    char *p = out;
    if (k < 0) {
        *p++ = '-';
        k = -k;
    }
    const long long ip = k / 1000;
    const int fp = (int)(k % 1000);
    char tmp[24];
    int n = 0;
    long long t = ip;
    do {
        tmp[n++] = (char)('0' + t % 10);
        t /= 10;
    } while (t);
    while (n) *p++ = tmp[--n];
    *p++ = '.';
    *p++ = (char)('0' + fp / 100);
    *p++ = (char)('0' + (fp / 10) % 10);
    *p++ = (char)('0' + fp % 10);
    *p++ = '\n';
    return (int)(p - out);
}

// One map value: "0" when |e| < PRECISION, else "%.3f" of
// (float)(floor(e*1000 + 0.5)/1000)  (main.cpp:3487-3490, HAVE_ROUND unset).
int format_map_value(float e32, char *out) {
    const double e = (double)e32;
    if (std::fabs(e) < kPrecision) {
        out[0] = '0';
        out[1] = '\n';
        return 2;
    }
    const float v = (float)(std::floor(e * 1000.0 + 0.5) / 1000.0);
    return format_fixed3(v, out);
}

// Floating grid values: "%.3f" of (float)round3dp(r_min), no zero rule.
static int format_rmin(float r, char *out) {
    const double e = (double)r;
    const float v = (float)(std::floor(e * 1000.0 + 0.5) / 1000.0);
    return format_fixed3(v, out);
}

template <typename F>
static std::string format_parallel(int64_t n, int n_threads, F &&one) {
    const int nt = std::max(1, std::min<int>(n_threads, (int)std::max<int64_t>(1, n / 65536)));
    std::vector<std::string> parts(nt);
    std::vector<std::thread> th;
    auto work = [&](int t) {
        const int64_t b = n * t / nt, e = n * (t + 1) / nt;
        std::string &s = parts[t];
        s.resize((size_t)(e - b) * 16 + 64);
        char *p = &s[0];
        for (int64_t i = b; i < e; i++) p += one(i, p);
        s.resize((size_t)(p - &s[0]));
    };
    for (int t = 1; t < nt; t++) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    size_t tot = 0;
    for (auto &s : parts) tot += s.size();
    std::string out;
    out.reserve(tot);
    for (auto &s : parts) out += s;
    return out;
}

void write_all_maps(Job &job, const float *maps, const float *fmap, const int32_t *cmap, int n_threads) {
    const int64_t P = job.points();
    bool problem = false;
    std::vector<double> emin(job.num_maps, kBig), emax(job.num_maps, -kBig);
    for (int j = 0; j < job.num_maps; j++) {
        const float *v = maps + (size_t)j * P;
        std::string text =
            format_parallel(P, n_threads, [v](int64_t i, char *o) { return format_map_value(v[i], o); });
        if (job.maps[j].fp && fwrite(text.data(), 1, text.size(), job.maps[j].fp) != text.size())
            problem = true;
        double lo = kBig, hi = -kBig;
        for (int64_t i = 0; i < P; i++) {
            lo = std::min(lo, (double)v[i]);
            hi = std::max(hi, (double)v[i]);
        }
        emin[j] = lo;
        emax[j] = hi;
    }
    if (job.floating_grid && fmap && job.fmap_fp) {
        std::string text = format_parallel(P, n_threads, [fmap](int64_t i, char *o) { return format_rmin(fmap[i], o); });
        if (fwrite(text.data(), 1, text.size(), job.fmap_fp) != text.size()) problem = true;
    }
    if (job.constriction_grid && cmap && job.cmap_fp) {
        std::string text = format_parallel(P, n_threads, [cmap](int64_t i, char *o) {
            return snprintf(o, 16, "%d\n", (int)cmap[i]);
        });
        if (fwrite(text.data(), 1, text.size(), job.cmap_fp) != text.size()) problem = true;
    }
    for (int j = 0; j < job.num_maps; j++)
        if (job.maps[j].fp && fflush(job.maps[j].fp) != 0) problem = true;
    if (problem) fail(AG_ERR_IO, "Problems writing grid maps - there may not be enough disk space.\n");

    job.logf("\nGrid\tAtom\tMinimum   \tMaximum\n");
    job.logf("Map \tType\tEnergy    \tEnergy \n");
    job.logf("\t\t(kcal/mol)\t(kcal/mol)\n");
    job.logf("____\t____\t_____________\t_____________\n");
    for (int i = 0; i < job.num_atom_maps; i++)
        job.logf(" %d\t %s\t  %6.2lf\t%9.2le\n", i + 1, job.maps[i].type.c_str(), emin[i], emax[i]);
    if (job.elecPE >= 0)
        job.logf(" %d\t %c\t  %6.2lf\t%9.2le\tElectrostatic Potential\n", job.elecPE + 1, 'e',
                 emin[job.elecPE], emax[job.elecPE]);
    if (job.dsolvPE >= 0)
        job.logf(" %d\t %c\t  %6.2lf\t%9.2le\tDesolvation Potential\n", job.dsolvPE + 1, 'd',
                 emin[job.dsolvPE], emax[job.dsolvPE]);
    job.logf("\n\n * Note:  Every pairwise-atomic interaction was clamped at %.2f\n\n", kEintClamp);
}

}  // namespace ag
