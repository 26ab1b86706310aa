// Map-file writer: the reference's value text (main.cpp:3481-3519), the
// extrema summary (main.cpp:3530-3547) and the completion line
// (main.cpp:3565).  Each map file is formatted and written by one host
// thread, as the reference does (main.cpp:3481).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
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

// "00".."99"
static const char kDigitPairs[201] =
    "0001020304050607080910111213141516171819202122232425262728293031323334353637383940414243444546474849"
    "5051525354555657585960616263646566676869707172737475767778798081828384858687888990919293949596979899";

// Writes n / 1000 with three decimals and a newline, for |n| < 16,384,000,
// without data-dependent branches: the integer part is formed as five
// zero-padded digits and the text is copied from its first significant one.
// Copies 11 bytes after the sign, so `out` needs 12 bytes of room.
static int format_milli(int n, char *out) {
    const int neg = n < 0;
    const unsigned k = neg ? 0u - (unsigned)n : (unsigned)n;
    const unsigned ip = k / 1000u, fp = k - ip * 1000u;   // ip < 16384
    const unsigned hi = ip / 100u, lo = ip - hi * 100u;   // hi < 164
    const unsigned top = hi / 100u, mid = hi - top * 100u;
    char t[16] = {};
    t[0] = (char)('0' + top);
    memcpy(t + 1, kDigitPairs + 2 * mid, 2);
    memcpy(t + 3, kDigitPairs + 2 * lo, 2);
    t[5] = '.';
    t[6] = (char)('0' + fp / 100u);
    memcpy(t + 7, kDigitPairs + 2 * (fp % 100u), 2);
    t[9] = '\n';
    const int nd = 1 + (ip >= 10u) + (ip >= 100u) + (ip >= 1000u) + (ip >= 10000u);
    out[0] = '-';
    memcpy(out + neg, t + 5 - nd, 16 - 5);
    return neg + nd + 5;
}

// One map value: "0" when |e| < PRECISION, else "%.3f" of
// (float)(floor(e*1000 + 0.5)/1000)  (main.cpp:3487-3490, HAVE_ROUND unset).
//
// With n = floor(e*1000 + 0.5) and |n| < 16,384,000 the text is n/1000 with
// three decimals: v = (float)(n/1000) is below 16384 in magnitude, so it is
// within half a float ulp (<= 2^-11) of n/1000 plus a double rounding of
// order 2^-40, and 1000*v (exact in double) is within 0.489 of n, which
// printf's "%.3f" therefore rounds back to n.  Only larger values take the
// float conversion.  n is never -0 (x + 0.5 with x = -0.5 is +0).  The
// floor of x there is its truncation, less one when that lies above x.
int format_map_value(float e32, char *out) {
    const double e = (double)e32;
    if (std::fabs(e) < kPrecision) {
        out[0] = '0';
        out[1] = '\n';
        return 2;
    }
    const double x = e * 1000.0 + 0.5;
    if (x >= -16383999.0 && x < 16384000.0) {
        int n = (int)x;
        n -= (double)n > x;
        return format_milli(n, out);
    }
    const float v = (float)(std::floor(x) / 1000.0);
    return format_fixed3(v, out);
}

// Floating grid values: "%.3f" of (float)round3dp(r_min), no zero rule.
static int format_rmin(float r, char *out) {
    const double e = (double)r;
    const float v = (float)(std::floor(e * 1000.0 + 0.5) / 1000.0);
    return format_fixed3(v, out);
}

MapExtrema write_all_maps(Job &job, const float *maps, const float *fmap, const int32_t *cmap, int n_threads,
                          const std::function<void(int)> &wait_ready) {
    const int64_t P = job.points();
    const int nm = job.num_maps;
    MapExtrema ext;
    ext.emin.assign(nm, kBig);
    ext.emax.assign(nm, -kBig);
    std::vector<double> &emin = ext.emin, &emax = ext.emax;
    // one task per output file, as the reference writes one file per thread
    // (main.cpp:3481); a task waits for its map to reach the host
    std::vector<int> tasks;
    for (int j = 0; j < nm; j++) tasks.push_back(j);
    if (job.floating_grid && fmap && job.fmap_fp) tasks.push_back(nm);
    if (job.constriction_grid && cmap && job.cmap_fp) tasks.push_back(nm + 1);
    std::atomic<int> next{0};
    std::atomic<bool> problem{false};
    // values are formatted in blocks into a per-thread buffer with room for
    // the longest text of a value (64 bytes), and each block is written out
    constexpr int64_t kBlock = 65536;
    constexpr int64_t kRoom = 64;
    auto worker = [&]() {
        std::unique_ptr<char[]> buf(new char[(size_t)(kBlock * kRoom + kRoom)]);
        for (;;) {
            const int t = next.fetch_add(1);
            if (t >= (int)tasks.size()) return;
            const int j = tasks[t];
            if (wait_ready) wait_ready(j);
            FILE *f = j < nm ? job.maps[j].fp : (j == nm ? job.fmap_fp : job.cmap_fp);
            double lo = kBig, hi = -kBig;
            for (int64_t i0 = 0; i0 < P; i0 += kBlock) {
                const int64_t i1 = std::min(P, i0 + kBlock);
                char *o = buf.get();
                if (j < nm) {
                    const float *v = maps + (size_t)j * P;
                    for (int64_t i = i0; i < i1; i++) {
                        o += format_map_value(v[i], o);
                        lo = std::min(lo, (double)v[i]);
                        hi = std::max(hi, (double)v[i]);
                    }
                } else if (j == nm) {
                    for (int64_t i = i0; i < i1; i++) o += format_rmin(fmap[i], o);
                } else {
                    for (int64_t i = i0; i < i1; i++) o += snprintf(o, 16, "%d\n", (int)cmap[i]);
                }
                const size_t n = (size_t)(o - buf.get());
                if (f && fwrite(buf.get(), 1, n, f) != n) problem = true;
            }
            if (j < nm) {
                emin[j] = lo;
                emax[j] = hi;
            }
            if (f && fflush(f) != 0) problem = true;
            // binary side-output: the FP32 values themselves, counts as float
            if (job.binary_maps && job.make_files) {
                const std::string fn = (j < nm ? job.maps[j].filename : (j == nm ? job.fmap_filename : job.cmap_filename))
                                       + ".bin";
                std::vector<float> cnt;
                const float *src = j < nm ? maps + (size_t)j * P : (j == nm ? fmap : nullptr);
                if (j > nm) {
                    cnt.resize((size_t)P);
                    for (int64_t i = 0; i < P; i++) cnt[(size_t)i] = (float)cmap[i];
                    src = cnt.data();
                }
                FILE *b = fopen(fn.c_str(), "wb");
                if (!b || fwrite(src, sizeof(float), (size_t)P, b) != (size_t)P) problem = true;
                if (b && fclose(b) != 0) problem = true;
            }
        }
    };
    const int nt = std::max(1, std::min(n_threads, (int)tasks.size()));
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(worker);
    worker();
    for (auto &x : th) x.join();
    if (problem) fail(AG_ERR_IO, "Problems writing grid maps - there may not be enough disk space.\n");
    return ext;
}

// Extrema table of main.cpp:3530-3547.
void log_map_summary(Job &job, const MapExtrema &ext) {
    const std::vector<double> &emin = ext.emin, &emax = ext.emax;
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

// The per-plane progress lines ADT parses from the log (main.cpp:2772-2776,
// 3436-3462).  The planes are computed together on the device, so every
// line reports the whole computation's time and nothing remaining.
void log_plane_progress(Job &job, double seconds) {
    if (job.outlev < 1) return;
    job.logf("                    Percent   Estimated Time  Time/this plane\n");
    job.logf("XY-plane  Z-coord   Done      Remaining       Real, User, System\n");
    job.logf("            /Ang              /sec            /sec\n");
    job.logf("________  ________  ________  ______________  __________________________\n\n");
    const double zmin = job.center[2] - job.spacing * job.ne[2];
    for (int done = 1; done <= job.n1[2]; done++) {
        job.logf(" %6d   %8.3lf   %5.1lf%%   ", done - job.ne[2] - 1, zmin + (double)(done - 1) * job.spacing,
                 100. * done / (double)job.n1[2]);
        job.logf("        %5.2fs", 0.0f);
        job.logf("  Real= %.2f,  CPU= %.2f,  System= %.2f\n", seconds / job.n1[2], 0.0, 0.0);
    }
}

}  // namespace ag
