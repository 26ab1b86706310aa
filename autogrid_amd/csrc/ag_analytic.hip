// Analytic (differentiable) map mode on gfx950.
//
// The map set is evaluated from the functional forms behind the reference's
// tables instead of the tables themselves (main.cpp:2040-2276 tables,
// main.cpp:2805-3431 per-point loop):
//   E_mt(r)  = min(EINTCLAMP, Phi_mt(s(r))), Phi the 12-6 / 12-10 pair
//              potential, s(r) the continuous form of the min-window
//              smoothing (r + w left of the well, r* inside, r - w right);
//   sol(r)   = coeff_desolv exp(-r^2 / (2 * 3.6^2));
//   e        = sum over all atoms of q coeff_estat (332 / eps(r)) / max(r, 0.5),
//              eps the Mehler-Solmajer dielectric at r;
// every 8 A term multiplied by the switch S(r) = 1 - u^3 (10 - 15 u + 6 u^2),
// u = clamp((r - 7.5) / 0.5, 0, 1), in place of the hard cutoff, so the maps
// are continuous in the receptor coordinates.  The per-point assembly (closest
// donor H, racc / rdon / Hramp, the H-bond case split with its min / max,
// desolvation, d map) is the reference's.
//
// Kernels (FP32):
//   k_an_near   one lane per grid point over the atoms within 8 A (host-built
//               cell list): affinity, covalent and d maps, plus the tape the
//               adjoint needs (closest H, the atoms that gave each point's
//               H-bond min and max);
//   k_an_e      runs of 13 or 16 points along x per lane over one chunk of
//               atoms, and k_an_esum the chunks' sums per point: the e map;
//   k_an_vh     one lane per grid point: the cotangent of the closest H's
//               bond vector through every Hramp at that point;
//   k_an_hsum   one workgroup per donor H: sums k_an_vh's vectors over the
//               points whose closest H it is (of the points where they are
//               nonzero, packed);
//   k_an_atom   one workgroup per atom: gathers over the grid points within
//               8 A of it the gradients of the 8 A terms with respect to its
//               position, bond vectors, charge, pair parameters and the
//               desolvation weight;
//   k_an_count / k_an_gscan / k_an_pack
//               points where a cotangent is nonzero, packed in point order
//               (the e map's for k_an_eadj, the bond-vector one for
//               k_an_hsum); k_an_gmask flags the points k_an_atom reads;
//   k_an_eadj   atoms x chunks of packed points: the e map's gradients,
//               written as per-chunk partial sums that k_an_ered adds in a
//               fixed order.
// Nothing accumulates through atomics: every gradient is a fixed-order sum,
// so results are bitwise repeatable.

#include <hip/hip_runtime.h>
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "ag_devcommon.h"
#include "ag_internal.h"

namespace {

constexpr int kAT = 256;                 // threads per workgroup
constexpr int kAMaxAff = 8;              // affinity maps per analytic set (registers, unrolled)
constexpr float kSwOn = 7.5f, kSwInvLen = 2.0f;  // switch from 7.5 A, over 0.5 A
constexpr float kCut2 = 64.0f;
constexpr float kInv2Sig2 = 1.0f / (2.0f * 3.6f * 3.6f);
constexpr float kClamp = 100000.f;       // EINTCLAMP
constexpr float kCellA = 4.0f;           // cell edge (A)
constexpr int kEAtoms = 128;             // atoms per chunk of the forward e kernel (at least)
constexpr int kEMaxChunks = 64;          // chunks of the forward e kernel (at most)
// Mehler-Solmajer constants (distdepdiel.cpp:30-57)
constexpr float kMsA = -8.5525f, kMsB = 78.4f + 8.5525f, kMsK = 7.7839f, kMsLB = 0.003627f * (78.4f + 8.5525f);

enum : int { CASE_PLAIN = 0, CASE_ACC = 1, CASE_MM_RACC = 2, CASE_MM_RDON = 3 };

struct AAtom {            // 64 bytes, in cell order
    float x, y, z, q;     // grid frame; charge
    float vol, solpar;
    int meta;             // type | hbond << 8 | rexp << 16
    int index;
    float rv[3];
    float pad0;
    float rv2[3];
    float pad1;
};
static_assert(sizeof(AAtom) == 64, "AAtom layout");

struct APair {            // one (map, receptor type) pair potential
    float R, eps;         // derived form (eps effective)
    float cA, cB;         // explicit form (nbp_coeffs)
    float rstar;          // minimum of Phi
    int a, b, expl;
};

struct AParams {
    int n1[3], ne[3];
    float h;                          // spacing
    long long P;
    // cells over the box padded by 8 A
    float corg[3];
    int nc[3];
    const int *cell_start;            // [ncells + 1]
    const AAtom *atoms;               // cell order
    const float *rv_by_index;         // [N][3] bond vectors by atom index
    int n_atoms;
    // maps
    int naff;
    int aff_slot[kAMaxAff], des_slot[kAMaxAff];
    int cases[kAMaxAff][7];           // case per atom H-bond class
    float sp[kAMaxAff], volp[kAMaxAff];  // desolvation parameters of des_slot
    int cov_slot[16], n_cov;
    double cov_pos[3], cov_hw, cov_barrier;
    int d_slot;
    const APair *pairs;               // [naff][T]
    int T;
    float w_smooth;                   // half window (A)
    float w_desolv, qasp;
    float *maps;                      // [n_maps][P]
    // tape
    int *closest_h;                   // [P]
    int *amin, *amax;                 // [naff][P]
    // adjoint
    const float *dmaps;               // cotangent [n_maps][P]
    float *vh;                        // [P][3]
};

__device__ __forceinline__ float switch_s(float r, float &dS) {
    const float u = fminf(fmaxf((r - kSwOn) * kSwInvLen, 0.f), 1.f);
    const float u2 = u * u;
    dS = -30.f * u2 * (1.f - u) * (1.f - u) * kSwInvLen;
    return 1.f - u2 * u * (10.f - 15.f * u + 6.f * u2);
}

__device__ __forceinline__ float powi(float x, int n) {
    float x2 = x * x, x4 = x2 * x2, x8 = x4 * x4;
    switch (n) {
        case 12: return x8 * x4;
        case 10: return x8 * x2;
        case 6: return x4 * x2;
        case 4: return x4;
        default: {
            float r = 1.f;
            for (int i = 0; i < n; i++) r *= x;
            return r;
        }
    }
}

// Smoothed, capped pair energy and its derivatives (FP32).
struct LJ {
    float E, dEdr, dEdR, dEdeps;
};
__device__ __forceinline__ LJ lj_eval(const APair &pp, float r, float w) {
    float s, dsdr;
    if (r < pp.rstar - w) { s = r + w; dsdr = 1.f; }
    else if (r > pp.rstar + w) { s = r - w; dsdr = 1.f; }
    else { s = pp.rstar; dsdr = 0.f; }
    const float a = (float)pp.a, b = (float)pp.b;
    LJ o;
    float dPds, dPdR = 0.f, dPdeps = 0.f;
    // hardware reciprocals (about 1 ulp; rho^12 carries 12 of them, ~1e-6
    // relative, far inside the tolerances); a - b is a small integer
    const float is = __builtin_amdgcn_rcpf(s);
    if (!pp.expl) {
        const float rho = pp.R * is;
        const float pa = powi(rho, pp.a), pb = powi(rho, pp.b);
        const float k = __builtin_amdgcn_rcpf(a - b);
        o.E = pp.eps * k * (b * pa - a * pb);
        dPds = pp.eps * k * a * b * (pb - pa) * is;
        if (dsdr != 0.f) dPdR = pp.eps * k * a * b * (pa - pb) * __builtin_amdgcn_rcpf(pp.R);
        dPdeps = k * (b * pa - a * pb);
    } else {
        const float pa = powi(is, pp.a), pb = powi(is, pp.b);
        o.E = pp.cA * pa - pp.cB * pb;
        dPds = (-a * pp.cA * pa + b * pp.cB * pb) * is;
    }
    if (o.E >= kClamp) {
        o.E = kClamp;
        o.dEdr = o.dEdR = o.dEdeps = 0.f;
        return o;
    }
    o.dEdr = dPds * dsdr;
    o.dEdR = dPdR;
    o.dEdeps = dPdeps;
    return o;
}

__device__ __forceinline__ float racc_of(float c, int rexp, float &dc) {
    if (!(c > 0.f)) { dc = 0.f; return 0.f; }
    if (rexp == 2) { dc = 2.f * c; return c * c; }
    if (rexp == 4) { const float c2 = c * c; dc = 4.f * c2 * c; return c2 * c2; }
    dc = 1.f;
    return c;
}

// Hramp of a D1 atom against the closest H: 0.5 - 0.5 cos(4/3 acos(ct)).
__device__ __forceinline__ float hramp_of(float ct_raw, float &dct) {
    if (ct_raw >= 1.f || ct_raw <= -1.f) {
        dct = 0.f;
        const float ct = fminf(fmaxf(ct_raw, -1.f), 1.f);
        return 0.5f - 0.5f * cosf(acosf(ct) * (4.f / 3.f));
    }
    const float th = acosf(ct_raw);
    const float st = sinf(th);
    dct = st > 1.0e-4f ? -(2.f / 3.f) * sinf(th * (4.f / 3.f)) / st : -8.f / 9.f;
    return 0.5f - 0.5f * cosf(th * (4.f / 3.f));
}

// Ordered A2 lone-pair rdon (main.cpp:3185-3246) with the identities
// cos(pi/2 - acos x) = sqrt(1 - x^2) and sin(2 |acos y - pi/2|) = 2 |y| sqrt(1 - y^2).
// With GRAD, adds gbar * d rdon / d(dn, rv, rv2) into gdn, grv, grv2.
template <bool GRAD>
__device__ __forceinline__ float a2_rdon(const float dn[3], const float rv[3], const float rv2[3], float gbar,
                                         float gdn[3], float grv[3], float grv2[3]) {
    const float cos_t = -(dn[0] * rv[0] + dn[1] * rv[1] + dn[2] * rv[2]);
    const float xr = dn[0] * rv2[0] + dn[1] * rv2[1] + dn[2] * rv2[2];
    const float x = fminf(fmaxf(xr, -1.f), 1.f);
    const float ct0 = sqrtf(fmaxf(1.f - x * x, 0.f));
    float rdon = 0.f, g_ct0 = 0.f;
    if (cos_t >= 0.f) {
        const float cr[3] = {dn[1] * rv2[2] - dn[2] * rv2[1], dn[2] * rv2[0] - dn[0] * rv2[2],
                             dn[0] * rv2[1] - dn[1] * rv2[0]};
        const float rd2r = cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2];
        const float rd2 = fmaxf(rd2r, 1.0e-6f);
        const float inv = 1.f / sqrtf(rd2);
        const float ch[3] = {cr[0] * inv, cr[1] * inv, cr[2] * inv};
        const float yr = ch[0] * rv[0] + ch[1] * rv[1] + ch[2] * rv[2];
        const float y = fminf(fmaxf(yr, -1.f), 1.f);
        const float sq = sqrtf(fmaxf(1.f - y * y, 0.f));
        const float sin2ti = 2.f * fabsf(y) * sq;
        rdon = (0.9f + 0.1f * sin2ti) * ct0;
        if (GRAD) {
            g_ct0 = gbar * (0.9f + 0.1f * sin2ti);
            const float g_s = gbar * 0.1f * ct0;
            float g_y = 0.f;
            if (yr > -1.f && yr < 1.f && sq > 1.0e-6f) g_y = g_s * 2.f * copysignf(1.f, y) * (1.f - 2.f * y * y) / sq;
            // y = ch . rv
            float gch[3];
            for (int k = 0; k < 3; k++) {
                grv[k] += g_y * ch[k];
                gch[k] = g_y * rv[k];
            }
            // ch = cr / |cr|
            float gcr[3];
            if (rd2r > 1.0e-6f) {
                const float dot = gch[0] * ch[0] + gch[1] * ch[1] + gch[2] * ch[2];
                for (int k = 0; k < 3; k++) gcr[k] = (gch[k] - dot * ch[k]) * inv;
            } else {
                for (int k = 0; k < 3; k++) gcr[k] = gch[k] * inv;
            }
            // cr = dn x rv2: d dn = rv2 x gcr, d rv2 = gcr x dn
            gdn[0] += rv2[1] * gcr[2] - rv2[2] * gcr[1];
            gdn[1] += rv2[2] * gcr[0] - rv2[0] * gcr[2];
            gdn[2] += rv2[0] * gcr[1] - rv2[1] * gcr[0];
            grv2[0] += gcr[1] * dn[2] - gcr[2] * dn[1];
            grv2[1] += gcr[2] * dn[0] - gcr[0] * dn[2];
            grv2[2] += gcr[0] * dn[1] - gcr[1] * dn[0];
        }
    } else if (cos_t >= -0.34202f) {
        const float u = 0.116978f - cos_t * cos_t;
        rdon = 562.25f * (u * u * u) * ct0;
        if (GRAD) {
            g_ct0 = gbar * 562.25f * (u * u * u);
            const float g_c = gbar * 562.25f * 3.f * u * u * (-2.f * cos_t) * ct0;
            for (int k = 0; k < 3; k++) {
                gdn[k] -= g_c * rv[k];
                grv[k] -= g_c * dn[k];
            }
        }
    }
    if (GRAD && g_ct0 != 0.f && xr > -1.f && xr < 1.f && ct0 > 1.0e-6f) {
        const float g_x = -g_ct0 * x / ct0;
        for (int k = 0; k < 3; k++) {
            gdn[k] += g_x * rv2[k];
            grv2[k] += g_x * dn[k];
        }
    }
    return rdon;
}

__device__ __forceinline__ void point_of(const AParams &p, long long pt, int &ix, int &iy, int &iz, float c[3]) {
    ix = (int)(pt % p.n1[0]);
    iy = (int)((pt / p.n1[0]) % p.n1[1]);
    iz = (int)(pt / ((long long)p.n1[0] * p.n1[1]));
    c[0] = (float)(ix - p.ne[0]) * p.h;
    c[1] = (float)(iy - p.ne[1]) * p.h;
    c[2] = (float)(iz - p.ne[2]) * p.h;
}

// Cell range covering the 8 A sphere around c.
__device__ __forceinline__ void cell_range(const AParams &p, const float c[3], int lo[3], int hi[3]) {
    for (int d = 0; d < 3; d++) {
        lo[d] = max(0, (int)floorf((c[d] - 8.f - p.corg[d]) / kCellA));
        hi[d] = min(p.nc[d] - 1, (int)floorf((c[d] + 8.f - p.corg[d]) / kCellA));
    }
}

// Closest donor H within 8 A of c (lowest index on ties), or -1.
__device__ __forceinline__ int closest_h(const AParams &p, const float c[3]) {
    int lo[3], hi[3];
    cell_range(p, c, lo, hi);
    float best = 1e30f;
    int bidx = -1;
    for (int gz = lo[2]; gz <= hi[2]; gz++)
        for (int gy = lo[1]; gy <= hi[1]; gy++) {
            const int row = p.nc[0] * (gy + p.nc[1] * gz);
            const int s0 = p.cell_start[row + lo[0]], s1 = p.cell_start[row + hi[0] + 1];
            for (int s = s0; s < s1; s++) {
                const AAtom &a = p.atoms[s];
                const int hb = (a.meta >> 8) & 0xff;
                if (hb != AG_HB_DS && hb != AG_HB_D1) continue;
                const float dx = a.x - c[0], dy = a.y - c[1], dz = a.z - c[2];
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 < best || (r2 == best && a.index < bidx)) {
                    best = r2;
                    bidx = a.index;
                }
            }
        }
    return best < kCut2 ? bidx : -1;
}

// Per-pair directional factors (forward).
struct Dir {
    float racc, rdon, Hramp;
};
__device__ __forceinline__ Dir dir_of(const AAtom &a, int hb, const float dn[3], int closest, const float *rvH) {
    Dir o{1.f, 1.f, 1.f};
    if (hb == AG_HB_D1) {
        const float c = -(dn[0] * a.rv[0] + dn[1] * a.rv[1] + dn[2] * a.rv[2]);
        float dc;
        o.racc = racc_of(c, (a.meta >> 16) & 0xff, dc);
        if (c > 0.f && a.index != closest && closest >= 0) {
            const float ct = rvH[0] * a.rv[0] + rvH[1] * a.rv[1] + rvH[2] * a.rv[2];
            float dct;
            o.Hramp = hramp_of(ct, dct);
        }
    } else if (hb == AG_HB_A1) {
        const float c = -(dn[0] * a.rv[0] + dn[1] * a.rv[1] + dn[2] * a.rv[2]);
        o.rdon = c > 0.f ? c * c : 0.f;
    } else if (hb == AG_HB_A2) {
        float g3[3] = {0, 0, 0};
        o.rdon = a2_rdon<false>(dn, a.rv, a.rv2, 0.f, g3, g3, g3);
    }
    return o;
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(kAT) k_an_near(AParams p) {
    const long long pt = (long long)blockIdx.x * kAT + threadIdx.x;
    if (pt >= p.P) return;
    int ix, iy, iz;
    float c[3];
    point_of(p, pt, ix, iy, iz, c);
    float hmin[kAMaxAff], hmax[kAMaxAff];
    int amin[kAMaxAff], amax[kAMaxAff];
    for (int k = 0; k < kAMaxAff; k++) {
        hmin[k] = 999999.f;
        hmax[k] = -999999.f;
        amin[k] = amax[k] = -1;
    }
    // slot-indexed accumulators live in the output maps; per affinity map the
    // E term goes to aff_slot and the desolvation to des_slot
    float accE[kAMaxAff], accD[kAMaxAff];
    for (int k = 0; k < kAMaxAff; k++) accE[k] = accD[k] = 0.f;
    float dacc = 0.f;
    const int ch = closest_h(p, c);
    float rvH[3] = {0.f, 0.f, 0.f};
    if (ch >= 0)
        for (int k = 0; k < 3; k++) rvH[k] = p.rv_by_index[3 * ch + k];
    int lo[3], hi[3];
    cell_range(p, c, lo, hi);
    for (int gz = lo[2]; gz <= hi[2]; gz++)
        for (int gy = lo[1]; gy <= hi[1]; gy++) {
            const int row = p.nc[0] * (gy + p.nc[1] * gz);
            const int s0 = p.cell_start[row + lo[0]], s1 = p.cell_start[row + hi[0] + 1];
            for (int s = s0; s < s1; s++) {
                const AAtom a = p.atoms[s];
                const float d[3] = {a.x - c[0], a.y - c[1], a.z - c[2]};
                const float r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (!(r2 < kCut2)) continue;
                // hardware rsq and exp2, about 1 ulp
                const float r2c = fmaxf(r2, 1.0e-12f);
                const float ir = __builtin_amdgcn_rsqf(r2c);
                const float r = r2c * ir;
                const float dn[3] = {d[0] * ir, d[1] * ir, d[2] * ir};
                const int type = a.meta & 0xff, hb = (a.meta >> 8) & 0xff;
                float dS;
                const float S = switch_s(r, dS);
                const float G = p.w_desolv * __builtin_amdgcn_exp2f(r2 * (-kInv2Sig2 * 1.44269504088896341f)) * S;
                const Dir dr = dir_of(a, hb, dn, ch, rvH);
                const float solq = a.solpar + p.qasp * fabsf(a.q);
#pragma unroll
                for (int k = 0; k < kAMaxAff; k++) {
                    if (k >= p.naff) continue;
                    const LJ lj = lj_eval(p.pairs[k * p.T + type], r, p.w_smooth);
                    const float E = lj.E * S;
                    const float rsph = fminf(fmaxf(E * 0.01f, 0.f), 1.f);
                    const int cs = p.cases[k][hb];
                    if (cs == CASE_PLAIN) {
                        accE[k] += E;
                    } else if (cs == CASE_ACC) {
                        accE[k] += E * dr.Hramp * (dr.racc + (1.f - dr.racc) * rsph);
                    } else {
                        const float f = cs == CASE_MM_RACC ? dr.racc : dr.rdon;
                        const float v = E * (f + (1.f - f) * rsph);
                        if (v < hmin[k]) { hmin[k] = v; amin[k] = a.index; }
                        if (v > hmax[k]) { hmax[k] = v; amax[k] = a.index; }
                    }
                    accD[k] += (p.sp[k] * a.vol + solq * p.volp[k]) * G;
                }
                dacc += p.qasp * a.vol * G;
            }
        }
    // assemble: a desolvation-only map belongs to exactly one affinity map
    // (separate_desolvation_maps), so slot writes do not collide
#pragma unroll
    for (int k = 0; k < kAMaxAff; k++) {
        if (k >= p.naff) continue;
        float v = accE[k];
        if (amin[k] >= 0) v += hmin[k] + fmaxf(hmax[k], 0.f);
        if (p.des_slot[k] == p.aff_slot[k]) {
            p.maps[(long long)p.aff_slot[k] * p.P + pt] = v + accD[k];
        } else {
            p.maps[(long long)p.aff_slot[k] * p.P + pt] = v;
            p.maps[(long long)p.des_slot[k] * p.P + pt] = accD[k];
        }
        p.amin[(long long)k * p.P + pt] = amin[k];
        p.amax[(long long)k * p.P + pt] = (amin[k] >= 0 && hmax[k] > 0.f) ? amax[k] : -1;
    }
    if (p.d_slot >= 0) p.maps[(long long)p.d_slot * p.P + pt] = dacc;
    p.closest_h[pt] = ch;
    for (int k = 0; k < p.n_cov; k++) {
        const double dx = p.cov_pos[0] - (double)(ix - p.ne[0]) * (double)p.h;
        const double dy = p.cov_pos[1] - (double)(iy - p.ne[1]) * (double)p.h;
        const double dz = p.cov_pos[2] - (double)(iz - p.ne[2]) * (double)p.h;
        double rc = sqrt(dx * dx + dy * dy + dz * dz) / p.cov_hw;
        if (rc < 1.0e-6) rc = 1.0e-6;
        p.maps[(long long)p.cov_slot[k] * p.P + pt] = (float)(p.cov_barrier * (1. - exp(-0.69314718055994530942 * rc * rc)));
    }
}

struct EParams {
    int n1[3], ne[3];
    float h;
    long long P;
    const float4 *atoms;      // (x, y, z, q * coeff_estat)
    int n_atoms;
    int dddiel;
    float inv_diel;           // 332 / diel for a constant dielectric
    float *emap;
    // forward in runs (k_an_e, k_an_esum)
    const float4 *fatoms;     // (cs x, cs y, cs z, w): coordinates scaled by cs, see k_an_e
    float sc;                 // cs * spacing
    float rcap;               // 2 / cs
    int echunk, nechunks;     // atoms per chunk, chunks
    float *epart;             // [nechunks][P] partial sums
    // adjoint
    const float *ge;          // cotangent of the e map
    float4 *partial;          // [chunks][n_atoms]: dx, dy, dz, dq
    int nchunks;
};

struct __align__(16) EAtom {
    float x, y, z, w;
};

template <int K>
struct ERun {
    static constexpr float mid = 0.5f * (K - 1);
    // (k - mid)^2 and the factor K - 1 - 2k of u * spacing
    static __device__ __forceinline__ constexpr float c2(int k) { return ((float)k - mid) * ((float)k - mid); }
    static __device__ __forceinline__ constexpr float m(int k) { return (float)(K - 1 - 2 * k); }
};

template <int K, bool DD>
__global__ void __launch_bounds__(64) k_an_e(EParams p) {
    const int run = blockIdx.x * 64 + threadIdx.x;
    const int rpr = (p.n1[0] + K - 1) / K;
    const int nruns = rpr * p.n1[1] * p.n1[2];
    const bool active = run < nruns;
    const int rr = active ? run : nruns - 1;   // an idle lane computes a real run and stores nothing
    const int row = rr / rpr;
    const int ix0 = (rr - row * rpr) * K;
    const int iy = row % p.n1[1], iz = row / p.n1[1];
    const float sc = p.sc;
    const float py = (float)(iy - p.ne[1]) * sc, pz = (float)(iz - p.ne[2]) * sc;
    const float cx = ((float)(ix0 - p.ne[0]) + ERun<K>::mid) * sc;
    const float s2 = sc * sc;
    const float bias = fmaxf(1.0e-12f, ERun<K>::mid * ERun<K>::mid * s2 * 0x1p-19f);
    const int a0 = blockIdx.y * p.echunk;
    const int a1 = min(a0 + p.echunk, p.n_atoms);
    const float rcap = p.rcap;   // 2 / cs
    float part[K];
#pragma unroll
    for (int k = 0; k < K; k++) part[k] = 0.f;
    // atoms are read-only here: scalar loads through the constant address space
    const __attribute__((address_space(4))) EAtom *atoms = (const __attribute__((address_space(4))) EAtom *)p.fatoms;
    auto ld = [&](int i) {
        const __attribute__((address_space(4))) EAtom &q = atoms[i];
        return EAtom{q.x, q.y, q.z, q.w};
    };
    if (a0 < a1) {
        EAtom nxt = ld(a0);
        for (int a = a0; a < a1; a++) {
            const EAtom at = nxt;
            nxt = ld(min(a + 1, a1 - 1));
            const float dy = at.y - py, dz = at.z - pz;
            const float dyz2 = __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, bias));
            const float u = at.x - cx;
            const float U = __builtin_fmaf(u, u, dyz2);
            const float us = u * sc;
#pragma unroll
            for (int k = 0; k < K; k++) {
                const float rb2 = __builtin_fmaf(us, ERun<K>::m(k), __builtin_fmaf(s2, ERun<K>::c2(k), U));
                const float rs = __builtin_amdgcn_rsqf(rb2);
                const float h = fminf(rs, rcap);
                if (DD) {
                    const float e2 = __builtin_amdgcn_exp2f(-(rb2 * rs));
                    const float den = __builtin_fmaf(kMsK, e2, 1.f);
                    const float g = den * __builtin_amdgcn_rcpf(__builtin_fmaf(kMsA, den, kMsB));
                    part[k] = __builtin_fmaf(at.w, h * g, part[k]);
                } else {
                    part[k] = __builtin_fmaf(at.w, h, part[k]);
                }
            }
        }
    }
    if (active) {
        float *out = p.epart + (long long)blockIdx.y * p.P + ix0 + (long long)p.n1[0] * (iy + (long long)p.n1[1] * iz);
#pragma unroll
        for (int k = 0; k < K; k++)
            if (ix0 + k < p.n1[0]) out[k] = part[k];
    }
}

// The e map: per point the chunks' partial sums in chunk order, Kahan-compensated.
__global__ void k_an_esum(const float *epart, int nchunk, long long P, float *emap) {
    const long long pt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pt >= P) return;
    float s = 0.f, c = 0.f;
    for (int j = 0; j < nchunk; j++) {
        const float y = epart[(long long)j * P + pt] - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    emap[pt] = s;
}

// ---------------------------------------------------------------------------
// adjoint
// ---------------------------------------------------------------------------

// Per point: the cotangent of the closest H's bond vector through the Hramp
// factor of every other D1 atom within 8 A.
__global__ void __launch_bounds__(kAT) k_an_vh(AParams p) {
    const long long pt = (long long)blockIdx.x * kAT + threadIdx.x;
    if (pt >= p.P) return;
    int ix, iy, iz;
    float c[3];
    point_of(p, pt, ix, iy, iz, c);
    const int ch = p.closest_h[pt];
    float v[3] = {0.f, 0.f, 0.f};
    if (ch >= 0) {
        float rvH[3];
        for (int k = 0; k < 3; k++) rvH[k] = p.rv_by_index[3 * ch + k];
        float g[kAMaxAff];
        bool any = false;
#pragma unroll
        for (int k = 0; k < kAMaxAff; k++) {
            g[k] = k < p.naff ? p.dmaps[(long long)p.aff_slot[k] * p.P + pt] : 0.f;
            any = any || (g[k] != 0.f && p.cases[k][AG_HB_D1] == CASE_ACC);
        }
        if (any) {
            int lo[3], hi[3];
            cell_range(p, c, lo, hi);
            for (int gz = lo[2]; gz <= hi[2]; gz++)
                for (int gy = lo[1]; gy <= hi[1]; gy++) {
                    const int row = p.nc[0] * (gy + p.nc[1] * gz);
                    const int s0 = p.cell_start[row + lo[0]], s1 = p.cell_start[row + hi[0] + 1];
                    for (int s = s0; s < s1; s++) {
                        const AAtom a = p.atoms[s];
                        const int hb = (a.meta >> 8) & 0xff;
                        if (hb != AG_HB_D1 || a.index == ch) continue;
                        const float d[3] = {a.x - c[0], a.y - c[1], a.z - c[2]};
                        const float r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                        if (!(r2 < kCut2)) continue;
                        const float r = fmaxf(sqrtf(r2), 1.0e-6f);
                        const float ir = 1.f / r;
                        const float dn[3] = {d[0] * ir, d[1] * ir, d[2] * ir};
                        const float cth = -(dn[0] * a.rv[0] + dn[1] * a.rv[1] + dn[2] * a.rv[2]);
                        if (!(cth > 0.f)) continue;
                        float dc;
                        const float racc = racc_of(cth, (a.meta >> 16) & 0xff, dc);
                        const float ct = rvH[0] * a.rv[0] + rvH[1] * a.rv[1] + rvH[2] * a.rv[2];
                        float dct;
                        hramp_of(ct, dct);
                        if (dct == 0.f) continue;
                        float dS;
                        const float S = switch_s(r, dS);
                        const int type = a.meta & 0xff;
                        float gH = 0.f;
#pragma unroll
                        for (int k = 0; k < kAMaxAff; k++) {
                            if (k >= p.naff || p.cases[k][AG_HB_D1] != CASE_ACC || g[k] == 0.f) continue;
                            const float E = lj_eval(p.pairs[k * p.T + type], r, p.w_smooth).E * S;
                            const float rsph = fminf(fmaxf(E * 0.01f, 0.f), 1.f);
                            gH += g[k] * E * (racc + (1.f - racc) * rsph);
                        }
                        for (int k = 0; k < 3; k++) v[k] += gH * dct * a.rv[k];
                    }
                }
        }
    }
    for (int k = 0; k < 3; k++) p.vh[3 * pt + k] = v[k];
}

// Block reduction of n floats per lane (fixed order): lane 0 of the block
// gets the sums.
template <int NV>
__device__ __forceinline__ void block_sum(float (&v)[NV], float *red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        float x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
        v[k] = x;
    }
    __syncthreads();
    if (lane == 0)
        for (int k = 0; k < NV; k++) red[wave * NV + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; k++) v[k] = red[k] + red[NV + k] + red[2 * NV + k] + red[3 * NV + k];
}

struct HsumParams {
    const int *closest_h;
    const float *vh;
    const int *list;         // points with a nonzero vh, in point order
    const int *count;        // their number (device)
    const int *h_atoms;      // donor H atom indices, one per workgroup
    float *out;              // [n][3]
};

__global__ void __launch_bounds__(kAT) k_an_hsum(HsumParams p) {
    __shared__ float red[4 * 3];
    const int h = p.h_atoms[blockIdx.x];
    const int n = *p.count;
    float v[3] = {0.f, 0.f, 0.f};
    for (int i = threadIdx.x; i < n; i += kAT) {
        const long long pt = p.list[i];
        if (p.closest_h[pt] == h)
            for (int k = 0; k < 3; k++) v[k] += p.vh[3 * pt + k];
    }
    block_sum<3>(v, red);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) p.out[3 * blockIdx.x + k] = v[k];
}

// Per-atom outputs of k_an_atom: x(3) rv(3) rv2(3) q wdes, then R and eps
// per affinity map.
constexpr int kAtomOut = 11 + 2 * kAMaxAff;

// Per point: 1 if any cotangent the atom gather reads there (affinity,
// desolvation and d slots) is nonzero.  Where all are zero a point adds
// exactly nothing to any atom's gradient, so k_an_atom skips it.
__global__ void __launch_bounds__(kAT) k_an_gmask(AParams p, unsigned char *mask) {
    const long long pt = (long long)blockIdx.x * kAT + threadIdx.x;
    if (pt >= p.P) return;
    bool any = p.d_slot >= 0 && p.dmaps[(long long)p.d_slot * p.P + pt] != 0.f;
    for (int k = 0; k < p.naff; k++)
        any = any || p.dmaps[(long long)p.aff_slot[k] * p.P + pt] != 0.f ||
              p.dmaps[(long long)p.des_slot[k] * p.P + pt] != 0.f;
    mask[pt] = any;
}

struct AtomAdjParams {
    AParams a;
    const AAtom *by_index;   // atoms by index
    const int *atom_list;    // atoms whose 8 A sphere meets the box, one per workgroup
    float *out;
    const unsigned char *mask;   // k_an_gmask              // [n_list][kAtomOut]
};

__global__ void __launch_bounds__(kAT) k_an_atom(AtomAdjParams q) {
    __shared__ float red[4 * kAtomOut];
    const AParams &p = q.a;
    const int idx = q.atom_list[blockIdx.x];
    const AAtom a = q.by_index[idx];
    const int type = a.meta & 0xff, hb = (a.meta >> 8) & 0xff, rexp = (a.meta >> 16) & 0xff;
    // points of the box within the 8 A cube around the atom
    int lo[3], hi[3];
    const float ac[3] = {a.x, a.y, a.z};
    for (int d = 0; d < 3; d++) {
        lo[d] = max(0, (int)ceilf((ac[d] - 8.f) / p.h) + p.ne[d]);
        hi[d] = min(p.n1[d] - 1, (int)floorf((ac[d] + 8.f) / p.h) + p.ne[d]);
    }
    const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
    const long long npt = (nx > 0 && ny > 0 && nz > 0) ? (long long)nx * ny * nz : 0;
    float acc[kAtomOut];
#pragma unroll
    for (int k = 0; k < kAtomOut; k++) acc[k] = 0.f;
    const float solq = a.solpar + p.qasp * fabsf(a.q);
    const float sgnq = a.q > 0.f ? 1.f : (a.q < 0.f ? -1.f : 0.f);
    // The points of the cube within 8 A whose cotangents are not all zero
    // (k_an_gmask) are appended to an LDS list in cube order, a tile of kAT
    // cube points at a time; each full kAT of the list is then processed
    // with every lane busy.  The rest add exactly nothing.
    __shared__ int list[2 * kAT];
    __shared__ int wcount[kAT / 64];
    int nlist = 0;
    for (long long t0 = 0; t0 < npt || nlist > 0; t0 += kAT) {
        if (t0 < npt) {
            const long long j = t0 + threadIdx.x;
            bool keep = false;
            int cell = 0;
            if (j < npt) {
                const int ix = lo[0] + (int)(j % nx), iy = lo[1] + (int)((j / nx) % ny),
                          iz = lo[2] + (int)(j / ((long long)nx * ny));
                const long long pt = (long long)ix + (long long)p.n1[0] * ((long long)iy + (long long)p.n1[1] * iz);
                const float dx = a.x - (float)(ix - p.ne[0]) * p.h, dy = a.y - (float)(iy - p.ne[1]) * p.h,
                            dz = a.z - (float)(iz - p.ne[2]) * p.h;
                keep = dx * dx + dy * dy + dz * dz < kCut2 && q.mask[pt];
                cell = (int)j;
            }
            const unsigned long long m = __ballot(keep);
            if ((threadIdx.x & 63) == 0) wcount[threadIdx.x >> 6] = __popcll(m);
            __syncthreads();
            int before = nlist;
            for (int w = 0; w < (int)(threadIdx.x >> 6); w++) before += wcount[w];
            if (keep) list[before + __popcll(m & ((1ull << (threadIdx.x & 63)) - 1))] = cell;
            int added = 0;
            for (int w = 0; w < kAT / 64; w++) added += wcount[w];
            nlist += added;
            __syncthreads();
            if (nlist < kAT && t0 + kAT < npt) continue;   // gather a full tile first
        }
        // process the first min(nlist, kAT) entries, then move the rest down
        const int take = min(nlist, kAT);
        if ((int)threadIdx.x < take) {
        const long long j = list[threadIdx.x];
        const int ix = lo[0] + (int)(j % nx), iy = lo[1] + (int)((j / nx) % ny), iz = lo[2] + (int)(j / ((long long)nx * ny));
        const long long pt = (long long)ix + (long long)p.n1[0] * ((long long)iy + (long long)p.n1[1] * iz);
        const float c[3] = {(float)(ix - p.ne[0]) * p.h, (float)(iy - p.ne[1]) * p.h, (float)(iz - p.ne[2]) * p.h};
        const float d[3] = {a.x - c[0], a.y - c[1], a.z - c[2]};
        const float r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const float r2c = fmaxf(r2, 1.0e-12f);
        const float ir = __builtin_amdgcn_rsqf(r2c);   // hardware rsq and exp2, about 1 ulp
        const float r = r2c * ir;
        const float dn[3] = {d[0] * ir, d[1] * ir, d[2] * ir};
        float dS;
        const float S = switch_s(r, dS);
        const float ex = __builtin_amdgcn_exp2f(r2 * (-kInv2Sig2 * 1.44269504088896341f));
        const float G0 = p.w_desolv * ex;               // sol(r)
        const float dG0 = -2.f * r * kInv2Sig2 * G0;    // d sol / dr
        const int ch = p.closest_h[pt];
        float rvH[3] = {0.f, 0.f, 0.f};
        if (ch >= 0)
            for (int k = 0; k < 3; k++) rvH[k] = p.rv_by_index[3 * ch + k];
        // directional factors and the pieces their derivatives need
        float racc = 1.f, rdon = 1.f, Hramp = 1.f, dracc_dc = 0.f, dH_dct = 0.f, cth = 0.f;
        if (hb == AG_HB_D1) {
            cth = -(dn[0] * a.rv[0] + dn[1] * a.rv[1] + dn[2] * a.rv[2]);
            racc = racc_of(cth, rexp, dracc_dc);
            if (cth > 0.f && idx != ch && ch >= 0) {
                const float ct = rvH[0] * a.rv[0] + rvH[1] * a.rv[1] + rvH[2] * a.rv[2];
                Hramp = hramp_of(ct, dH_dct);
            }
        } else if (hb == AG_HB_A1) {
            cth = -(dn[0] * a.rv[0] + dn[1] * a.rv[1] + dn[2] * a.rv[2]);
            rdon = cth > 0.f ? cth * cth : 0.f;
        } else if (hb == AG_HB_A2) {
            float g3[3] = {0, 0, 0};
            rdon = a2_rdon<false>(dn, a.rv, a.rv2, 0.f, g3, g3, g3);
        }
        float g_r = 0.f, g_racc = 0.f, g_rdon = 0.f, g_H = 0.f;
#pragma unroll
        for (int k = 0; k < kAMaxAff; k++) {
            if (k >= p.naff) continue;
            const float gE_up = p.dmaps[(long long)p.aff_slot[k] * p.P + pt];
            const float gD_up = p.dmaps[(long long)p.des_slot[k] * p.P + pt];
            const LJ lj = lj_eval(p.pairs[k * p.T + type], r, p.w_smooth);
            const float E = lj.E * S;
            const float rsph = fminf(fmaxf(E * 0.01f, 0.f), 1.f);
            const float drsph = (E > 0.f && E < 100.f) ? 0.01f : 0.f;
            const int cs = p.cases[k][hb];
            float gE = 0.f;   // dL/dE of this map's pair term
            if (cs == CASE_PLAIN) {
                gE = gE_up;
            } else if (cs == CASE_ACC) {
                const float mix = racc + (1.f - racc) * rsph;
                gE = gE_up * Hramp * (mix + E * (1.f - racc) * drsph);
                g_racc += gE_up * E * Hramp * (1.f - rsph);
                g_H += gE_up * E * mix;
            } else {
                const float wmm = (p.amin[(long long)k * p.P + pt] == idx ? 1.f : 0.f) +
                                  (p.amax[(long long)k * p.P + pt] == idx ? 1.f : 0.f);
                if (wmm != 0.f) {
                    const float f = cs == CASE_MM_RACC ? racc : rdon;
                    const float mix = f + (1.f - f) * rsph;
                    const float gu = gE_up * wmm;
                    gE = gu * (mix + E * (1.f - f) * drsph);
                    if (cs == CASE_MM_RACC) g_racc += gu * E * (1.f - rsph);
                    else g_rdon += gu * E * (1.f - rsph);
                }
            }
            // E = lj.E * S
            g_r += gE * (lj.dEdr * S + lj.E * dS);
            acc[11 + k] += gE * lj.dEdR * S;
            acc[11 + kAMaxAff + k] += gE * lj.dEdeps * S;
            // desolvation (sp vol + solq volp) sol S
            const float coef = p.sp[k] * a.vol + solq * p.volp[k];
            g_r += gD_up * coef * (dG0 * S + G0 * dS);
            acc[9] += gD_up * p.qasp * sgnq * p.volp[k] * G0 * S;
            acc[10] += gD_up * coef * ex * S;   // d / d coeff_desolv
        }
        if (p.d_slot >= 0) {
            const float gd = p.dmaps[(long long)p.d_slot * p.P + pt];
            g_r += gd * p.qasp * a.vol * (dG0 * S + G0 * dS);
            acc[10] += gd * p.qasp * a.vol * ex * S;
        }
        // back through the directional factors to dn and the bond vectors
        float gdn[3] = {0.f, 0.f, 0.f};
        if (hb == AG_HB_D1) {
            const float g_c = g_racc * dracc_dc;
            for (int k = 0; k < 3; k++) {
                gdn[k] -= g_c * a.rv[k];
                acc[3 + k] -= g_c * dn[k];
                acc[3 + k] += g_H * dH_dct * rvH[k];
            }
        } else if (hb == AG_HB_A1) {
            const float g_c = cth > 0.f ? g_rdon * 2.f * cth : 0.f;
            for (int k = 0; k < 3; k++) {
                gdn[k] -= g_c * a.rv[k];
                acc[3 + k] -= g_c * dn[k];
            }
        } else if (hb == AG_HB_A2 && g_rdon != 0.f) {
            float grv[3] = {0.f, 0.f, 0.f}, grv2[3] = {0.f, 0.f, 0.f};
            a2_rdon<true>(dn, a.rv, a.rv2, g_rdon, gdn, grv, grv2);
            for (int k = 0; k < 3; k++) {
                acc[3 + k] += grv[k];
                acc[6 + k] += grv2[k];
            }
        }
        // dn = d / r, r = |d|, d = x - c
        const float dot = gdn[0] * dn[0] + gdn[1] * dn[1] + gdn[2] * dn[2];
        for (int k = 0; k < 3; k++) acc[k] += (gdn[k] - dot * dn[k]) * ir + g_r * dn[k];
        }
        __syncthreads();
        const int rest = nlist - take;
        const int moved = (int)threadIdx.x < rest ? list[take + threadIdx.x] : 0;
        __syncthreads();
        if ((int)threadIdx.x < rest) list[threadIdx.x] = moved;
        nlist = rest;
        __syncthreads();
        if (t0 >= npt && nlist == 0) break;
    }
    block_sum<kAtomOut>(acc, red);
    if (threadIdx.x == 0)
        for (int k = 0; k < kAtomOut; k++) q.out[(long long)blockIdx.x * kAtomOut + k] = acc[k];
}

// e-map adjoint: lane = atom, workgroup = (atom block, point chunk).
// The e map's adjoint gathers, per atom, over the points where the
// cotangent g is nonzero (the poses' trilinear footprint, a fraction of the
// box).  Those points are packed first, in point order, as (x, y, z, g):
// k_an_gcount counts them per block of kGPack points, k_an_gscan turns the
// counts into offsets (one workgroup) and k_an_gpack writes them.
constexpr int kGPack = 1024;             // points per packing block (4 per lane)
constexpr int kEAdjChunk = 1024;         // packed points per e-adjoint chunk

// Points where a predicate holds, packed in point order: k_an_count counts
// them per block of kGPack points, k_an_gscan turns the counts into offsets
// (one workgroup) and k_an_pack writes their indices.
struct NonzeroE {             // the e map's cotangent
    const float *ge;
    __device__ bool operator()(long long pt) const { return ge[pt] != 0.f; }
};
struct NonzeroVh {            // the closest H's bond-vector cotangent (k_an_vh)
    const float *vh;
    __device__ bool operator()(long long pt) const {
        return vh[3 * pt] != 0.f || vh[3 * pt + 1] != 0.f || vh[3 * pt + 2] != 0.f;
    }
};

template <typename Pred>
__global__ void __launch_bounds__(256) k_an_count(Pred pred, long long P, int *counts) {
    __shared__ int wc[4];
    const long long b = (long long)blockIdx.x * kGPack;
    int n = 0;
    for (int k = 0; k < 4; k++) {
        const long long pt = b + k * 256 + threadIdx.x;
        n += pt < P && pred(pt);
    }
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o, 64);
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

// Exclusive offsets of the block counts, and their total in offs[n].
__global__ void __launch_bounds__(1024) k_an_gscan(const int *counts, int n, int *offs) {
    __shared__ int ws[16];
    const int per = (n + 1023) / 1024;
    const int lo = min(n, threadIdx.x * per), hi = min(n, lo + per);
    int t = 0;
    for (int i = lo; i < hi; i++) t += counts[i];
    // inclusive scan of the lane sums: within each wave, then across waves
    int v = t;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    if (lane == 63) ws[w] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 1; k < 16; k++) ws[k] += ws[k - 1];
    __syncthreads();
    int run = v - t + (w ? ws[w - 1] : 0);
    for (int i = lo; i < hi; i++) {
        offs[i] = run;
        run += counts[i];
    }
    if (threadIdx.x == 1023) offs[n] = ws[15];
}

template <typename Pred>
__global__ void __launch_bounds__(256) k_an_pack(Pred pred, long long P, const int *offs, int *packed) {
    __shared__ int wc[4][4];
    const long long b = (long long)blockIdx.x * kGPack;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned long long m[4];
    for (int k = 0; k < 4; k++) {
        const long long pt = b + k * 256 + threadIdx.x;
        m[k] = __ballot(pt < P && pred(pt));
        if (lane == 0) wc[k][w] = __popcll(m[k]);
    }
    __syncthreads();
    int base = offs[blockIdx.x];
    for (int k = 0; k < 4; k++) {
        int before = base;
        for (int j = 0; j < w; j++) before += wc[k][j];
        if ((m[k] >> lane) & 1) packed[before + __popcll(m[k] & ((1ull << lane) - 1))] = (int)(b + k * 256 + threadIdx.x);
        for (int j = 0; j < 4; j++) base += wc[k][j];
    }
}

// Count, offsets and indices of the points where pred holds; the count
// lands in offs[nblk] (device).
template <typename Pred>
static void pack_points(Pred pred, long long P, int *counts, int *offs, int *packed, hipStream_t st) {
    const int nblk = (int)((P + kGPack - 1) / kGPack);
    k_an_count<<<nblk, 256, 0, st>>>(pred, P, counts);
    k_an_gscan<<<1, 1024, 0, st>>>(counts, nblk, offs);
    k_an_pack<<<nblk, 256, 0, st>>>(pred, P, offs, packed);
}

// One wave per (64 atoms, chunk of kEAdjChunk packed points): per atom
// sum_p g f(r) and sum_p g d(q w f(r))/d(atom position), f = 332 / (eps(r)
// max(r, 0.5)) (inv_diel / max(r, 0.5) for a constant dielectric), with
// hardware rsq / exp2 / rcp.  Chunks past the packed count write zeros, so
// k_an_ered's fixed-order sum needs no count.
template <bool DD>
__global__ void __launch_bounds__(64) k_an_eadj(EParams p, const int *packed, const int *count) {
    __shared__ float4 sp[256];
    const int ai = blockIdx.x * 64 + threadIdx.x;
    const bool act = ai < p.n_atoms;
    const float4 at = act ? p.atoms[ai] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int npk = *count;
    const int p0 = blockIdx.y * kEAdjChunk, p1 = min(npk, p0 + kEAdjChunk);
    float gx = 0.f, gy = 0.f, gz = 0.f, gq = 0.f;
    constexpr float cs = kMsLB * 1.44269504088896341f;      // exp(-lambda B r) = 2^(-cs r)
    constexpr float dk = -kMsB * kMsK * kMsLB;             // d(1/eps)/dr = dk e y^2 (see below)
    for (int b = p0; b < p1; b += 256) {
        __syncthreads();
        if (b + (int)threadIdx.x * 4 < p1)
            for (int k = 0; k < 4; k++) {
                const int q = b + threadIdx.x * 4 + k;
                float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                if (q < p1) {
                    const long long pt = packed[q];
                    const int ix = (int)(pt % p.n1[0]), iy = (int)((pt / p.n1[0]) % p.n1[1]),
                              iz = (int)(pt / ((long long)p.n1[0] * p.n1[1]));
                    v = make_float4((float)(ix - p.ne[0]) * p.h, (float)(iy - p.ne[1]) * p.h,
                                    (float)(iz - p.ne[2]) * p.h, p.ge[pt]);
                }
                sp[threadIdx.x * 4 + k] = v;
            }
        __syncthreads();
        const int nb = min(256, p1 - b);
        for (int j = 0; j < nb; j++) {
            const float4 P = sp[j];
            const float dx = at.x - P.x, dy = at.y - P.y, dz = at.z - P.z;
            const float r2 = fmaxf(__builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz)), 1.0e-12f);
            const float rs = __builtin_amdgcn_rsqf(r2);
            const bool beyond = rs < 2.f;                   // r > 0.5
            const float rm = beyond ? rs : 2.f;             // 1 / max(r, 0.5)
            const float drm = beyond ? -rs * rs : 0.f;      // d(1 / max(r, 0.5))/dr
            float f, dfdr;
            if (DD) {
                // 1/eps = den y, den = 1 + k e, y = 1 / (A den + B), e = exp(-lambda B r):
                // d(1/eps)/dr = B y^2 dden/dr = -B k lambda B e y^2
                const float e = __builtin_amdgcn_exp2f(-cs * (r2 * rs));
                const float den = __builtin_fmaf(kMsK, e, 1.f);
                const float y = __builtin_amdgcn_rcpf(__builtin_fmaf(kMsA, den, kMsB));
                const float G = den * y;
                const float dG = dk * e * (y * y);
                f = 332.f * G * rm;
                dfdr = 332.f * __builtin_fmaf(dG, rm, G * drm);
            } else {
                f = p.inv_diel * rm;
                dfdr = p.inv_diel * drm;
            }
            gq = __builtin_fmaf(P.w, f, gq);
            const float sc = P.w * at.w * dfdr * rs;       // g q w f'(r) / r
            gx = __builtin_fmaf(sc, dx, gx);
            gy = __builtin_fmaf(sc, dy, gy);
            gz = __builtin_fmaf(sc, dz, gz);
        }
    }
    if (act) p.partial[(long long)blockIdx.y * p.n_atoms + ai] = make_float4(gx, gy, gz, gq);
}

__global__ void k_an_ered(const float4 *partial, int n_atoms, int nchunks, double *out) {
    const int ai = blockIdx.x * blockDim.x + threadIdx.x;
    if (ai >= n_atoms) return;
    double s[4] = {0, 0, 0, 0};
    for (int c = 0; c < nchunks; c++) {
        const float4 v = partial[(long long)c * n_atoms + ai];
        s[0] += v.x;
        s[1] += v.y;
        s[2] += v.z;
        s[3] += v.w;
    }
    for (int k = 0; k < 4; k++) out[4 * ai + k] = s[k];
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct ag_analytic {
    ag_ctx *ctx = nullptr;
    int N = 0, M = 0, T = 0, n_maps = 0, naff = 0;
    ag_grid grid{};
    long long P = 0;
    // map set
    std::vector<ag_map_desc> maps;
    int elec = -1, dsolv = -1, dddiel = 0, separate = 0, hydrogen_type = -1;
    double inv_diel = 0, qasp = 0;
    double cov_pos[3] = {0, 0, 0}, cov_hw = 1, cov_barrier = 0;
    std::vector<int> aff_slot, des_slot, cov_slots;
    std::vector<std::array<int, 7>> cases;
    // setup (bond vectors, pair forms)
    ag_setup setup{};
    std::vector<ag_pair> pair_form;    // xA, xB, hbonder, explicit cA/cB ([M][T])
    double r_smooth = 0.5;
    // parameters
    std::vector<double> coord, charge, vol, solpar;
    std::vector<int32_t> type, hbond;
    std::vector<double> rij, eps0;     // [M][T]; eps0 unweighted
    double w[4] = {1, 1, 1, 1};        // vdW, H-bond, electrostatic, desolvation weights
    // derived
    std::vector<int32_t> rexp;
    std::vector<double> rv, rv2;
    std::vector<uint8_t> dis;
    std::vector<ag::BondRecipe> recipes;
    std::vector<int> h_atoms, atom_list;
    // device
    DevBuf d_atoms, d_by_index, d_rv, d_cells, d_pairs, d_eatoms, d_closest, d_amin, d_amax, d_vh, d_hout,
        d_hatoms, d_alist, d_aout, d_partial, d_eout, d_fatoms, d_epart, d_gcount, d_gpack, d_vcount, d_vpack,
        d_gmask;
    int eK = 16;                       // run length of the forward e kernel
    AParams ap{};
    EParams ep{};
    int nchunks = 0;
    bool uploaded = false;
    float *last_maps = nullptr;
};

namespace agdev {

// (probe class, atom class) -> case of main.cpp:3286-3374 without disorder.
static int an_case(int mh, int ah) {
    const bool atom_don = (ah == 1 || ah == 2 || ah == 6);
    if ((mh == 3 || mh == 5 || mh == 6) && atom_don) return CASE_ACC;
    if ((mh == 4 || mh == 6) && atom_don) return CASE_MM_RACC;
    if ((mh == 1 || mh == 2 || mh == 6) && (ah > 2 || ah == 6)) return CASE_MM_RDON;
    return CASE_PLAIN;
}

ag_analytic *analytic_create(ag_ctx *c, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                             const ag_setup *setup, const double w0[4]) {
    if (ms->disorder_h) ag::fail(AG_ERR_UNSUPPORTED, "disorder_h is not part of the analytic model");
    if (ms->bhtree) ag::fail(AG_ERR_UNSUPPORTED, "the tree build's semantics are not part of the analytic model");
    if (setup->n_atom_maps != ms->n_atom_maps) ag::fail(AG_ERR_BAD_INPUT, "setup and map set disagree");
    if (rec->n_atoms < 0 || rec->n_atoms > AG_MAX_ATOMS) ag::fail(AG_ERR_BAD_INPUT, "atom count out of range");
    for (int d = 0; d < 3; d++)
        if (grid->n1[d] < 1 || grid->n1[d] > AG_MAX_GRID_PTS) ag::fail(AG_ERR_BAD_INPUT, "grid out of range");
    for (int k = 0; k < 4; k++)
        if (!(w0[k] > 0)) ag::fail(AG_ERR_BAD_INPUT, "weights must be positive");
    std::unique_ptr<ag_analytic> a(new ag_analytic());
    a->ctx = c;
    const int N = a->N = rec->n_atoms;
    const int M = a->M = ms->n_atom_maps;
    const int T = a->T = setup->n_rec_types;
    a->n_maps = ms->n_maps;
    a->grid = *grid;
    a->P = (long long)grid->n1[0] * grid->n1[1] * grid->n1[2];
    a->maps.assign(ms->maps, ms->maps + M);
    a->elec = ms->elec_index;
    a->dsolv = ms->dsolv_index;
    a->dddiel = ms->dddiel;
    a->separate = ms->separate_desolvation;
    a->hydrogen_type = ms->hydrogen_type;
    a->inv_diel = ms->inv_diel_cal;
    a->qasp = ms->solpar_q;
    for (int d = 0; d < 3; d++) a->cov_pos[d] = ms->cov_pos[d];
    a->cov_hw = ms->cov_half_width;
    a->cov_barrier = ms->cov_barrier;
    a->setup = *setup;
    a->r_smooth = setup->r_smooth;
    a->pair_form.assign(setup->pairs, setup->pairs + (size_t)M * T);
    a->setup.pairs = a->pair_form.data();
    a->coord.assign(rec->coord, rec->coord + 3 * (size_t)N);
    a->charge.assign(rec->charge, rec->charge + N);
    a->vol.assign(rec->vol, rec->vol + N);
    a->solpar.assign(rec->solpar, rec->solpar + N);
    a->type.assign(rec->type, rec->type + N);
    a->hbond.assign(rec->hbond, rec->hbond + N);
    for (int i = 0; i < N; i++)
        if (a->type[i] < 0 || a->type[i] >= T || a->hbond[i] < 0 || a->hbond[i] > 6)
            ag::fail(AG_ERR_BAD_INPUT, "atom type or H-bond class out of range");
    for (int k = 0; k < 4; k++) a->w[k] = w0[k];
    a->rij.assign((size_t)M * T, 0.0);
    a->eps0.assign((size_t)M * T, 0.0);
    for (size_t k = 0; k < (size_t)M * T; k++) {
        const ag_pair &pp = a->pair_form[k];
        a->rij[k] = pp.Rij;
        a->eps0[k] = pp.eps / (pp.hbonder ? w0[1] : w0[0]);
    }
    // map roles: affinity maps carry E terms, their desolvation goes to the
    // map itself or, with separate_desolvation_maps, to the next one
    for (int m = 0; m < M; m++) {
        if (a->maps[m].is_covalent) {
            a->cov_slots.push_back(m);
            continue;
        }
        a->aff_slot.push_back(m);
        a->des_slot.push_back(a->separate ? m + 1 : m);
        std::array<int, 7> cs{};
        for (int ah = 0; ah <= 6; ah++) cs[ah] = a->maps[m].is_hbonder ? an_case(a->maps[m].hbond, ah) : CASE_PLAIN;
        a->cases.push_back(cs);
        if (a->separate) m++;
    }
    a->naff = (int)a->aff_slot.size();
    if (a->naff > kAMaxAff) ag::fail(AG_ERR_UNSUPPORTED, "more than 8 affinity maps in analytic mode");
    if (a->cov_slots.size() > 16) ag::fail(AG_ERR_UNSUPPORTED, "more than 16 covalent maps");
    return a.release();
}

void analytic_destroy(ag_analytic *a) {
    if (!a) return;
    (void)hipSetDevice(a->ctx->device);
    (void)hipStreamSynchronize(a->ctx->stream);
    delete a;
}

void analytic_set(ag_analytic *a, const double *coord, const double *charge, const double *rij, const double *eps0,
                  const double *w) {
    if (coord) a->coord.assign(coord, coord + 3 * (size_t)a->N);
    if (charge) a->charge.assign(charge, charge + a->N);
    if (rij) a->rij.assign(rij, rij + (size_t)a->M * a->T);
    if (eps0) a->eps0.assign(eps0, eps0 + (size_t)a->M * a->T);
    if (w)
        for (int k = 0; k < 4; k++) {
            if (!(w[k] > 0)) ag::fail(AG_ERR_BAD_INPUT, "weights must be positive");
            a->w[k] = w[k];
        }
    a->uploaded = false;
}

// Host precompute and upload of everything the kernels read.
static void analytic_upload(ag_analytic *a) {
    ag_ctx *c = a->ctx;
    hipStream_t st = c->stream;
    const int N = a->N, T = a->T;
    const ag_grid &g = a->grid;
    // bond vectors with their construction recipes
    a->rexp.assign(N, 0);
    a->rv.assign(3 * (size_t)N, 0.0);
    a->rv2.assign(3 * (size_t)N, 0.0);
    a->dis.assign(N, 0);
    a->recipes.assign(N, ag::BondRecipe{});
    ag::build_bond_vectors(a->setup, N, a->coord.data(), a->type.data(), a->hbond.data(), a->rexp.data(),
                           a->rv.data(), a->rv2.data(), a->dis.data(), nullptr, a->recipes.data());
    // cells over the box padded by 8 A
    float corg[3];
    int nc[3];
    for (int d = 0; d < 3; d++) {
        const double lo = -g.ne[d] * g.spacing - 8.1, hi = (g.n1[d] - 1 - g.ne[d]) * g.spacing + 8.1;
        corg[d] = (float)lo;
        nc[d] = (int)std::ceil((hi - lo) / kCellA) + 1;
    }
    const int ncells = nc[0] * nc[1] * nc[2];
    std::vector<int> cell_of(N, -1), count(ncells + 1, 0);
    std::vector<AAtom> by_index(N);
    for (int i = 0; i < N; i++) {
        AAtom &t = by_index[i];
        memset(&t, 0, sizeof t);
        t.x = (float)a->coord[3 * i];
        t.y = (float)a->coord[3 * i + 1];
        t.z = (float)a->coord[3 * i + 2];
        t.q = (float)a->charge[i];
        t.vol = (float)a->vol[i];
        t.solpar = (float)a->solpar[i];
        t.meta = (a->type[i] & 0xff) | ((a->hbond[i] & 0xff) << 8) | ((a->rexp[i] & 0xff) << 16);
        t.index = i;
        for (int k = 0; k < 3; k++) {
            t.rv[k] = (float)a->rv[3 * i + k];
            t.rv2[k] = (float)a->rv2[3 * i + k];
        }
        int cc[3];
        bool in = true;
        const float pos[3] = {t.x, t.y, t.z};
        for (int d = 0; d < 3; d++) {
            cc[d] = (int)std::floor((pos[d] - corg[d]) / kCellA);
            in = in && cc[d] >= 0 && cc[d] < nc[d];
        }
        if (in) {
            cell_of[i] = cc[0] + nc[0] * (cc[1] + nc[1] * cc[2]);
            count[cell_of[i] + 1]++;
        }
    }
    for (int k = 0; k < ncells; k++) count[k + 1] += count[k];
    std::vector<AAtom> sorted(count[ncells]);
    {
        std::vector<int> cur(count.begin(), count.end() - 1);
        for (int i = 0; i < N; i++)   // atom index order within each cell
            if (cell_of[i] >= 0) sorted[cur[cell_of[i]]++] = by_index[i];
    }
    std::vector<float> rvf(3 * (size_t)std::max(N, 1), 0.f);
    for (size_t k = 0; k < 3 * (size_t)N; k++) rvf[k] = (float)a->rv[k];
    // pair potentials of the affinity maps, effective eps
    std::vector<APair> pairs((size_t)std::max(a->naff, 1) * std::max(T, 1));
    for (int k = 0; k < a->naff; k++)
        for (int t = 0; t < T; t++) {
            const int m = a->aff_slot[k];
            const ag_pair &pf = a->pair_form[(size_t)m * T + t];
            APair &pp = pairs[(size_t)k * T + t];
            memset(&pp, 0, sizeof pp);
            pp.a = pf.xA;
            pp.b = pf.xB;
            if (pf.xA == pf.xB || pf.xA <= 0 || pf.xB <= 0)
                ag::fail(AG_ERR_BAD_INPUT, "pair exponents must differ and be positive");
            if (pf.cA == 0 && pf.cB == 0) {
                pp.R = (float)a->rij[(size_t)m * T + t];
                pp.eps = (float)(a->eps0[(size_t)m * T + t] * (pf.hbonder ? a->w[1] : a->w[0]));
                pp.rstar = pp.R;
                pp.expl = 0;
            } else {
                pp.cA = (float)pf.cA;
                pp.cB = (float)pf.cB;
                pp.rstar = (float)std::pow(pf.xA * pf.cA / (pf.xB * pf.cB), 1.0 / (pf.xA - pf.xB));
                pp.expl = 1;
            }
        }
    std::vector<float4> eat(std::max(N, 1));
    for (int i = 0; i < N; i++)
        eat[i] = make_float4((float)a->coord[3 * i], (float)a->coord[3 * i + 1], (float)a->coord[3 * i + 2],
                             (float)(a->charge[i] * a->w[2]));
    // the forward's atoms: coordinates scaled by cs, weights with the
    // electrostatic constants (k_an_e)
    const double cs = a->dddiel ? (double)kMsLB * 1.4426950408889634 : 1.0;
    const double wk = a->dddiel ? 332.0 * cs : a->inv_diel;
    std::vector<float4> fat(std::max(N, 1));
    for (int i = 0; i < N; i++)
        fat[i] = make_float4((float)(cs * a->coord[3 * i]), (float)(cs * a->coord[3 * i + 1]),
                             (float)(cs * a->coord[3 * i + 2]), (float)(a->charge[i] * a->w[2] * wk));
    // donor H atoms and atoms whose 8 A sphere meets the box (adjoint work lists)
    a->h_atoms.clear();
    a->atom_list.clear();
    for (int i = 0; i < N; i++) {
        bool near = true;
        for (int d = 0; d < 3; d++) {
            const double lo = -g.ne[d] * g.spacing - 8.0, hi = (g.n1[d] - 1 - g.ne[d]) * g.spacing + 8.0;
            near = near && a->coord[3 * i + d] >= lo && a->coord[3 * i + d] <= hi;
        }
        if (!near) continue;
        a->atom_list.push_back(i);
        if (a->hbond[i] == AG_HB_D1 || a->hbond[i] == AG_HB_DS) a->h_atoms.push_back(i);
    }
    auto up = [&](DevBuf &b, const void *src, size_t bytes) {
        void *d = b.get(bytes);
        if (bytes) HIP_CHECK(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, st));
        return d;
    };
    AParams &p = a->ap;
    memset(&p, 0, sizeof p);
    for (int d = 0; d < 3; d++) {
        p.n1[d] = g.n1[d];
        p.ne[d] = g.ne[d];
        p.corg[d] = corg[d];
        p.nc[d] = nc[d];
    }
    p.h = (float)g.spacing;
    p.P = a->P;
    p.cell_start = (const int *)up(a->d_cells, count.data(), sizeof(int) * count.size());
    p.atoms = (const AAtom *)up(a->d_atoms, sorted.data(), sizeof(AAtom) * sorted.size());
    up(a->d_by_index, by_index.data(), sizeof(AAtom) * by_index.size());
    p.rv_by_index = (const float *)up(a->d_rv, rvf.data(), sizeof(float) * rvf.size());
    p.n_atoms = N;
    p.naff = a->naff;
    for (int k = 0; k < a->naff; k++) {
        p.aff_slot[k] = a->aff_slot[k];
        p.des_slot[k] = a->des_slot[k];
        for (int h = 0; h < 7; h++) p.cases[k][h] = a->cases[k][h];
        const ag_map_desc &md = a->maps[a->des_slot[k]];
        p.sp[k] = (float)md.solpar_probe;
        p.volp[k] = (float)md.vol_probe;
    }
    p.n_cov = (int)a->cov_slots.size();
    for (int k = 0; k < p.n_cov; k++) p.cov_slot[k] = a->cov_slots[k];
    for (int d = 0; d < 3; d++) p.cov_pos[d] = a->cov_pos[d];
    p.cov_hw = a->cov_hw;
    p.cov_barrier = a->cov_barrier;
    p.d_slot = a->dsolv;
    p.pairs = (const APair *)up(a->d_pairs, pairs.data(), sizeof(APair) * pairs.size());
    p.T = T;
    p.w_smooth = (float)(std::floor(a->r_smooth * AG_A_DIV / 2.) / AG_A_DIV);
    p.w_desolv = (float)a->w[3];
    p.qasp = (float)a->qasp;
    p.closest_h = (int *)a->d_closest.get(sizeof(int) * (size_t)a->P);
    p.amin = (int *)a->d_amin.get(sizeof(int) * (size_t)a->P * std::max(a->naff, 1));
    p.amax = (int *)a->d_amax.get(sizeof(int) * (size_t)a->P * std::max(a->naff, 1));
    p.vh = (float *)a->d_vh.get(sizeof(float) * 3 * (size_t)a->P);
    EParams &e = a->ep;
    memset(&e, 0, sizeof e);
    for (int d = 0; d < 3; d++) {
        e.n1[d] = g.n1[d];
        e.ne[d] = g.ne[d];
    }
    e.h = (float)g.spacing;
    e.P = a->P;
    e.atoms = (const float4 *)up(a->d_eatoms, eat.data(), sizeof(float4) * eat.size());
    e.n_atoms = N;
    e.dddiel = a->dddiel;
    e.inv_diel = (float)a->inv_diel;
    e.fatoms = (const float4 *)up(a->d_fatoms, fat.data(), sizeof(float4) * fat.size());
    e.sc = (float)(cs * g.spacing);
    e.rcap = (float)(2.0 / cs);
    e.echunk = std::max(kEAtoms, (N + kEMaxChunks - 1) / kEMaxChunks);
    e.nechunks = std::max(1, (N + e.echunk - 1) / e.echunk);
    e.epart = (float *)a->d_epart.get(sizeof(float) * (size_t)e.nechunks * (size_t)a->P);
    // run length: 13 or 16 points, whichever leaves fewer idle run slots
    a->eK = 13 * ((g.n1[0] + 12) / 13) < 16 * ((g.n1[0] + 15) / 16) ? 13 : 16;
    a->nchunks = (int)((a->P + kEAdjChunk - 1) / kEAdjChunk);
    e.nchunks = a->nchunks;
    e.partial = (float4 *)a->d_partial.get(sizeof(float4) * (size_t)a->nchunks * std::max(N, 1));
    up(a->d_hatoms, a->h_atoms.data(), sizeof(int) * a->h_atoms.size());
    up(a->d_alist, a->atom_list.data(), sizeof(int) * a->atom_list.size());
    HIP_CHECK(hipStreamSynchronize(st));   // host staging goes out of scope
    a->uploaded = true;
}

void analytic_forward(ag_analytic *a, float *d_maps) {
    ag_ctx *c = a->ctx;
    HIP_CHECK(hipSetDevice(c->device));
    if (!a->uploaded) analytic_upload(a);
    hipStream_t st = c->stream;
    AParams p = a->ap;
    p.maps = d_maps;
    const unsigned grid = (unsigned)((a->P + kAT - 1) / kAT);
    HIP_CHECK(hipEventRecord(c->ev[0], st));
    k_an_near<<<grid, kAT, 0, st>>>(p);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(c->ev[1], st));
    if (a->elec >= 0) {
        EParams e = a->ep;
        e.emap = d_maps + (size_t)a->elec * a->P;
        const long long nruns = (long long)a->grid.n1[1] * a->grid.n1[2] * ((a->grid.n1[0] + a->eK - 1) / a->eK);
        const dim3 eg((unsigned)((nruns + 63) / 64), (unsigned)e.nechunks);
        if (a->eK == 13) {
            if (a->dddiel) k_an_e<13, true><<<eg, 64, 0, st>>>(e);
            else k_an_e<13, false><<<eg, 64, 0, st>>>(e);
        } else {
            if (a->dddiel) k_an_e<16, true><<<eg, 64, 0, st>>>(e);
            else k_an_e<16, false><<<eg, 64, 0, st>>>(e);
        }
        k_an_esum<<<(unsigned)((a->P + 255) / 256), 256, 0, st>>>(e.epart, e.nechunks, a->P, e.emap);
        HIP_CHECK(hipGetLastError());
    }
    HIP_CHECK(hipEventRecord(c->ev[2], st));
    HIP_CHECK(hipEventRecord(c->ev[3], st));
    HIP_CHECK(hipStreamSynchronize(st));
    HIP_CHECK(hipEventElapsedTime(&c->last_ms[1], c->ev[0], c->ev[1]));
    HIP_CHECK(hipEventElapsedTime(&c->last_ms[2], c->ev[1], c->ev[2]));
    HIP_CHECK(hipEventElapsedTime(&c->last_ms[3], c->ev[0], c->ev[3]));
    c->last_ms[0] = 0.f;
    a->last_maps = d_maps;
}

// Gradients of L with respect to the parameters, given dL/dmaps on the
// device (the maps of the last forward are the linearisation point).
void analytic_adjoint(ag_analytic *a, const float *d_dmaps, double *g_coord, double *g_charge, double *g_rij,
                      double *g_eps0, double g_w[4]) {
    ag_ctx *c = a->ctx;
    HIP_CHECK(hipSetDevice(c->device));
    if (!a->uploaded || !a->last_maps) ag::fail(AG_ERR_BAD_INPUT, "ag_analytic_adjoint needs a forward first");
    hipStream_t st = c->stream;
    const int N = a->N, T = a->T, M = a->M;
    const unsigned grid = (unsigned)((a->P + kAT - 1) / kAT);
    AParams p = a->ap;
    p.dmaps = d_dmaps;
    HIP_CHECK(hipEventRecord(c->ev[0], st));
    const int nh = (int)a->h_atoms.size(), nl = (int)a->atom_list.size();
    float *d_hout = (float *)a->d_hout.get(sizeof(float) * 3 * (size_t)std::max(nh, 1));
    float *d_aout = (float *)a->d_aout.get(sizeof(float) * kAtomOut * (size_t)std::max(nl, 1));
    double *d_eout = (double *)a->d_eout.get(sizeof(double) * 4 * (size_t)std::max(N, 1));
    if (nh > 0) {
        k_an_vh<<<grid, kAT, 0, st>>>(p);
        const int nblk = (int)((a->P + kGPack - 1) / kGPack);
        int *counts = (int *)a->d_vcount.get(sizeof(int) * (2 * (size_t)nblk + 1));
        int *vlist = (int *)a->d_vpack.get(sizeof(int) * (size_t)a->P);
        pack_points(NonzeroVh{p.vh}, a->P, counts, counts + nblk, vlist, st);
        HsumParams hp{p.closest_h, p.vh, vlist, counts + 2 * nblk, (const int *)a->d_hatoms.ptr, d_hout};
        k_an_hsum<<<nh, kAT, 0, st>>>(hp);
    }
    if (nl > 0) {
        AtomAdjParams q;
        q.a = p;
        q.by_index = (const AAtom *)a->d_by_index.ptr;
        q.atom_list = (const int *)a->d_alist.ptr;
        q.out = d_aout;
        unsigned char *mask = (unsigned char *)a->d_gmask.get((size_t)a->P);
        k_an_gmask<<<grid, kAT, 0, st>>>(p, mask);
        q.mask = mask;
        k_an_atom<<<nl, kAT, 0, st>>>(q);
    }
    HIP_CHECK(hipGetLastError());
    const bool want_e = a->elec >= 0 && N > 0;
    if (want_e) {
        EParams e = a->ep;
        e.ge = d_dmaps + (size_t)a->elec * a->P;
        const int nblk = (int)((a->P + kGPack - 1) / kGPack);
        int *counts = (int *)a->d_gcount.get(sizeof(int) * (2 * (size_t)nblk + 1));
        int *offs = counts + nblk;
        int *packed = (int *)a->d_gpack.get(sizeof(int) * (size_t)a->P);
        pack_points(NonzeroE{e.ge}, a->P, counts, offs, packed, st);
        const dim3 eg((unsigned)((N + 63) / 64), (unsigned)a->nchunks);
        if (a->dddiel) k_an_eadj<true><<<eg, 64, 0, st>>>(e, packed, offs + nblk);
        else k_an_eadj<false><<<eg, 64, 0, st>>>(e, packed, offs + nblk);
        k_an_ered<<<(N + 255) / 256, 256, 0, st>>>(e.partial, N, a->nchunks, d_eout);
        HIP_CHECK(hipGetLastError());
    }
    HIP_CHECK(hipEventRecord(c->ev[3], st));
    std::vector<float> hout(3 * (size_t)nh), aout((size_t)kAtomOut * nl);
    std::vector<double> eout(4 * (size_t)N, 0.0);
    if (nh) HIP_CHECK(hipMemcpyAsync(hout.data(), d_hout, sizeof(float) * hout.size(), hipMemcpyDeviceToHost, st));
    if (nl) HIP_CHECK(hipMemcpyAsync(aout.data(), d_aout, sizeof(float) * aout.size(), hipMemcpyDeviceToHost, st));
    if (want_e) HIP_CHECK(hipMemcpyAsync(eout.data(), d_eout, sizeof(double) * eout.size(), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    HIP_CHECK(hipEventElapsedTime(&c->last_ms[3], c->ev[0], c->ev[3]));
    // assemble in FP64
    std::vector<double> gx(3 * (size_t)N, 0.0), grv(3 * (size_t)N, 0.0), grv2(3 * (size_t)N, 0.0), gq(N, 0.0);
    std::vector<double> gR((size_t)M * T, 0.0), gE((size_t)M * T, 0.0);
    double gwd = 0.0;
    for (int l = 0; l < nl; l++) {
        const int i = a->atom_list[l];
        const float *o = aout.data() + (size_t)l * kAtomOut;
        for (int k = 0; k < 3; k++) {
            gx[3 * i + k] += o[k];
            grv[3 * i + k] += o[3 + k];
            grv2[3 * i + k] += o[6 + k];
        }
        gq[i] += o[9];
        gwd += o[10];
        for (int k = 0; k < a->naff; k++) {
            const size_t pm = (size_t)a->aff_slot[k] * T + a->type[i];
            gR[pm] += o[11 + k];
            gE[pm] += o[11 + kAMaxAff + k];
        }
    }
    for (int l = 0; l < nh; l++) {
        const int h = a->h_atoms[l];
        for (int k = 0; k < 3; k++) grv[3 * h + k] += hout[3 * l + k];
    }
    double gwe = 0.0;
    for (int i = 0; i < N; i++) {
        // eout: d/dx of sum_p g q w f (the float4 weight holds q w), and
        // sum_p g f
        for (int k = 0; k < 3; k++) gx[3 * i + k] += eout[4 * i + k];
        gq[i] += a->w[2] * eout[4 * i + 3];
        gwe += a->charge[i] * eout[4 * i + 3];
    }
    // bond vectors -> coordinates
    ag::bond_vectors_vjp(N, a->coord.data(), a->recipes.data(), grv.data(), grv2.data(), gx.data());
    // effective eps -> unweighted eps and the vdW / H-bond weights
    double gwv = 0.0, gwh = 0.0;
    for (int m = 0; m < M; m++)
        for (int t = 0; t < T; t++) {
            const size_t k = (size_t)m * T + t;
            const ag_pair &pf = a->pair_form[k];
            if (!(pf.cA == 0 && pf.cB == 0)) {
                gR[k] = gE[k] = 0.0;
                continue;
            }
            const double wk = pf.hbonder ? a->w[1] : a->w[0];
            if (pf.hbonder) gwh += gE[k] * a->eps0[k];
            else gwv += gE[k] * a->eps0[k];
            gE[k] *= wk;
        }
    if (g_coord) std::copy(gx.begin(), gx.end(), g_coord);
    if (g_charge) std::copy(gq.begin(), gq.end(), g_charge);
    if (g_rij) std::copy(gR.begin(), gR.end(), g_rij);
    if (g_eps0) std::copy(gE.begin(), gE.end(), g_eps0);
    if (g_w) {
        g_w[0] = gwv;
        g_w[1] = gwh;
        g_w[2] = gwe;
        g_w[3] = gwd;
    }
}

void analytic_info(const ag_analytic *a, int *n_atoms, int *n_maps, int64_t *points, int *n_pairs) {
    if (n_atoms) *n_atoms = a->N;
    if (n_maps) *n_maps = a->n_maps;
    if (points) *points = a->P;
    if (n_pairs) *n_pairs = a->M * a->T;
}

}  // namespace agdev
