// Device side of the MI355X AutoGrid4 engine (gfx950, wave64).
//
// One call of ag_maps_plain runs, on the context's stream:
//   1. cell list   : atoms within 8 A of the box binned into cubic cells by a
//                    counting sort; each cell's atoms are then ordered by atom
//                    index so every later traversal is deterministic.
//   2. near kernel : one workgroup per 8x8x4 brick of grid points.  It stages
//                    the atoms within 8.03 A of the brick (cells in fixed
//                    order, ballot compaction) and evaluates, per point, the
//                    8 A part of the reference's per-point loop
//                    (main.cpp:2866-3431): covalent maps, closest donor H,
//                    directional H-bond factors, vdW/H-bond table terms,
//                    desolvation, the d map, the H-bond min/max fix-up, and
//                    the electrostatic terms of every atom the far kernel
//                    leaves out.  Bin indices floor(128 r) are taken from an
//                    FP32 distance unless 128 r lies within 1e-3 of an
//                    integer, where the reference's FP64 distance is
//                    recomputed, so the table bins match the reference's.
//   3. far kernel  : the electrostatic sum over all atoms (main.cpp:2940-2982)
//                    for every pair whose FP32 bin is >= 1025 (r > 8 A); the
//                    dielectric table sits in LDS with zeros below bin 1025,
//                    atoms stream through scalar registers, each lane owns a
//                    run of 8 points along x, and the partial sums of every
//                    64-atom chunk are added with Kahan compensation onto
//                    the near kernel's value.
// No atomics touch map values; every sum has a fixed order.

#include <hip/hip_runtime.h>
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "ag_internal.h"

#define HIP_CHECK(x)                                                                    \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess)                                                           \
            ag::fail(AG_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_) + \
                                        " at " #x);                                     \
    } while (0)

namespace {

constexpr int kBX = 8, kBY = 8, kBZ = 4;          // brick of the near kernel
constexpr int kNearThreads = kBX * kBY * kBZ;     // 256
constexpr int kMapSlots = 16;                     // maps per near-kernel pass
constexpr int kPlainSlots = 12;                   // of which without H-bond terms
constexpr int kHbSlots = kMapSlots - kPlainSlots; // of which H-bonding probes
constexpr int kCandCap = 2048;                    // candidates per brick
constexpr int kHCap = 256;                        // donor-H candidates per brick
constexpr int kChunk = 128;                       // candidate records staged per pass
constexpr int kNearRepsLen = 1040;                // electrostatic bins used by the near kernel
constexpr float kNearRadius = 8.03f;              // staging radius (A)
constexpr int kFarK = 8;                          // points per lane, far kernel
constexpr int kFarThreads = 512;
constexpr int kFarTabMax = 16384;                 // LDS table entries without the clamp
constexpr int kFarAtomChunk = 64;                 // Kahan block
constexpr int kFarBinMin = 1025;                  // far kernel takes bins >= this
constexpr float kFarBias = 1.0e-3f;               // added to (128 r)^2, keeps rsq finite
constexpr float kGuardBins = 1.0e-3f;             // FP64 recompute band around bin edges

// Map roles.
enum : int { ROLE_NONE = 0, ROLE_AFF = 1, ROLE_DES = 2, ROLE_COV = 4 };
// Coefficient bits of an H-bonding slot for one atom class.
enum : int { HC_PLAIN = 1, HC_HRAMP = 2, HC_MINMAX = 4, HC_RACC = 8 };
// H-bond cases (main.cpp:3286-3390).
enum : int { HBC_PLAIN = 0, HBC_ACC_PROBE = 1, HBC_MINMAX_RACC = 2, HBC_DON_PROBE = 3 };

struct __align__(16) NearAtom {   // 96 bytes
    double x, y, z;               // grid frame, as the reference holds them
    float qe;                     // charge * coeff_estat
    float vol;
    float solq;                   // solpar + solpar_q * |charge|
    float dvol;                   // solpar_q * vol (d map)
    float rv[3];
    int meta;                     // type | hbond << 8 | rexp << 16 | disorder << 24
    float rv2[3];
    int index;                    // atom index in the receptor
    float xE, yE, zE;             // the far kernel's coordinates (128 * coord, float)
    float pad;
};
static_assert(sizeof(NearAtom) == 96, "NearAtom layout");

struct NearParams {
    int n1[3], ne[3];
    double spacing;
    float s128;                   // (float)(spacing * 128)
    int nb[3], nbricks;
    // cell list
    double cell_origin[3];
    double cell_size;
    int nc[3];
    const int *cell_start;        // [ncells + 1]
    const NearAtom *atoms;        // sorted by (cell, index)
    const double *coord64;        // [n][3] by atom index
    const double *rvec64;         // [n][3]
    const double *rvec2_64;       // [n][3]
    // tables
    const float *tab;             // [T][1024][kMapSlots] for this pass's maps
    const float *sol;             // [1025]
    const float *reps;            // [8192] electrostatic factor (r_eps or 332/diel)
    // map slots of this pass: [0, kPlainSlots) maps without H-bond terms,
    // [kPlainSlots, kMapSlots) H-bonding probes
    float kE[kMapSlots];          // 1: add the vdW/H-bond table value; 0: desolvation only or unused
    float sp[kMapSlots];          // solpar_probe, 0 when the slot carries no desolvation
    float volp[kMapSlots];        // vol_probe, 0 when the slot carries no desolvation
    int hcode[kHbSlots][7];       // per atom H-bond class: bits 0-3 ordered atom, 4-7 disordered
                                  // atom; each nibble kP | kH << 1 | kM << 2 | use_racc << 3
    int out_slot[kMapSlots];      // output map of each slot, -1 if unused
    int e_slot, d_slot;           // -1 when not produced in this pass
    int has_affinity;             // any slot in use, or the d map
    int disorder_h, hydrogen_type;
    int *cmap;                    // constriction count per point, or null (pass 0 only)
    double ccut;                  // constriction distance cutoff (<= 8 A)
    float *maps;
    long long P;
    int *status;
};

struct FarParams {
    int n1[3], ne[3];
    float s128;
    int runs_per_row, nruns;
    const float4 *atoms;          // (xE, yE, zE, w = q * coeff_estat * 128)
    int n_atoms;
    const float *tab;             // [tab_len]
    int tab_len;
    float tmax;                   // tab_len - 1
    float *emap;                  // e map: in = near part, out = total
};

__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    // Consecutive workgroup ids go round-robin over the 8 XCDs; give each XCD
    // a contiguous range of tiles so neighbouring tiles share its L2.  XCD x
    // receives q + (x < r) workgroups and owns tiles [x*q + min(x, r), ...).
    const int q = nblocks / 8, r = nblocks % 8;
    const int x = bid % 8, k = bid / 8;
    return x * q + min(x, r) + k;
}

// The far kernel's per-pair bin input, reproduced bit for bit by the near
// kernel when it decides which atoms the far kernel left out.
__device__ __forceinline__ float far_rb(float ax, float ay, float az, float px, float py, float pz) {
#pragma clang fp contract(off)
    const float dy = ay - py, dz = az - pz;
    const float dyz2 = __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, kFarBias));
    const float dx = ax - px;
    const float rb2 = __builtin_fmaf(dx, dx, dyz2);
    return rb2 * __builtin_amdgcn_rsqf(rb2);
}

// ---------------------------------------------------------------------------
// cell list
// ---------------------------------------------------------------------------

struct CellParams {
    double origin[3];
    double cs;
    int nc[3];
    int n_atoms;
};

__device__ __forceinline__ int cell_of(const CellParams &c, const double *p) {
    int ix = (int)floor((p[0] - c.origin[0]) / c.cs);
    int iy = (int)floor((p[1] - c.origin[1]) / c.cs);
    int iz = (int)floor((p[2] - c.origin[2]) / c.cs);
    if (ix < 0 || iy < 0 || iz < 0 || ix >= c.nc[0] || iy >= c.nc[1] || iz >= c.nc[2]) return -1;
    return ix + c.nc[0] * (iy + c.nc[1] * iz);
}

__global__ void k_cell_count(CellParams c, const double *coord, int *atom_cell, int *count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n_atoms) return;
    int cell = cell_of(c, coord + 3 * (size_t)i);
    atom_cell[i] = cell;
    if (cell >= 0) atomicAdd(&count[cell], 1);
}

// Exclusive scan of count[0..n) into start[0..n], one workgroup of 1024.
__global__ void __launch_bounds__(1024) k_scan(const int *count, int *start, int n) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b = t * per, e = min(n, b + per);
    int s = 0;
    for (int i = b; i < e; i++) s += count[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int i = b; i < e; i++) {
        start[i] = run;
        run += count[i];
    }
    if (t == 1023) start[n] = part[1023];
}

__global__ void k_cell_scatter(int n, const int *atom_cell, const int *start, int *cursor, int *slot_atom) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cell = atom_cell[i];
    if (cell < 0) return;
    int pos = start[cell] + atomicAdd(&cursor[cell], 1);
    slot_atom[pos] = i;
}

// Order each cell's atoms by atom index (insertion sort; cells hold few atoms).
__global__ void k_cell_sort(int ncells, const int *start, int *slot_atom) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    int b = start[c], e = start[c + 1];
    for (int i = b + 1; i < e; i++) {
        int v = slot_atom[i];
        int j = i - 1;
        while (j >= b && slot_atom[j] > v) {
            slot_atom[j + 1] = slot_atom[j];
            j--;
        }
        slot_atom[j + 1] = v;
    }
}

struct AtomSrc {
    const double *coord, *charge, *vol, *solpar, *rvec, *rvec2;
    const int *type, *hbond, *rexp;
    const unsigned char *disorder;
    double coeff_estat, solpar_q;
};

// nslots = start[ncells], the number of atoms that landed in a cell.
__global__ void k_gather(const int *nslots, const int *slot_atom, AtomSrc s, NearAtom *out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *nslots) return;
    int i = slot_atom[k];
    NearAtom a;
    a.x = s.coord[3 * i];
    a.y = s.coord[3 * i + 1];
    a.z = s.coord[3 * i + 2];
    double q = s.charge[i];
    a.qe = (float)(q * s.coeff_estat);
    a.vol = (float)s.vol[i];
    a.solq = (float)(s.solpar[i] + s.solpar_q * fabs(q));
    a.dvol = (float)(s.solpar_q * s.vol[i]);
    for (int d = 0; d < 3; d++) {
        a.rv[d] = (float)s.rvec[3 * i + d];
        a.rv2[d] = (float)s.rvec2[3 * i + d];
    }
    a.meta = (s.type[i] & 0xff) | ((s.hbond[i] & 0xff) << 8) | ((s.rexp[i] & 0xff) << 16) |
             ((s.disorder[i] ? 1 : 0) << 24);
    a.index = i;
    a.xE = (float)(a.x * 128.0);
    a.yE = (float)(a.y * 128.0);
    a.zE = (float)(a.z * 128.0);
    a.pad = 0.f;
    out[k] = a;
}

__global__ void k_far_atoms(int n, const double *coord, const double *charge, double coeff_estat,
                            float4 *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = make_float4((float)(coord[3 * i] * 128.0), (float)(coord[3 * i + 1] * 128.0),
                         (float)(coord[3 * i + 2] * 128.0), (float)(charge[i] * coeff_estat * 128.0));
}

// ---------------------------------------------------------------------------
// near kernel
// ---------------------------------------------------------------------------

// Block-wide ordered compaction helper: returns this thread's output position
// among flagged threads and the block total (256 threads = 4 waves).
__device__ __forceinline__ int block_compact(bool flag, int *wave_tot, int &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long m = __ballot(flag);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wave_tot[wave] = __popcll(m);
    __syncthreads();
    int base = 0;
    total = 0;
    for (int w = 0; w < kNearThreads / 64; w++) {
        if (w < wave) base += wave_tot[w];
        total += wave_tot[w];
    }
    __syncthreads();
    return base + before;
}

__device__ __forceinline__ double ref_dist(double ax, double ay, double az, double cx, double cy, double cz,
                                           double &dx, double &dy, double &dz) {
#pragma clang fp contract(off)
    // hypotenuse(): sqrt(dx*dx + dy*dy + dz*dz), no fused operations
    dx = ax - cx;
    dy = ay - cy;
    dz = az - cz;
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// Uniform per-atom values: every lane reads the same LDS record, so its
// fields are moved to scalar registers and branches on them stay scalar.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float unif(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

template <int NPLAIN>
__global__ void __launch_bounds__(kNearThreads) k_near(NearParams p) {
    __shared__ int cand[kCandCap];
    __shared__ int hcand[kHCap];
    __shared__ NearAtom rec[kChunk];
    __shared__ double hpos[kHCap][3];
    __shared__ int hidx[kHCap];
    __shared__ float hrvs[kHCap][3];
    __shared__ float reps_s[kNearRepsLen];
    __shared__ float sol_s[AG_NEINT + 1];
    __shared__ int wave_tot[kNearThreads / 64];
    __shared__ int cell_off[kNearThreads + 1];
    __shared__ int cstart_s[kNearThreads];
    __shared__ int ncand_s, nh_s;

    const int tid = threadIdx.x;
    const int brick = xcd_remap(blockIdx.x, gridDim.x);
    const int bx = brick % p.nb[0];
    const int by = (brick / p.nb[0]) % p.nb[1];
    const int bz = brick / (p.nb[0] * p.nb[1]);
    const int ix = bx * kBX + (tid % kBX);
    const int iy = by * kBY + ((tid / kBX) % kBY);
    const int iz = bz * kBZ + tid / (kBX * kBY);
    const bool active = ix < p.n1[0] && iy < p.n1[1] && iz < p.n1[2];

    for (int k = tid; k < kNearRepsLen; k += kNearThreads) reps_s[k] = p.reps[k];
    for (int k = tid; k <= AG_NEINT; k += kNearThreads) sol_s[k] = p.sol[k];

    // brick box in the grid frame
    const int ix0 = bx * kBX, iy0 = by * kBY, iz0 = bz * kBZ;
    const int ix1 = min(ix0 + kBX, p.n1[0]) - 1, iy1 = min(iy0 + kBY, p.n1[1]) - 1,
              iz1 = min(iz0 + kBZ, p.n1[2]) - 1;
    const double bmin[3] = {(ix0 - p.ne[0]) * p.spacing, (iy0 - p.ne[1]) * p.spacing, (iz0 - p.ne[2]) * p.spacing};
    const double bmax[3] = {(ix1 - p.ne[0]) * p.spacing, (iy1 - p.ne[1]) * p.spacing, (iz1 - p.ne[2]) * p.spacing};

    // point, reference arithmetic: c = (double)(i - ne) * spacing
    const double cx = (double)(ix - p.ne[0]) * p.spacing;
    const double cy = (double)(iy - p.ne[1]) * p.spacing;
    const double cz = (double)(iz - p.ne[2]) * p.spacing;
    // the far kernel's point coordinates
    const float pxE = (float)(ix - p.ne[0]) * p.s128;
    const float pyE = (float)(iy - p.ne[1]) * p.s128;
    const float pzE = (float)(iz - p.ne[2]) * p.s128;

    // ---- stage candidate atoms: cells overlapping the brick box + radius ----
    int c0[3], c1[3];
    for (int d = 0; d < 3; d++) {
        c0[d] = (int)floor((bmin[d] - kNearRadius - p.cell_origin[d]) / p.cell_size);
        c1[d] = (int)floor((bmax[d] + kNearRadius - p.cell_origin[d]) / p.cell_size);
        c0[d] = max(c0[d], 0);
        c1[d] = min(c1[d], p.nc[d] - 1);
    }
    const int ncx = c1[0] - c0[0] + 1, ncy = c1[1] - c0[1] + 1, ncz = c1[2] - c0[2] + 1;
    const int ncell = (ncx > 0 && ncy > 0 && ncz > 0) ? ncx * ncy * ncz : 0;
    if (tid == 0) { ncand_s = 0; nh_s = 0; }
    __syncthreads();
    const float r2cut = kNearRadius * kNearRadius;
    for (int cbase = 0; cbase < ncell; cbase += kNearThreads) {
        const int ci = cbase + tid;
        int cstart = 0, ccount = 0;
        if (ci < ncell) {
            const int gx = c0[0] + ci % ncx, gy = c0[1] + (ci / ncx) % ncy, gz = c0[2] + ci / (ncx * ncy);
            const int g = gx + p.nc[0] * (gy + p.nc[1] * gz);
            cstart = p.cell_start[g];
            ccount = p.cell_start[g + 1] - cstart;
        }
        cell_off[tid] = ccount;
        __syncthreads();
        for (int off = 1; off < kNearThreads; off <<= 1) {
            int v = tid >= off ? cell_off[tid - off] : 0;
            __syncthreads();
            cell_off[tid] += v;
            __syncthreads();
        }
        const int total = cell_off[kNearThreads - 1];
        const int my_end = cell_off[tid];
        __syncthreads();
        cell_off[tid + 1] = my_end;
        if (tid == 0) cell_off[0] = 0;
        cstart_s[tid] = cstart;
        __syncthreads();
        for (int j0 = 0; j0 < total; j0 += kNearThreads) {
            const int j = j0 + tid;
            bool ok = false, is_h = false;
            int slot = -1;
            if (j < total) {
                int lo = 0, hi = kNearThreads;  // last cell with cell_off[c] <= j
                while (hi - lo > 1) {
                    int mid = (lo + hi) >> 1;
                    if (cell_off[mid] <= j) lo = mid; else hi = mid;
                }
                slot = cstart_s[lo] + (j - cell_off[lo]);
                const NearAtom &a = p.atoms[slot];
                float dd = 0.f;
                const double av[3] = {a.x, a.y, a.z};
                for (int d = 0; d < 3; d++) {
                    double t = av[d] < bmin[d] ? bmin[d] - av[d] : (av[d] > bmax[d] ? av[d] - bmax[d] : 0.0);
                    dd += (float)(t * t);
                }
                ok = dd <= r2cut;
                const int hb = (a.meta >> 8) & 0xff;
                is_h = ok && (hb == AG_HB_DS || hb == AG_HB_D1);
            }
            int tot;
            const int pos = block_compact(ok, wave_tot, tot);
            const int base = ncand_s;
            if (ok && base + pos < kCandCap) cand[base + pos] = slot;
            int htot;
            const int hpos_i = block_compact(is_h, wave_tot, htot);
            const int hbase = nh_s;
            if (is_h && hbase + hpos_i < kHCap) hcand[hbase + hpos_i] = slot;
            __syncthreads();
            if (tid == 0) { ncand_s = base + tot; nh_s = hbase + htot; }
            __syncthreads();
        }
    }
    const int ncand = ncand_s, nh = nh_s;
    if (ncand > kCandCap || nh > kHCap) {
        if (tid == 0) atomicOr(p.status, 1);
        return;
    }
    // donor-H candidates into LDS for the closest-H search
    for (int h = tid; h < nh; h += kNearThreads) {
        const NearAtom &a = p.atoms[hcand[h]];
        hpos[h][0] = a.x; hpos[h][1] = a.y; hpos[h][2] = a.z;
        hidx[h] = a.index;
        hrvs[h][0] = a.rv[0]; hrvs[h][1] = a.rv[1]; hrvs[h][2] = a.rv[2];
    }
    __syncthreads();

    float acc[kMapSlots];
#pragma unroll
    for (int m = 0; m < kMapSlots; m++) acc[m] = 0.f;
    // min/max of the H-bond pairs of each H-bonding slot (main.cpp:3300-3374)
    float hmin[kHbSlots], hmax[kHbSlots];
#pragma unroll
    for (int k = 0; k < kHbSlots; k++) {
        hmin[k] = 999999.f;
        hmax[k] = -999999.f;
    }

    // ---- closest donor H (main.cpp:2985-3015), FP64, lowest index on ties ----
    double best_r = 999999.0;
    int best_idx = 0x7fffffff, best_h = -1;
    for (int h = 0; h < nh; h++) {
        double dx, dy, dz;
        const double r = ref_dist(hpos[h][0], hpos[h][1], hpos[h][2], cx, cy, cz, dx, dy, dz);
        const int ix_h = hidx[h];
        if (r < best_r || (r == best_r && ix_h < best_idx)) {
            best_r = r;
            best_idx = ix_h;
            best_h = h;
        }
    }
    float hrv[3] = {0.f, 0.f, 0.f};
    if (best_h >= 0) {
        hrv[0] = hrvs[best_h][0]; hrv[1] = hrvs[best_h][1]; hrv[2] = hrvs[best_h][2];
    }

    float e_near = 0.f, d_acc = 0.f;
    int ccount = 0;
    const float *tab = p.tab;
    const bool want_e = p.e_slot >= 0;
    // cos and sin of 1.24791 rad, the disordered-hydroxyl offset angle
    const float cdis = 0.31936680f, sdis = 0.94763168f;

    for (int cbase = 0; cbase < ncand; cbase += kChunk) {
        const int nc = min(kChunk, ncand - cbase);
        __syncthreads();
        for (int j = tid; j < nc; j += kNearThreads) rec[j] = p.atoms[cand[cbase + j]];
        __syncthreads();
        for (int j = 0; j < nc; j++) {
            const NearAtom &a = rec[j];
            // geometry: FP64 difference as the reference forms it, FP32 after
            const double dx64 = a.x - cx, dy64 = a.y - cy, dz64 = a.z - cz;
            const float dxf = (float)dx64, dyf = (float)dy64, dzf = (float)dz64;
            const float r2 = __builtin_fmaf(dxf, dxf, __builtin_fmaf(dyf, dyf, dzf * dzf));
            if (r2 > 64.33f) continue;  // beyond 8.02 A: neither cutoff pass nor near e term
            // hardware rsq: 128 r is good to ~3e-4 bins here, inside the guard band
            const float r2c = fmaxf(r2, 1.0e-12f);
            float r = r2c * __builtin_amdgcn_rsqf(r2c);
            const float g = r * 128.f;
            int bin;
            bool cut;
            if (fabsf(g - rintf(g)) < kGuardBins) {
                double rdx, rdy, rdz;
                const double r64 = ref_dist(a.x, a.y, a.z, cx, cy, cz, rdx, rdy, rdz);
                bin = (int)(r64 * 128.0);
                cut = r64 > 8.0;
                r = (float)r64;
            } else {
                bin = (int)g;
                cut = r > 8.0f;
            }
            // constriction grid: atoms within the cutoff (main.cpp:2969-2973)
            if (p.cmap) {
                if (fabsf(r - (float)p.ccut) < 1.0e-5f) {
                    double rdx, rdy, rdz;
                    ccount += ref_dist(a.x, a.y, a.z, cx, cy, cz, rdx, rdy, rdz) <= p.ccut;
                } else {
                    ccount += r <= (float)p.ccut;
                }
            }
            // electrostatics of pairs the far kernel leaves out
            if (want_e) {
                bool near_e;
                if (g < 1022.f) near_e = true;
                else if (g > 1027.f) near_e = false;
                else near_e = far_rb(a.xE, a.yE, a.zE, pxE, pyE, pzE) < (float)kFarBinMin;
                if (near_e) {
                    // 1/max(r, 0.5) with r clamped at APPROX_ZERO
                    const float inv_rmax = fminf(__builtin_amdgcn_rcpf(fmaxf(r, 1.0e-6f)), 2.f);
                    e_near += unif(a.qe) * inv_rmax * reps_s[min(bin, kNearRepsLen - 1)];
                }
            }
            if (cut || !p.has_affinity) continue;
            const int meta = uni(a.meta);
            const int type = meta & 0xff, hb = (meta >> 8) & 0xff, rexp = (meta >> 16) & 0xff;
            const bool dis = (meta >> 24) & 1;
            if (p.disorder_h && dis && type == p.hydrogen_type) continue;
            const float inv_r = __builtin_amdgcn_rcpf(fmaxf(r, 1.0e-6f));
            const float dn0 = dxf * inv_r, dn1 = dyf * inv_r, dn2 = dzf * inv_r;
            const float rv0 = unif(a.rv[0]), rv1 = unif(a.rv[1]), rv2 = unif(a.rv[2]);
            const int aidx = uni(a.index);
            float racc = 1.f, rdon = 1.f, Hramp = 1.f;
            if (hb == AG_HB_D1 || (hb == AG_HB_A2 && !dis)) {
                // cos(theta) of the D1 / ordered-A2 branches; FP64 near 0
                float cos_t = -(dn0 * rv0) - dn1 * rv1 - dn2 * rv2;
                bool pos;
                if (fabsf(cos_t) < 1.0e-5f) {
                    double rdx, rdy, rdz;
                    double r64 = ref_dist(a.x, a.y, a.z, cx, cy, cz, rdx, rdy, rdz);
                    if (r64 < 1.0e-6) r64 = 1.0e-6;
                    const double ir = 1. / r64;
                    const double *rv64 = p.rvec64 + 3 * (size_t)aidx;
                    double c = 0.;
                    c -= rdx * ir * rv64[0];
                    c -= rdy * ir * rv64[1];
                    c -= rdz * ir * rv64[2];
                    pos = hb == AG_HB_D1 ? c > 0. : c >= 0.;
                    cos_t = (float)c;
                } else {
                    pos = hb == AG_HB_D1 ? cos_t > 0.f : cos_t >= 0.f;
                }
                if (hb == AG_HB_D1) {
                    if (!pos) {
                        racc = 0.f;
                    } else {
                        const float c2 = cos_t * cos_t;
                        racc = rexp == 2 ? c2 : (rexp == 4 ? c2 * c2 : cos_t);
                        if (aidx != best_idx) {
                            float ct = hrv[0] * rv0 + hrv[1] * rv1 + hrv[2] * rv2;
                            ct = fminf(fmaxf(ct, -1.f), 1.f);
                            Hramp = 0.5f - 0.5f * cosf(acosf(ct) * (120.f / 90.f));
                        }
                    }
                } else {
                    // Goodford lone-pair terms.  With x = dn.rv2 and
                    // y = (dn x rv2).rv / |dn x rv2|:
                    //   t0 = pi/2 - acos(x)         -> cos(t0) = sqrt(1 - x^2)
                    //   ti = |acos(y) - pi/2|       -> sin(2 ti) = 2|y| sqrt(1 - y^2)
                    const float rw0 = unif(a.rv2[0]), rw1 = unif(a.rv2[1]), rw2 = unif(a.rv2[2]);
                    float t0 = dn0 * rw0 + dn1 * rw1 + dn2 * rw2;
                    t0 = fminf(fmaxf(t0, -1.f), 1.f);
                    const float cos_t0 = __builtin_amdgcn_sqrtf(fmaxf(1.f - t0 * t0, 0.f));
                    rdon = 0.f;
                    if (pos) {
                        const float cr0 = dn1 * rw2 - dn2 * rw1;
                        const float cr1 = dn2 * rw0 - dn0 * rw2;
                        const float cr2 = dn0 * rw1 - dn1 * rw0;
                        float rd2 = cr0 * cr0 + cr1 * cr1 + cr2 * cr2;
                        if (rd2 < 1.0e-6f) rd2 = 1.0e-6f;
                        const float inv_rd = __builtin_amdgcn_rsqf(rd2);
                        float ti = (cr0 * inv_rd) * rv0 + (cr1 * inv_rd) * rv1 + (cr2 * inv_rd) * rv2;
                        ti = fminf(fmaxf(ti, -1.f), 1.f);
                        const float sin2ti = 2.f * fabsf(ti) * __builtin_amdgcn_sqrtf(fmaxf(1.f - ti * ti, 0.f));
                        rdon = (0.9f + 0.1f * sin2ti) * cos_t0;
                    } else if (cos_t >= -0.34202f) {
                        const float u = 0.116978f - cos_t * cos_t;
                        rdon = 562.25f * (u * u * u) * cos_t0;
                    }
                }
            } else if (hb == AG_HB_A1) {
                const float cos_t = -(dn0 * rv0) - dn1 * rv1 - dn2 * rv2;
                rdon = cos_t <= 0.f ? 0.f : cos_t * cos_t;
            } else if (hb == AG_HB_A2 && dis) {
                // cos(theta - 1.24791) = c cos(1.24791) + sqrt(1 - c^2) sin(1.24791);
                // theta <= 1.24791 + pi/2  <=>  c >= -sin(1.24791)
                float c = -(dn0 * rv0) - dn1 * rv1 - dn2 * rv2;
                c = fminf(fmaxf(c, -1.f), 1.f);
                racc = 0.f;
                rdon = 0.f;
                if (c >= -sdis) {
                    const float cc = c * cdis + __builtin_amdgcn_sqrtf(fmaxf(1.f - c * c, 0.f)) * sdis;
                    rdon = (cc * cc) * (cc * cc);
                    racc = rdon;
                }
            }
            const int indx_n = min(bin, AG_NEINT - 1);
            const int indx_r = min(bin, AG_NEINT);  // r <= 8 here
            const float S = sol_s[indx_r];
            const float avol = unif(a.vol), asolq = unif(a.solq);
            // vdW/H-bond table values of this pair for the pass's slots
            const float *row = tab + ((size_t)type * AG_NEINT + indx_n) * kMapSlots;
            float E[kMapSlots];
#pragma unroll
            for (int q4 = 0; q4 < kMapSlots / 4; q4++) {
                if (q4 < NPLAIN / 4 || q4 == kMapSlots / 4 - 1) {
                    const float4 v = reinterpret_cast<const float4 *>(row)[q4];
                    E[4 * q4] = v.x; E[4 * q4 + 1] = v.y; E[4 * q4 + 2] = v.z; E[4 * q4 + 3] = v.w;
                } else {
                    E[4 * q4] = E[4 * q4 + 1] = E[4 * q4 + 2] = E[4 * q4 + 3] = 0.f;
                }
            }
            // plain slots: table value + desolvation (main.cpp:3375-3390)
#pragma unroll
            for (int m = 0; m < NPLAIN; m++) {
                const float t = __builtin_fmaf(p.sp[m], avol, p.volp[m] * asolq);
                acc[m] = __builtin_fmaf(t, S, __builtin_fmaf(p.kE[m], E[m], acc[m]));
            }
            // H-bonding slots (main.cpp:3286-3374): the (probe, atom) case is a
            // set of 0/1 coefficients chosen per atom class, applied without
            // branches.  A disordered partner uses the hydrogen-type row at
            // r - 1.10 A (main.cpp:3313-3349), FP64 like the reference.
            float Eh[kHbSlots] = {0.f, 0.f, 0.f, 0.f};
            if (dis && p.hydrogen_type >= 0) {
                double rdx, rdy, rdz;
                const double r64 = ref_dist(a.x, a.y, a.z, cx, cy, cz, rdx, rdy, rdz);
                const int hb_bin = min(max(0, (int)((r64 - 1.10) * 128.0)), AG_NEINT - 1);
                const float4 v = reinterpret_cast<const float4 *>(
                    tab + ((size_t)p.hydrogen_type * AG_NEINT + hb_bin) * kMapSlots)[kPlainSlots / 4];
                Eh[0] = v.x; Eh[1] = v.y; Eh[2] = v.z; Eh[3] = v.w;
            }
            const float dfac = racc - rdon;
            const int dsh = dis ? 4 : 0;
#pragma unroll
            for (int k = 0; k < kHbSlots; k++) {
                const int m = kPlainSlots + k;
                const int code = (p.hcode[k][hb] >> dsh) & 15;
                const float kP = (code & HC_PLAIN) ? 1.f : 0.f;
                const float kH = (code & HC_HRAMP) ? 1.f : 0.f;
                const float kF = (code & HC_RACC) ? 1.f : 0.f;
                const bool kM = (code & HC_MINMAX) != 0;
                const float e = E[m];
                const float rsph = __builtin_amdgcn_fmed3f(e * 0.01f, 0.f, 1.f);
                const float f = __builtin_fmaf(kF, dfac, rdon);
                const float mix = __builtin_fmaf(1.f - f, rsph, f);
                const float eh = dis ? Eh[k] : e;
                acc[m] = __builtin_fmaf(kP, e, acc[m]);
                acc[m] = __builtin_fmaf(kH, (eh * Hramp) * mix, acc[m]);
                const float v = e * mix;
                hmin[k] = kM ? fminf(hmin[k], v) : hmin[k];
                hmax[k] = kM ? fmaxf(hmax[k], v) : hmax[k];
                const float t = __builtin_fmaf(p.sp[m], avol, p.volp[m] * asolq);
                acc[m] = __builtin_fmaf(t, S, acc[m]);
            }
            if (p.d_slot >= 0) d_acc += unif(a.dvol) * S;
        }
    }

    if (!active) return;
    const long long pt = (long long)ix + (long long)p.n1[0] * ((long long)iy + (long long)p.n1[1] * iz);
#pragma unroll
    for (int m = 0; m < kMapSlots; m++) {
        if ((m >= NPLAIN && m < kPlainSlots) || p.out_slot[m] < 0) continue;
        float v = acc[m];
        if (m >= kPlainSlots && hmin[m - kPlainSlots] < 999999.f)
            v += hmin[m - kPlainSlots] + fmaxf(hmax[m - kPlainSlots], 0.f);
        p.maps[(long long)p.out_slot[m] * p.P + pt] = v;
    }
    if (want_e) p.maps[(long long)p.e_slot * p.P + pt] = e_near;
    if (p.cmap) p.cmap[pt] = ccount;
    if (p.d_slot >= 0) p.maps[(long long)p.d_slot * p.P + pt] = d_acc;
}

__global__ void k_fill(float *out, long long n, float v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = v;
}

// Covalent maps (main.cpp:2866-2879): a Gaussian well around the
// attachment point, no receptor atoms involved.
struct CovParams {
    int n1[3], ne[3];
    double spacing;
    double pos[3], half_width, barrier;
    int n_cov;
    int slot[16];
    float *maps;
    long long P;
};

__global__ void k_covalent(CovParams c) {
    const long long pt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pt >= c.P) return;
    const int ix = (int)(pt % c.n1[0]);
    const int iy = (int)((pt / c.n1[0]) % c.n1[1]);
    const int iz = (int)(pt / ((long long)c.n1[0] * c.n1[1]));
    const double dx = c.pos[0] - (double)(ix - c.ne[0]) * c.spacing;
    const double dy = c.pos[1] - (double)(iy - c.ne[1]) * c.spacing;
    const double dz = c.pos[2] - (double)(iz - c.ne[2]) * c.spacing;
    double rcov = sqrt(dx * dx + dy * dy + dz * dz) / c.half_width;
    if (rcov < 1.0e-6) rcov = 1.0e-6;
    const float v = (float)(c.barrier * (1. - exp(-0.69314718055994530942 * rcov * rcov)));
    for (int k = 0; k < c.n_cov; k++) c.maps[(long long)c.slot[k] * c.P + pt] = v;
}

// Floating grid (main.cpp:2962-2966): the distance to the nearest receptor
// atom, each distance clamped at APPROX_ZERO.  One lane owns kFarK points
// along x; atoms stream through scalar registers.
struct FminParams {
    int n1[3], ne[3];
    float spacing;
    int runs_per_row, nruns;
    const float4 *atoms;          // (x, y, z, 0) grid frame, A
    int n_atoms;
    float *fmap;
};

__global__ void __launch_bounds__(256) k_fmin(FminParams p) {
    const int run = blockIdx.x * 256 + threadIdx.x;
    if (run >= p.nruns) return;
    const int row = run / p.runs_per_row;
    const int ix0 = (run % p.runs_per_row) * kFarK;
    const int iy = row % p.n1[1];
    const int iz = row / p.n1[1];
    const float py = (float)(iy - p.ne[1]) * p.spacing;
    const float pz = (float)(iz - p.ne[2]) * p.spacing;
    float px[kFarK], m2[kFarK];
#pragma unroll
    for (int k = 0; k < kFarK; k++) {
        px[k] = (float)(ix0 + k - p.ne[0]) * p.spacing;
        m2[k] = 3.0e38f;
    }
    const float4 *__restrict__ atoms = p.atoms;
#pragma unroll 2
    for (int a = 0; a < p.n_atoms; a++) {
        const float4 at = atoms[a];
        const float dy = at.y - py, dz = at.z - pz;
        const float dyz2 = __builtin_fmaf(dy, dy, dz * dz);
#pragma unroll
        for (int k = 0; k < kFarK; k++) {
            const float dx = at.x - px[k];
            m2[k] = fminf(m2[k], __builtin_fmaf(dx, dx, dyz2));
        }
    }
    const long long base = (long long)ix0 + (long long)p.n1[0] * ((long long)iy + (long long)p.n1[1] * iz);
#pragma unroll
    for (int k = 0; k < kFarK; k++)
        if (ix0 + k < p.n1[0])
            p.fmap[base + k] = p.n_atoms > 0 ? fmaxf(__builtin_amdgcn_sqrtf(m2[k]), 1.0e-6f) : 1.0e12f;
}

__global__ void k_fmin_atoms(int n, const double *coord, float4 *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = make_float4((float)coord[3 * i], (float)coord[3 * i + 1], (float)coord[3 * i + 2], 0.f);
}

// Bin audit (diagnostic, not part of a map set): for every (point, atom)
// pair compare the bin floor(128 r) the device path uses with the one the
// reference's FP64 distance gives.
//   far pairs  (the far kernel's FP32 bin >= 1025): the FP32 bin is used
//              as is, so every mismatch is a table-bin flip;
//   near pairs: the near kernel recomputes in FP64 whenever 128 r is within
//              kGuardBins of an integer, so a mismatch outside that band
//              would be an unguarded flip (expected: none).
struct AuditParams {
    int n1[3], ne[3];
    double spacing;
    float s128;
    const float4 *far_atoms;     // the far kernel's atoms
    const double *coord;         // [n][3]
    int n_atoms;
    long long P;
    unsigned long long *out;     // [0] far pairs, [1] far flips, [2] near unguarded flips, [3] max flips/point
};

__global__ void __launch_bounds__(256) k_audit(AuditParams p) {
    const long long pt = (long long)blockIdx.x * 256 + threadIdx.x;
    unsigned long long far_pairs = 0, far_flips = 0, near_bad = 0;
    if (pt < p.P) {
        const int ix = (int)(pt % p.n1[0]);
        const int iy = (int)((pt / p.n1[0]) % p.n1[1]);
        const int iz = (int)(pt / ((long long)p.n1[0] * p.n1[1]));
        const double cx = (double)(ix - p.ne[0]) * p.spacing, cy = (double)(iy - p.ne[1]) * p.spacing,
                     cz = (double)(iz - p.ne[2]) * p.spacing;
        const float pxE = (float)(ix - p.ne[0]) * p.s128, pyE = (float)(iy - p.ne[1]) * p.s128,
                    pzE = (float)(iz - p.ne[2]) * p.s128;
        for (int a = 0; a < p.n_atoms; a++) {
            const float4 at = p.far_atoms[a];
            const float rb = far_rb(at.x, at.y, at.z, pxE, pyE, pzE);
            double dx, dy, dz;
            const double r64 = ref_dist(p.coord[3 * a], p.coord[3 * a + 1], p.coord[3 * a + 2], cx, cy, cz, dx, dy, dz);
            const int b64 = min((int)(r64 * 128.0), AG_NDIEL - 1);
            if (rb >= (float)kFarBinMin) {
                far_pairs++;
                far_flips += min((int)rb, AG_NDIEL - 1) != b64;
            } else {
                const float dxf = (float)dx, dyf = (float)dy, dzf = (float)dz;
                const float r2 = fmaxf(__builtin_fmaf(dxf, dxf, __builtin_fmaf(dyf, dyf, dzf * dzf)), 1.0e-12f);
                const float g = r2 * __builtin_amdgcn_rsqf(r2) * 128.f;
                if (fabsf(g - rintf(g)) >= kGuardBins && min((int)g, AG_NDIEL - 1) != b64) near_bad++;
            }
        }
    }
    // wave sums, then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        far_pairs += __shfl_xor(far_pairs, off);
        near_bad += __shfl_xor(near_bad, off);
    }
    unsigned long long mx = far_flips;
    unsigned long long sum_f = far_flips;
    for (int off = 32; off > 0; off >>= 1) {
        sum_f += __shfl_xor(sum_f, off);
        const unsigned long long o = __shfl_xor(mx, off);
        mx = o > mx ? o : mx;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&p.out[0], far_pairs);
        atomicAdd(&p.out[1], sum_f);
        atomicAdd(&p.out[2], near_bad);
        atomicMax(&p.out[3], mx);
    }
}

// ---------------------------------------------------------------------------
// far kernel
// ---------------------------------------------------------------------------

// CLAMP: the table stops at bin 8191 and indices are clamped as the
// reference clamps them (main.cpp:2956); without it the table is extended
// to the largest bin any pair can reach, every entry past 8191 holding the
// reference's last value, and the clamp instruction disappears.
template <bool CLAMP>
__global__ void __launch_bounds__(kFarThreads) k_far(FarParams p) {
    extern __shared__ float T[];
    for (int i = threadIdx.x; i < p.tab_len; i += kFarThreads) T[i] = p.tab[i];
    __syncthreads();
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int run = blk * kFarThreads + threadIdx.x;
    if (run >= p.nruns) return;
    const int row = run / p.runs_per_row;
    const int ix0 = (run % p.runs_per_row) * kFarK;
    const int iy = row % p.n1[1];
    const int iz = row / p.n1[1];
    const float py = (float)(iy - p.ne[1]) * p.s128;
    const float pz = (float)(iz - p.ne[2]) * p.s128;
    float px[kFarK], s[kFarK], c[kFarK];
    const long long base = (long long)ix0 + (long long)p.n1[0] * ((long long)iy + (long long)p.n1[1] * iz);
#pragma unroll
    for (int k = 0; k < kFarK; k++) {
        px[k] = (float)(ix0 + k - p.ne[0]) * p.s128;
        s[k] = (ix0 + k < p.n1[0]) ? p.emap[base + k] : 0.f;
        c[k] = 0.f;
    }
    const float4 *__restrict__ atoms = p.atoms;
    for (int a0 = 0; a0 < p.n_atoms; a0 += kFarAtomChunk) {
        const int a1 = min(a0 + kFarAtomChunk, p.n_atoms);
        float part[kFarK];
#pragma unroll
        for (int k = 0; k < kFarK; k++) part[k] = 0.f;
#pragma unroll 2
        for (int a = a0; a < a1; a++) {
#pragma clang fp contract(off)
            const float4 at = atoms[a];
            const float dy = at.y - py, dz = at.z - pz;
            const float dyz2 = __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, kFarBias));
            float rs[kFarK];
            int bin[kFarK];
#pragma unroll
            for (int k = 0; k < kFarK; k++) {
                const float dx = at.x - px[k];
                const float rb2 = __builtin_fmaf(dx, dx, dyz2);
                rs[k] = __builtin_amdgcn_rsqf(rb2);
                const float rb = rb2 * rs[k];
                bin[k] = CLAMP ? (int)fminf(rb, p.tmax) : (int)rb;
            }
            float tv[kFarK];
#pragma unroll
            for (int k = 0; k < kFarK; k++) tv[k] = T[bin[k]];
#pragma unroll
            for (int k = 0; k < kFarK; k++) part[k] = __builtin_fmaf(at.w, rs[k] * tv[k], part[k]);
        }
#pragma unroll
        for (int k = 0; k < kFarK; k++) {
#pragma clang fp contract(off)
            const float y = part[k] - c[k];
            const float t = s[k] + y;
            c[k] = (t - s[k]) - y;
            s[k] = t;
        }
    }
#pragma unroll
    for (int k = 0; k < kFarK; k++)
        if (ix0 + k < p.n1[0]) p.emap[base + k] = s[k];
}

}  // namespace

// ---------------------------------------------------------------------------
// context, plans and host orchestration
// ---------------------------------------------------------------------------

struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    void *get(size_t bytes) {
        if (bytes == 0) bytes = 16;
        if (bytes > cap) {
            if (ptr) HIP_CHECK(hipFree(ptr));
            ptr = nullptr;
            cap = 0;
            if (hipMalloc(&ptr, bytes) != hipSuccess) {
                ptr = nullptr;
                ag::fail(AG_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
            }
            cap = bytes;
        }
        return ptr;
    }
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (ptr) (void)hipFree(ptr);
    }
};

struct ag_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    DevBuf maps;                       // output staging of ag_maps_plain
    hipEvent_t ev[4] = {};
    float last_ms[4] = {0, 0, 0, 0};   // cell list, near, far, total of the last execute
};

// A receptor, grid, map set and tables resident on one device.
struct ag_plan {
    ag_ctx *ctx = nullptr;
    int N = 0, M = 0, T = 0, n_maps = 0, groups = 0;
    long long P = 0;
    int elec_index = -1;
    DevBuf coord, charge, vol, solpar, rvec, rvec2, type, hbond, rexp, disorder;
    DevBuf atom_cell, cell_count, cell_start, cell_cursor, slot_atom, near_atoms, far_atoms;
    DevBuf tab, sol, reps, fartab, status;
    CellParams cp{};
    int ncells = 0;
    AtomSrc src{};
    std::vector<NearParams> near;      // one per pass of 16 map slots
    int nplain = 12;                   // plain slots in use (4, 8 or 12)
    CovParams cov{};
    bool run_cov = false;
    bool want_fmap = false, want_cmap = false;
    ag_grid grid{};
    double coeff_estat = 0.0;
    bool run_far_atoms_ready = false;
    DevBuf fmin_atoms;
    FminParams fmin{};
    FarParams far{};
    bool run_far = false;
    bool far_clamp = true;
    std::vector<int> zero_slots;       // output maps that are all zero (Vina e/d)
};

namespace agdev {

ag_status context_create(ag_ctx **out, int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        ag::fail(AG_ERR_DEVICE, "no HIP device is available (this engine has no CPU fallback)");
    if (device < 0 || device >= n) ag::fail(AG_ERR_DEVICE, "HIP device index out of range");
    HIP_CHECK(hipSetDevice(device));
    ag_ctx *c = new ag_ctx();
    c->device = device;
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto &e : c->ev) HIP_CHECK(hipEventCreate(&e));
    *out = c;
    return AG_OK;
}

void context_destroy(ag_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

void context_timing(ag_ctx *c, float *ms) {
    for (int i = 0; i < 4; i++) ms[i] = c->last_ms[i];
}

template <typename T>
static T *upload(hipStream_t st, DevBuf &b, const T *src, size_t n) {
    T *d = (T *)b.get(sizeof(T) * n);
    if (n) HIP_CHECK(hipMemcpyAsync(d, src, sizeof(T) * n, hipMemcpyHostToDevice, st));
    return d;
}

static void validate(const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms, const ag_tables *tab,
                     bool vina) {
    const int N = rec->n_atoms, M = ms->n_atom_maps, T = tab->n_rec_types;
    if (N < 0 || N > AG_MAX_ATOMS) ag::fail(AG_ERR_BAD_INPUT, "atom count out of range");
    if (M < 0 || M + 2 > AG_MAX_MAPS) ag::fail(AG_ERR_BAD_INPUT, "map count out of range");
    if (T < 0 || T > AG_MAX_REC_TYPES) ag::fail(AG_ERR_BAD_INPUT, "receptor type count out of range");
    for (int d = 0; d < 3; d++)
        if (grid->n1[d] < 1 || grid->n1[d] > AG_MAX_GRID_PTS || grid->ne[d] < 0)
            ag::fail(AG_ERR_BAD_INPUT, "grid dimensions out of range");
    if (!(grid->spacing > 0)) ag::fail(AG_ERR_BAD_INPUT, "grid spacing must be positive");
    if (ms->constriction_grid && !(ms->constriction_cutoff <= 8.0))
        ag::fail(AG_ERR_UNSUPPORTED, "constriction_distance_cutoff above 8 A is not supported on the device");
    const int n_maps = M + (ms->elec_index >= 0) + (ms->dsolv_index >= 0);
    if (ms->n_maps != n_maps) ag::fail(AG_ERR_BAD_INPUT, "n_maps does not match the map set");
    if ((ms->elec_index >= 0 && ms->elec_index >= n_maps) || (ms->dsolv_index >= 0 && ms->dsolv_index >= n_maps))
        ag::fail(AG_ERR_BAD_INPUT, "e/d map slot out of range");
    for (int i = 0; i < N; i++) {
        if (rec->type[i] < 0 || rec->type[i] >= T) ag::fail(AG_ERR_BAD_INPUT, "atom type index out of range");
        if (rec->hbond[i] < 0 || rec->hbond[i] > 6) ag::fail(AG_ERR_BAD_INPUT, "atom hbond class out of range");
    }
    if (vina && ms->separate_desolvation)
        ag::fail(AG_ERR_UNSUPPORTED, "Vina-potential maps do not combine with separate_desolvation_maps");
    if (M > 0 && !vina) {
        bool any_h = false;
        for (int i = 0; i < N && !any_h; i++) any_h = rec->hbond[i] == AG_HB_DS || rec->hbond[i] == AG_HB_D1;
        if (!any_h) ag::fail(AG_ERR_NO_DONOR_H, "no closestH atom was found");
    }
    if (ms->hydrogen_type < -1 || ms->hydrogen_type >= T) ag::fail(AG_ERR_BAD_INPUT, "hydrogen type index out of range");
    if (M > 0 && !ms->maps) ag::fail(AG_ERR_BAD_INPUT, "map descriptors missing");
}

// Role of map m: the second map of each separate-desolvation pair carries
// that type's desolvation (main.cpp:3385).
static int map_role(const ag_mapset *ms, int m) {
    if (ms->maps[m].is_covalent) return ROLE_COV;
    if (!ms->separate_desolvation) return ROLE_AFF | ROLE_DES;
    int k = 0;
    while (k < m) k += ms->maps[k].is_covalent ? 1 : 2;
    return k == m ? ROLE_AFF : ROLE_DES;
}

// H-bond case of (probe class mh, atom class ah) (main.cpp:3286-3374).
static int hb_case(int mh, int ah) {
    const bool atom_don = (ah == 1 || ah == 2 || ah == 6);
    if ((mh == 3 || mh == 5 || mh == 6) && atom_don) return HBC_ACC_PROBE;
    if ((mh == 4 || mh == 6) && atom_don) return HBC_MINMAX_RACC;
    if ((mh == 1 || mh == 2 || mh == 6) && (ah > 2 || ah == 6)) return HBC_DON_PROBE;
    return HBC_PLAIN;
}

ag_plan *plan_create(ag_ctx *c, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                     const ag_tables *tab, int potential) {
    const bool vina = potential == AG_POTENTIAL_VINA;
    if (potential != AG_POTENTIAL_AD4 && !vina) ag::fail(AG_ERR_BAD_INPUT, "unknown potential");
    validate(rec, grid, ms, tab, vina);
    HIP_CHECK(hipSetDevice(c->device));
    std::unique_ptr<ag_plan> pl(new ag_plan());
    ag_plan &p = *pl;
    p.ctx = c;
    const int N = p.N = rec->n_atoms;
    const int M = p.M = ms->n_atom_maps;
    const int T = p.T = tab->n_rec_types;
    p.n_maps = ms->n_maps;
    p.elec_index = ms->elec_index;
    p.grid = *grid;
    p.coeff_estat = ms->coeff_estat;
    const long long P = p.P = (long long)grid->n1[0] * grid->n1[1] * grid->n1[2];
    hipStream_t st = c->stream;

    double *d_coord = upload(st, p.coord, rec->coord, 3 * (size_t)N);
    double *d_charge = upload(st, p.charge, rec->charge, N);
    double *d_vol = upload(st, p.vol, rec->vol, N);
    double *d_solpar = upload(st, p.solpar, rec->solpar, N);
    double *d_rvec = upload(st, p.rvec, rec->rvector, 3 * (size_t)N);
    double *d_rvec2 = upload(st, p.rvec2, rec->rvector2, 3 * (size_t)N);
    int *d_type = upload(st, p.type, rec->type, N);
    int *d_hbond = upload(st, p.hbond, rec->hbond, N);
    int *d_rexp = upload(st, p.rexp, rec->rexp, N);
    unsigned char *d_dis = upload(st, p.disorder, rec->disorder, N);
    p.src = AtomSrc{d_coord, d_charge, d_vol, d_solpar, d_rvec, d_rvec2, d_type, d_hbond, d_rexp, d_dis,
                    ms->coeff_estat, ms->solpar_q};

    // cell list geometry: the box padded by the staging radius
    const double cs = 4.0;
    p.cp.cs = cs;
    p.cp.n_atoms = N;
    for (int d = 0; d < 3; d++) {
        const double lo = -grid->ne[d] * grid->spacing - 8.1;
        const double hi = (grid->n1[d] - 1 - grid->ne[d]) * grid->spacing + 8.1;
        p.cp.origin[d] = lo;
        p.cp.nc[d] = (int)ceil((hi - lo) / cs) + 1;
    }
    p.ncells = p.cp.nc[0] * p.cp.nc[1] * p.cp.nc[2];
    p.atom_cell.get(sizeof(int) * (size_t)N);
    p.cell_count.get(sizeof(int) * (size_t)p.ncells);
    p.cell_start.get(sizeof(int) * ((size_t)p.ncells + 1));
    p.cell_cursor.get(sizeof(int) * (size_t)p.ncells);
    p.slot_atom.get(sizeof(int) * (size_t)N);
    p.near_atoms.get(sizeof(NearAtom) * (size_t)N);
    p.status.get(sizeof(int));

    // electrostatic factor per bin: 332/eps for the distance-dependent
    // dielectric, 332/diel for a constant one (main.cpp:2960-2978)
    std::vector<float> reps(AG_NDIEL);
    for (int k = 0; k < AG_NDIEL; k++) reps[k] = ms->dddiel ? (float)tab->r_eps[k] : (float)ms->inv_diel_cal;
    float *d_reps = upload(st, p.reps, reps.data(), AG_NDIEL);
    std::vector<float> sol(AG_NEINT + 1);
    for (int k = 0; k <= AG_NEINT; k++) sol[k] = (float)tab->sol_fn[k];
    float *d_sol = upload(st, p.sol, sol.data(), AG_NEINT + 1);

    // Map slots of the near-kernel passes.  Maps without H-bond terms (and
    // the desolvation halves of separate_desolvation_maps) take plain slots,
    // H-bonding probes take the H-bond slots; covalent maps take none.  With
    // the Vina potential every map is a plain table sum: no directional
    // H-bond terms, no desolvation, no e or d map.
    std::vector<int> plain_maps, hb_maps;
    std::vector<int> cov_maps;
    for (int m = 0; m < M; m++) {
        const int role = map_role(ms, m);
        if (role == ROLE_COV) cov_maps.push_back(m);
        else if (!vina && (role & ROLE_AFF) && ms->maps[m].is_hbonder) hb_maps.push_back(m);
        else plain_maps.push_back(m);
    }
    p.groups = std::max<int>(1, std::max((int)(plain_maps.size() + kPlainSlots - 1) / kPlainSlots,
                                         (int)(hb_maps.size() + kHbSlots - 1) / kHbSlots));
    const size_t gstride = (size_t)(T > 0 ? T : 1) * AG_NEINT * kMapSlots;
    std::vector<float> gtab(gstride * p.groups, 0.f);
    std::vector<std::array<int, kMapSlots>> slot_map(p.groups);
    int nplain_max = 0;
    for (int g = 0; g < p.groups; g++) {
        slot_map[g].fill(-1);
        int np_g = 0;
        for (int j = 0; j < kPlainSlots; j++) {
            const size_t i = (size_t)g * kPlainSlots + j;
            if (i < plain_maps.size()) { slot_map[g][j] = plain_maps[i]; np_g = j + 1; }
        }
        nplain_max = std::max(nplain_max, np_g);
        for (int k = 0; k < kHbSlots; k++) {
            const size_t i = (size_t)g * kHbSlots + k;
            if (i < hb_maps.size()) slot_map[g][kPlainSlots + k] = hb_maps[i];
        }
        for (int t = 0; t < T; t++)
            for (int j = 0; j < kMapSlots; j++) {
                const int m = slot_map[g][j];
                if (m < 0) continue;
                const float *srcrow = tab->e_vdw_hb + ((size_t)t * M + m) * AG_NEINT;
                float *dst = gtab.data() + g * gstride + (size_t)t * AG_NEINT * kMapSlots + j;
                for (int k = 0; k < AG_NEINT; k++) dst[(size_t)k * kMapSlots] = srcrow[k];
            }
    }
    p.nplain = nplain_max <= 4 ? 4 : (nplain_max <= 8 ? 8 : 12);
    float *d_tab = upload(st, p.tab, gtab.data(), gtab.size());

    // near-kernel parameters, one set per pass
    NearParams np;
    memset(&np, 0, sizeof np);
    for (int d = 0; d < 3; d++) {
        np.n1[d] = grid->n1[d];
        np.ne[d] = grid->ne[d];
        np.cell_origin[d] = p.cp.origin[d];
        np.nc[d] = p.cp.nc[d];
    }
    np.spacing = grid->spacing;
    np.s128 = (float)(grid->spacing * 128.0);
    np.nb[0] = (grid->n1[0] + kBX - 1) / kBX;
    np.nb[1] = (grid->n1[1] + kBY - 1) / kBY;
    np.nb[2] = (grid->n1[2] + kBZ - 1) / kBZ;
    np.nbricks = np.nb[0] * np.nb[1] * np.nb[2];
    np.cell_size = cs;
    np.cell_start = (const int *)p.cell_start.ptr;
    np.atoms = (const NearAtom *)p.near_atoms.ptr;
    np.coord64 = d_coord;
    np.rvec64 = d_rvec;
    np.rvec2_64 = d_rvec2;
    np.sol = d_sol;
    np.reps = d_reps;
    np.disorder_h = ms->disorder_h;
    np.hydrogen_type = ms->hydrogen_type;
    np.P = P;
    np.status = (int *)p.status.ptr;
    for (int g = 0; g < p.groups; g++) {
        NearParams q = np;
        bool used = false;
        for (int j = 0; j < kMapSlots; j++) {
            const int m = slot_map[g][j];
            q.out_slot[j] = m;
            q.kE[j] = 0.f;
            q.sp[j] = 0.f;
            q.volp[j] = 0.f;
            if (m < 0) continue;
            used = true;
            const int role = map_role(ms, m);
            const ag_map_desc &md = ms->maps[m];
            q.kE[j] = (role & ROLE_AFF) ? 1.f : 0.f;
            if ((role & ROLE_DES) && !vina) {
                q.sp[j] = (float)md.solpar_probe;
                q.volp[j] = (float)md.vol_probe;
            }
            if (j >= kPlainSlots) {
                const int k = j - kPlainSlots;
                for (int ah = 0; ah <= 6; ah++) {
                    const int cse = hb_case(md.hbond, ah);
                    int ordered = 0, disordered = 0;
                    if (cse == HBC_PLAIN) ordered = disordered = HC_PLAIN;
                    else if (cse == HBC_ACC_PROBE) ordered = disordered = HC_HRAMP | HC_RACC;
                    else if (cse == HBC_MINMAX_RACC) ordered = disordered = HC_MINMAX | HC_RACC;
                    else {  // HBC_DON_PROBE
                        ordered = HC_MINMAX;
                        disordered = HC_HRAMP;
                    }
                    q.hcode[k][ah] = ordered | (disordered << 4);
                }
            }
        }
        q.tab = d_tab + g * gstride;
        q.e_slot = (g == 0 && !vina) ? ms->elec_index : -1;
        q.d_slot = (g == 0 && !vina) ? ms->dsolv_index : -1;
        q.has_affinity = used || q.d_slot >= 0;
        p.near.push_back(q);
    }

    // floating and constriction grids
    p.want_fmap = ms->floating_grid != 0;
    p.want_cmap = ms->constriction_grid != 0;
    if (p.want_cmap) p.near[0].ccut = ms->constriction_cutoff;
    if (p.want_fmap) {
        float4 *d_fa = (float4 *)p.fmin_atoms.get(sizeof(float4) * (size_t)std::max(N, 1));
        if (N > 0) k_fmin_atoms<<<(N + 255) / 256, 256, 0, st>>>(N, d_coord, d_fa);
        HIP_CHECK(hipGetLastError());
        FminParams &fm = p.fmin;
        for (int d = 0; d < 3; d++) {
            fm.n1[d] = grid->n1[d];
            fm.ne[d] = grid->ne[d];
        }
        fm.spacing = (float)grid->spacing;
        fm.runs_per_row = (grid->n1[0] + kFarK - 1) / kFarK;
        fm.nruns = fm.runs_per_row * grid->n1[1] * grid->n1[2];
        fm.atoms = d_fa;
        fm.n_atoms = N;
    }

    // covalent maps
    if (!cov_maps.empty()) {
        if (cov_maps.size() > 16) ag::fail(AG_ERR_UNSUPPORTED, "more than 16 covalent maps");
        CovParams &cv = p.cov;
        memset(&cv, 0, sizeof cv);
        for (int d = 0; d < 3; d++) {
            cv.n1[d] = grid->n1[d];
            cv.ne[d] = grid->ne[d];
            cv.pos[d] = ms->cov_pos[d];
        }
        cv.spacing = grid->spacing;
        cv.half_width = ms->cov_half_width;
        cv.barrier = ms->cov_barrier;
        cv.n_cov = (int)cov_maps.size();
        for (int k = 0; k < cv.n_cov; k++) cv.slot[k] = cov_maps[k];
        cv.P = P;
        p.run_cov = true;
    }

    // far kernel (e map): per-atom float4 and the table with bins < 1025 zeroed
    // Vina: the e and d maps stay zero (main.cpp:2968, 3391)
    if (vina) {
        if (ms->elec_index >= 0) p.zero_slots.push_back(ms->elec_index);
        if (ms->dsolv_index >= 0) p.zero_slots.push_back(ms->dsolv_index);
    }
    p.run_far = !vina && ms->elec_index >= 0 && N > 0;
    if (p.run_far) {
        float4 *d_far = (float4 *)p.far_atoms.get(sizeof(float4) * (size_t)N);
        k_far_atoms<<<(N + 255) / 256, 256, 0, st>>>(N, d_coord, d_charge, ms->coeff_estat, d_far);
        HIP_CHECK(hipGetLastError());
        p.run_far_atoms_ready = true;
        // largest bin any (point, atom) pair reaches: farthest box corner
        double rmax = 0.0;
        for (int i = 0; i < N; i++) {
            double q = 0.0;
            for (int d = 0; d < 3; d++) {
                const double lo = -grid->ne[d] * grid->spacing, hi = (grid->n1[d] - 1 - grid->ne[d]) * grid->spacing;
                const double a = rec->coord[3 * (size_t)i + d];
                const double e = std::max(std::fabs(a - lo), std::fabs(a - hi));
                q += e * e;
            }
            rmax = std::max(rmax, q);
        }
        const long long need = (long long)(std::sqrt(rmax) * 128.0) + 8;
        p.far_clamp = need > kFarTabMax;
        const int tab_len = p.far_clamp ? AG_NDIEL : (int)std::max<long long>(need, kFarBinMin + 1);
        std::vector<float> ft(tab_len);
        for (int k = 0; k < tab_len; k++) ft[k] = k < kFarBinMin ? 0.f : reps[std::min(k, AG_NDIEL - 1)];
        float *d_ft = upload(st, p.fartab, ft.data(), ft.size());
        FarParams &fp = p.far;
        memset(&fp, 0, sizeof fp);
        for (int d = 0; d < 3; d++) {
            fp.n1[d] = grid->n1[d];
            fp.ne[d] = grid->ne[d];
        }
        fp.s128 = (float)(grid->spacing * 128.0);
        fp.runs_per_row = (grid->n1[0] + kFarK - 1) / kFarK;
        fp.nruns = fp.runs_per_row * grid->n1[1] * grid->n1[2];
        fp.atoms = d_far;
        fp.n_atoms = N;
        fp.tab = d_ft;
        fp.tab_len = tab_len;
        fp.tmax = (float)(tab_len - 1);
    }
    HIP_CHECK(hipStreamSynchronize(st));   // host staging vectors go out of scope
    return pl.release();
}

void plan_destroy(ag_plan *p) {
    if (!p) return;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    delete p;
}

// Cell list + near kernel (+ far kernel) into d_maps ([n_maps][P], device).
// Asynchronous on the context's stream up to the final status check.
void plan_execute(ag_plan *p, float *d_maps, float *d_fmap, int32_t *d_cmap) {
    if ((p->want_fmap && !d_fmap) || (p->want_cmap && !d_cmap))
        ag::fail(AG_ERR_BAD_INPUT, "the map set asks for a floating/constriction grid but no output was given");
    ag_ctx *c = p->ctx;
    HIP_CHECK(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const int N = p->N;
    const int tpb = 256;
    int *d_atom_cell = (int *)p->atom_cell.ptr, *d_count = (int *)p->cell_count.ptr;
    int *d_start = (int *)p->cell_start.ptr, *d_cursor = (int *)p->cell_cursor.ptr;
    int *d_slot_atom = (int *)p->slot_atom.ptr;
    int *d_status = (int *)p->status.ptr;
    HIP_CHECK(hipEventRecord(c->ev[0], st));
    HIP_CHECK(hipMemsetAsync(d_count, 0, sizeof(int) * (size_t)p->ncells, st));
    HIP_CHECK(hipMemsetAsync(d_cursor, 0, sizeof(int) * (size_t)p->ncells, st));
    HIP_CHECK(hipMemsetAsync(d_status, 0, sizeof(int), st));
    if (N > 0) k_cell_count<<<(N + tpb - 1) / tpb, tpb, 0, st>>>(p->cp, (const double *)p->coord.ptr, d_atom_cell, d_count);
    k_scan<<<1, 1024, 0, st>>>(d_count, d_start, p->ncells);
    if (N > 0) k_cell_scatter<<<(N + tpb - 1) / tpb, tpb, 0, st>>>(N, d_atom_cell, d_start, d_cursor, d_slot_atom);
    k_cell_sort<<<(p->ncells + tpb - 1) / tpb, tpb, 0, st>>>(p->ncells, d_start, d_slot_atom);
    if (N > 0)
        k_gather<<<(N + tpb - 1) / tpb, tpb, 0, st>>>(d_start + p->ncells, d_slot_atom, p->src,
                                                      (NearAtom *)p->near_atoms.ptr);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(c->ev[1], st));
    bool first = true;
    for (auto &q : p->near) {
        NearParams np = q;
        np.maps = d_maps;
        np.cmap = (first && p->want_cmap) ? d_cmap : nullptr;
        first = false;
        if (!(np.has_affinity || np.e_slot >= 0 || np.cmap)) continue;
        if (p->nplain == 4) k_near<4><<<np.nbricks, kNearThreads, 0, st>>>(np);
        else if (p->nplain == 8) k_near<8><<<np.nbricks, kNearThreads, 0, st>>>(np);
        else k_near<12><<<np.nbricks, kNearThreads, 0, st>>>(np);
        HIP_CHECK(hipGetLastError());
    }
    if (p->want_fmap) {
        FminParams fm = p->fmin;
        fm.fmap = d_fmap;
        k_fmin<<<(fm.nruns + 255) / 256, 256, 0, st>>>(fm);
        HIP_CHECK(hipGetLastError());
    }
    if (p->run_cov) {
        CovParams cv = p->cov;
        cv.maps = d_maps;
        k_covalent<<<(unsigned)((p->P + 255) / 256), 256, 0, st>>>(cv);
        HIP_CHECK(hipGetLastError());
    }
    for (int m : p->zero_slots) {
        k_fill<<<(unsigned)((p->P + 255) / 256), 256, 0, st>>>(d_maps + (size_t)m * p->P, p->P, 0.f);
        HIP_CHECK(hipGetLastError());
    }
    HIP_CHECK(hipEventRecord(c->ev[2], st));
    if (p->run_far) {
        FarParams fp = p->far;
        fp.emap = d_maps + (size_t)p->elec_index * p->P;
        const int blocks = (fp.nruns + kFarThreads - 1) / kFarThreads;
        if (p->far_clamp) k_far<true><<<blocks, kFarThreads, sizeof(float) * fp.tab_len, st>>>(fp);
        else k_far<false><<<blocks, kFarThreads, sizeof(float) * fp.tab_len, st>>>(fp);
        HIP_CHECK(hipGetLastError());
    }
    HIP_CHECK(hipEventRecord(c->ev[3], st));
    int status = 0;
    HIP_CHECK(hipMemcpyAsync(&status, d_status, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    HIP_CHECK(hipEventElapsedTime(&c->last_ms[0], c->ev[0], c->ev[1]));
    HIP_CHECK(hipEventElapsedTime(&c->last_ms[1], c->ev[1], c->ev[2]));
    HIP_CHECK(hipEventElapsedTime(&c->last_ms[2], c->ev[2], c->ev[3]));
    HIP_CHECK(hipEventElapsedTime(&c->last_ms[3], c->ev[0], c->ev[3]));
    if (status & 1) ag::fail(AG_ERR_UNSUPPORTED, "more than 2048 atoms within 8 A of one 8x8x4 brick of grid points");
}

// One map set, handed to the host map by map: each map is copied into host
// memory as soon as the kernels are done and on_ready(j) is called for it
// (j = n_maps for fmap, n_maps + 1 for cmap), so the caller can format and
// write map j while later maps are still in flight.
void maps_plain_streamed(ag_ctx *c, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                         const ag_tables *tab, float *maps_out, float *fmap, int32_t *cmap,
                         const std::function<void(int)> &on_ready, int potential) {
    std::unique_ptr<ag_plan, void (*)(ag_plan *)> p(plan_create(c, rec, grid, ms, tab, potential), plan_destroy);
    const size_t pbytes = sizeof(float) * (size_t)p->P;
    const size_t bytes = pbytes * (size_t)p->n_maps;
    float *d_maps = (float *)c->maps.get(bytes + 2 * pbytes + 256);
    char *extra = (char *)d_maps + bytes;
    float *d_fmap = ms->floating_grid ? (float *)extra : nullptr;
    int32_t *d_cmap = ms->constriction_grid ? (int32_t *)(extra + pbytes) : nullptr;
    plan_execute(p.get(), d_maps, d_fmap, d_cmap);
    for (int j = 0; j < p->n_maps; j++) {
        HIP_CHECK(hipMemcpyAsync(maps_out + (size_t)j * p->P, d_maps + (size_t)j * p->P, pbytes,
                                 hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        on_ready(j);
    }
    if (d_fmap) {
        HIP_CHECK(hipMemcpyAsync(fmap, d_fmap, pbytes, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        on_ready(p->n_maps);
    }
    if (d_cmap) {
        HIP_CHECK(hipMemcpyAsync(cmap, d_cmap, pbytes, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        on_ready(p->n_maps + 1);
    }
}

void plan_audit(ag_plan *p, long long out[4]) {
    ag_ctx *c = p->ctx;
    HIP_CHECK(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    if (!p->run_far_atoms_ready) {
        float4 *d_far = (float4 *)p->far_atoms.get(sizeof(float4) * (size_t)std::max(p->N, 1));
        if (p->N > 0)
            k_far_atoms<<<(p->N + 255) / 256, 256, 0, st>>>(p->N, (const double *)p->coord.ptr,
                                                            (const double *)p->charge.ptr, p->coeff_estat, d_far);
        p->run_far_atoms_ready = true;
    }
    DevBuf acc;
    unsigned long long *d_out = (unsigned long long *)acc.get(4 * sizeof(unsigned long long));
    HIP_CHECK(hipMemsetAsync(d_out, 0, 4 * sizeof(unsigned long long), st));
    AuditParams ap;
    memset(&ap, 0, sizeof ap);
    for (int d = 0; d < 3; d++) {
        ap.n1[d] = p->grid.n1[d];
        ap.ne[d] = p->grid.ne[d];
    }
    ap.spacing = p->grid.spacing;
    ap.s128 = (float)(p->grid.spacing * 128.0);
    ap.far_atoms = (const float4 *)p->far_atoms.ptr;
    ap.coord = (const double *)p->coord.ptr;
    ap.n_atoms = p->N;
    ap.P = p->P;
    ap.out = d_out;
    k_audit<<<(unsigned)((p->P + 255) / 256), 256, 0, st>>>(ap);
    HIP_CHECK(hipGetLastError());
    unsigned long long h[4];
    HIP_CHECK(hipMemcpyAsync(h, d_out, sizeof h, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    for (int i = 0; i < 4; i++) out[i] = (long long)h[i];
}

void maps_plain(ag_ctx *c, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms, const ag_tables *tab,
                float *maps_out, float *fmap, int32_t *cmap, int outputs_on_device, int potential) {
    if ((ms->floating_grid && !fmap) || (ms->constriction_grid && !cmap))
        ag::fail(AG_ERR_BAD_INPUT, "the map set asks for a floating/constriction grid but no output was given");
    std::unique_ptr<ag_plan, void (*)(ag_plan *)> p(plan_create(c, rec, grid, ms, tab, potential), plan_destroy);
    const size_t bytes = sizeof(float) * (size_t)p->n_maps * p->P;
    const size_t pbytes = sizeof(float) * (size_t)p->P;
    float *d_maps = outputs_on_device ? maps_out : (float *)c->maps.get(bytes + 2 * pbytes + 256);
    float *d_fmap = nullptr;
    int32_t *d_cmap = nullptr;
    if (outputs_on_device) {
        d_fmap = ms->floating_grid ? fmap : nullptr;
        d_cmap = ms->constriction_grid ? cmap : nullptr;
    } else {
        char *base = (char *)d_maps + bytes;
        if (ms->floating_grid) d_fmap = (float *)base;
        if (ms->constriction_grid) d_cmap = (int32_t *)(base + pbytes);
    }
    plan_execute(p.get(), d_maps, d_fmap, d_cmap);
    if (!outputs_on_device) {
        HIP_CHECK(hipMemcpyAsync(maps_out, d_maps, bytes, hipMemcpyDeviceToHost, c->stream));
        if (d_fmap) HIP_CHECK(hipMemcpyAsync(fmap, d_fmap, pbytes, hipMemcpyDeviceToHost, c->stream));
        if (d_cmap) HIP_CHECK(hipMemcpyAsync(cmap, d_cmap, pbytes, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    }
}

}  // namespace agdev
