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
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <type_traits>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "ag_internal.h"
#include "ag_devcommon.h"

// Device scratch of the allocation-free entry point (ag_maps_plain_noalloc):
// one array per GPU, present from the library's load until the process
// ends.  Host code never uses this name as an address: each device's
// address comes from hipGetSymbolAddress at context creation (ag_ctx::scratch).
static __device__ __align__(256) unsigned char scratch[AG_NOALLOC_SCRATCH_BYTES];

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
constexpr int kFarK = 16;                         // points per lane, far kernel
constexpr int kFarRuns = 256;                     // runs (lanes of kFarK points) per far-kernel item
constexpr int kFarThreads = 2 * kFarRuns;         // two halves of the atom list per run
constexpr int kFarTabMax = 16380;                 // LDS table entries without the clamp (two workgroups per CU)
constexpr int kFarAtomChunk = 128;                // Kahan block
enum : int { kFarTable = 0, kFarTfree = 1, kFarMixed = 2 };   // far-kernel paths of a launch
constexpr int kFarBinMin = 1025;                  // far kernel takes bins >= this
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

// One receptor conformation on its own grid, as the kernels see it.  A plan
// holds one; a batch holds many, their atoms, cells, bricks and far-kernel
// tiles concatenated in conformation order.
struct ConfDev {
    int n1[3], ne[3];
    double spacing;
    float s128;                   // (float)(spacing * 128)
    int nb[3];                    // bricks per axis
    double cell_origin[3];
    int nc[3];
    int cell_base;                // first cell of this conformation in the cell arrays
    int atom_base, n_atoms;       // atom range in the concatenated atom arrays
    long long P;
    long long map_off;            // offset of its [n_maps][P] block in the map output
    long long pt_off;             // offset of its [P] block in the fmap / cmap outputs
    int runs_per_row, nruns;      // far-kernel runs of kFarK points along x
    int far_const;                // far kernel: the table-free path for atoms past bin 8192 of a wave's box
    int far_clamp;                // far kernel: some pair may pass the table's end
};

// Read-only device tables (conformation records, work prefixes) are read
// through the constant address space: no kernel writes them, so uniform
// reads become scalar loads the compiler may repeat instead of holding the
// values in registers across the hot loops.
#define AG_CONST __attribute__((address_space(4)))
// float4 layout without HIP_vector_type's constructors, for copies out of
// the constant address space.
struct __align__(16) PlainF4 {
    float x, y, z, w;
};
__device__ __forceinline__ PlainF4 ld4(const AG_CONST PlainF4 *p, int i) {
    const AG_CONST PlainF4 &q = p[i];
    return PlainF4{q.x, q.y, q.z, q.w};
}
template <typename T>
__device__ __forceinline__ const AG_CONST T *as_const(const T *p) {
    return (const AG_CONST T *)(p);
}

// Work lists: item k of a launch belongs to conformation c with
// prefix[c] <= k < prefix[c + 1].
__device__ __forceinline__ int find_conf(const AG_CONST int *prefix, int nconf, int item) {
    int lo = 0, hi = nconf;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= item) lo = mid;
        else hi = mid;
    }
    return lo;
}

struct NearParams {
    // conformations and work distribution
    const ConfDev *confs;
    const int *brick_prefix;      // [nconf + 1]
    int nconf, total_bricks;
    int item_base;                // first brick of this launch
    double cell_size;
    const int *cell_start;        // [total cells + 1]
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
    int bhtree;                   // ag_mapset.bhtree: Loop 2 and cmap by the tree's FP32 test
    unsigned char *collide;       // tree build, collisions tested: 1 per point within 1.5 A of an atom (pass 0)
    float *maps;
    int *status;
};

struct FarParams {
    const ConfDev *confs;
    const int *tile_prefix;       // [nconf + 1], items of kFarRuns runs
    int nconf, total_tiles;
    int item_base, item_end;      // tiles [item_base, item_end) of this launch
    int *queue;                   // null: one tile per workgroup; else a persistent work counter
    const float4 *atoms;          // (xE, yE, zE, w = q * coeff_estat * 128)
    const float *tab;             // [tab_len]
    int tab_len;
    unsigned tmax4;               // 4 * (tab_len - 1): byte offset of the last entry
    float c_last;                 // the table's last value, r_eps[8191] (every bin >= 8191)
    float *maps;                  // e map slot: in = near part, out = total
    int e_slot;
};

// Consecutive workgroup ids go round-robin over the 8 XCDs; here XCD x takes
// chunks x, x + 8, x + 16, ... of C consecutive items, so that every XCD
// gets items from all over the grid (the work per item varies with the
// local atom density, and in the far kernel's table-free path with the
// distance to the atoms) while the items of a chunk share its L2.  Ids past
// the last whole round of 8 chunks map to themselves.
template <int C>
__device__ __forceinline__ int xcd_chunked(int bid, int nblocks) {
    const int full = nblocks / (8 * C) * (8 * C);
    if (bid >= full) return bid;
    const int x = bid % 8, k = bid / 8;
    return ((k / C) * 8 + x) * C + k % C;
}

// Table offset of the far kernel's bin: with r = rb2 * rsq(rb2) (128 r in
// bins), y = fma(rb2, rsq(rb2), 2^21 - 1/8) lies in [2^21, 2^22), where the
// float spacing is 1/4, so the low 23 bits of y hold m = round(4 r - 1/2)
// (ties to even).  For n <= r < n + 1, 4n - 1/2 <= 4r - 1/2 < 4n + 7/2 puts m
// in {4n, ..., 4n + 3} (the tie at r = n goes to the even 4n), so
// m & ~3 = 4 floor(r): the byte offset of bin floor(r) in a float table, in
// one fma and one and.  Valid for r < 2^21 - 1.
constexpr float kBinMagic = 2097151.875f;         // 2^21 - 1/8
__device__ __forceinline__ unsigned bin4_of(float rb2, float rs) {
    return __float_as_uint(__builtin_fmaf(rb2, rs, kBinMagic)) & 0x7FFFFCu;
}

// Far-kernel geometry of a run of kFarK points along x.  The run's points
// sit at cx + dk, dk = (k - kFarMid) s128 (k = 0..kFarK-1, cx the run's
// centre, s128 = 128 spacing), and with u = ax - cx
//   rb2 = (u - dk)^2 + dy^2 + dz^2 + bias
//       = fma(u s128, 15 - 2k, fma(s128^2, (j + 1/2)^2, fma(u, u, dyz2))),
// j = |k - kFarMid| - 1/2: the points k and kFarK-1-k share the middle
// term, and both factors 15 - 2k and (j + 1/2)^2 are exact small constants
// the instructions carry as literals, so no register holds them.  Per point
// that is 1.5 operations instead of a subtract and an fma.  The roundings
// of the three products and sums are below 2^-21 (7.5 s128)^2 in total when
// u = dk, where the exact value is only dyz2; the bias (a floor for dyz2)
// is 2^-19 (7.5 s128)^2, twice that, which keeps rb2 > 0 and rsq finite.
// In bins (128 r) it moves r by bias / (2 r), about 1e-4 of a bin at the far
// field's r >= 1025 bins and 0.375 A spacing.
constexpr float kFarMid = 0.5f * (kFarK - 1);

__device__ __forceinline__ float far_bias(float s128) {
    const float d = kFarMid * s128;
    return fmaxf(1.0e-3f, d * d * 0x1p-19f);
}
__device__ __forceinline__ float far_cx(int ix0, int ne0, float s128) { return ((float)(ix0 - ne0) + kFarMid) * s128; }
// (j + 1/2)^2 for point k, and the factor 15 - 2k of u s128
__device__ __forceinline__ constexpr float far_c2(int k) {
    return (k >= kFarK / 2 ? (float)(k - kFarK / 2) + 0.5f : (float)(kFarK / 2 - 1 - k) + 0.5f) *
           (k >= kFarK / 2 ? (float)(k - kFarK / 2) + 0.5f : (float)(kFarK / 2 - 1 - k) + 0.5f);
}
__device__ __forceinline__ constexpr float far_m(int k) { return (float)(kFarK - 1 - 2 * k); }

// Row q (0 <= q < n1y * n1z) of the far kernel's run order -> (iy, iz).
// Rows go in blocks of 8 y by 8 z (smaller at the grid's upper edges), y
// fastest inside a block, blocks in y order within each band of 8 z.  The
// 64 lanes of a wave then hold runs of one x-run from an 8 x 8 patch of
// rows rather than eight whole x-rows, so their distances to an atom, and
// the table entries they read, spread less: fewer LDS bank conflicts.
__device__ __forceinline__ void far_row(int q, int n1y, int n1z, int &iy, int &iz) {
    const int bz = q / (8 * n1y);
    const int o = q - bz * 8 * n1y;
    const int nzb = min(8, n1z - 8 * bz);   // >= 1 for q in range
    const int by = o / (8 * nzb);
    const int o2 = o - by * 8 * nzb;
    const int nyb = min(8, n1y - 8 * by);   // >= 1 for q in range
    iy = 8 * by + o2 % nyb;
    iz = 8 * bz + o2 / nyb;
}

// The far kernel's per-pair table offset for point k of the run centred at
// cx, reproduced bit for bit by the near kernel when it decides which atoms
// the far kernel left out, and by the audit (same operations in the same
// order; the build turns off FP contraction).
__device__ __forceinline__ unsigned far_bin4(float ax, float ay, float az, float cx, int k, float s128, float py,
                                             float pz, float bias) {
    const float dy = ay - py, dz = az - pz;
    const float dyz2 = __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, bias));
    const float u = ax - cx;
    const float U = __builtin_fmaf(u, u, dyz2);
    const float W = __builtin_fmaf(s128 * s128, far_c2(k), U);
    const float rb2 = __builtin_fmaf(u * s128, far_m(k), W);
    return bin4_of(rb2, __builtin_amdgcn_rsqf(rb2));
}

// ---------------------------------------------------------------------------
// cell list
// ---------------------------------------------------------------------------

__device__ __forceinline__ int cell_of(const ConfDev &c, double cs, const double *p) {
    int ix = (int)floor((p[0] - c.cell_origin[0]) / cs);
    int iy = (int)floor((p[1] - c.cell_origin[1]) / cs);
    int iz = (int)floor((p[2] - c.cell_origin[2]) / cs);
    if (ix < 0 || iy < 0 || iz < 0 || ix >= c.nc[0] || iy >= c.nc[1] || iz >= c.nc[2]) return -1;
    return c.cell_base + ix + c.nc[0] * (iy + c.nc[1] * iz);
}

// atom_conf: conformation of each atom, or null when there is one.
__global__ void k_cell_count(const ConfDev *confs, const int *atom_conf, int n, double cs, const double *coord,
                             int *atom_cell, int *count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cell = cell_of(confs[atom_conf ? atom_conf[i] : 0], cs, coord + 3 * (size_t)i);
    atom_cell[i] = cell;
    if (cell >= 0) atomicAdd(&count[cell], 1);
}

// Clears the cell counts and cursors [0, n), the status word and the nq
// work counters of one map set, in one launch: a kernel rather than memset
// nodes, since a plan's map set is captured in a graph.
__global__ void k_zero_cells(int *count, int *cursor, int n, int *status, int *queue, int nq) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        count[i] = 0;
        cursor[i] = 0;
    }
    if (i < nq) queue[i] = 0;
    if (i == 0) *status = 0;
}

// Exclusive scan of count[0..n) into start[0..n], one workgroup of 1024
// walking tiles of 1024 * kScanPer cells.  Each thread takes kScanPer
// consecutive cells of a tile (16-byte loads and stores, so a wave covers
// one contiguous 4 KiB span); the thread sums are scanned within each wave
// by shuffles and across the 16 waves through LDS, and the total of the
// tiles before is carried in a register.
constexpr int kScanPer = 16;

__global__ void __launch_bounds__(1024) k_scan(const int *count, int *start, int n) {
    __shared__ int wave_sum[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int carry = 0;
    for (int base = 0; base < n; base += 1024 * kScanPer) {   // uniform: every thread reaches both barriers
        const int i0 = base + t * kScanPer;
        const bool whole = i0 + kScanPer <= n;
        int v[kScanPer];
        if (whole) {
            const int4 *q = reinterpret_cast<const int4 *>(count + i0);
#pragma unroll
            for (int j = 0; j < kScanPer / 4; j++) {
                const int4 a = q[j];
                v[4 * j] = a.x;
                v[4 * j + 1] = a.y;
                v[4 * j + 2] = a.z;
                v[4 * j + 3] = a.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < kScanPer; j++) v[j] = i0 + j < n ? count[i0 + j] : 0;
        }
        int s = 0;
#pragma unroll
        for (int j = 0; j < kScanPer; j++) s += v[j];
        int x = s;   // inclusive scan over the wave
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) wave_sum[w] = x;
        __syncthreads();
        int before = 0, total = 0;
        for (int k = 0; k < 16; k++) {
            const int ws = wave_sum[k];
            before += k < w ? ws : 0;
            total += ws;
        }
        __syncthreads();   // the next tile rewrites wave_sum
        int run = carry + before + x - s;
        if (whole) {
            int4 *o = reinterpret_cast<int4 *>(start + i0);
#pragma unroll
            for (int j = 0; j < kScanPer / 4; j++) {
                int4 r;
                r.x = run;
                r.y = r.x + v[4 * j];
                r.z = r.y + v[4 * j + 1];
                r.w = r.z + v[4 * j + 2];
                run = r.w + v[4 * j + 3];
                o[j] = r;
            }
        } else {
#pragma unroll
            for (int j = 0; j < kScanPer; j++) {
                if (i0 + j < n) start[i0 + j] = run;
                run += v[j];
            }
        }
        carry += total;
    }
    if (t == 0) start[n] = carry;
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

// A thread's point inside the 8x8x4 brick: each wave takes an 8x8x1 slab
// (CUBE false) or a 4x4x4 cube (CUBE true) of points.
template <bool CUBE>
__device__ __forceinline__ void brick_lane(int tid, int &lx, int &ly, int &lz) {
    if (CUBE) {
        const int wave = tid >> 6, lane = tid & 63;
        lx = (wave & 1) * 4 + (lane & 3);
        ly = (wave >> 1) * 4 + ((lane >> 2) & 3);
        lz = lane >> 4;
    } else {
        lx = tid % kBX;
        ly = (tid / kBX) % kBY;
        lz = tid / (kBX * kBY);
    }
}

// CUBE selects the lane layout inside the 8x8x4 brick (brick_lane).  FORM:
//   kNearPlain   neither the tree build's queries, a constriction count
//                nor disordered hydroxyls (their branches compiled out);
//   kNearGeneral every map-set option;
//   kNearPaged   every option, and candidate / donor-H lists longer than
//                one LDS page.
// The unpaged forms leave a brick whose candidates (or donor H) do not fit
// one LDS list: they set the status word, and the host runs the whole map
// set again with the paged form (whose page loops keep state beside the
// accumulators, with spills the common case should not pay for).
enum : int { kNearPlain = 0, kNearGeneral = 1, kNearPaged = 2 };
template <int NPLAIN, bool CUBE, int FORM>
__global__ void __launch_bounds__(kNearThreads) __attribute__((amdgpu_waves_per_eu(5))) k_near(NearParams p) {
    constexpr bool PAGED = FORM == kNearPaged;
    constexpr bool GENERAL = FORM != kNearPlain;
    __shared__ int cand[kCandCap];
    __shared__ int hcand[kHCap];
    __shared__ double hpos[kHCap][3];
    __shared__ int hidx[kHCap];
    __shared__ float hrvs[kHCap][3];
    __shared__ float2 rs_s[kNearRepsLen];   // (electrostatic factor, sol_fn) per bin
    __shared__ int wave_tot[kNearThreads / 64];
    __shared__ int cell_off[kNearThreads + 1];
    __shared__ int cstart_s[kNearThreads];
    __shared__ int ncand_s, nh_s;

    const int tid = threadIdx.x;
    for (int k = tid; k < kNearRepsLen; k += kNearThreads) rs_s[k] = make_float2(p.reps[k], p.sol[min(k, AG_NEINT)]);
    int lx, ly, lz;
    brick_lane<CUBE>(tid, lx, ly, lz);

    // one workgroup per brick of all conformations, in chunks of 16 bricks
    // spread over the XCDs: a brick's time follows the local atom density,
    // and contiguous eighths of the grid per XCD left the XCDs that hold a
    // protein's dense core finishing last.  (A persistent loop
    // over a work counter keeps more uniform state live than the scalar
    // register file holds and spills; the far kernel, whose state is
    // small, is the persistent one.)
    const int item = p.item_base + xcd_chunked<16>(blockIdx.x, gridDim.x);
    if (item >= p.total_bricks) return;
    const int ci = find_conf(as_const(p.brick_prefix), p.nconf, item);
    const AG_CONST ConfDev &cf = as_const(p.confs)[ci];
    const int brick = item - as_const(p.brick_prefix)[ci];
    const int n1x = cf.n1[0], n1y = cf.n1[1], n1z = cf.n1[2];
    const int nex = cf.ne[0], ney = cf.ne[1], nez = cf.ne[2];
    const double spacing = cf.spacing;
    const float s128 = cf.s128;
    const int bx = brick % cf.nb[0];
    const int by = (brick / cf.nb[0]) % cf.nb[1];
    const int bz = brick / (cf.nb[0] * cf.nb[1]);
    const int ix = bx * kBX + lx;
    const int iy = by * kBY + ly;
    const int iz = bz * kBZ + lz;

    // brick box in the grid frame
    const int ix0 = bx * kBX, iy0 = by * kBY, iz0 = bz * kBZ;
    const int ix1 = min(ix0 + kBX, n1x) - 1, iy1 = min(iy0 + kBY, n1y) - 1, iz1 = min(iz0 + kBZ, n1z) - 1;
    const double bmin[3] = {(ix0 - nex) * spacing, (iy0 - ney) * spacing, (iz0 - nez) * spacing};
    const double bmax[3] = {(ix1 - nex) * spacing, (iy1 - ney) * spacing, (iz1 - nez) * spacing};

    // point, reference arithmetic: c = (double)(i - ne) * spacing
    const double cx = (double)(ix - nex) * spacing;
    const double cy = (double)(iy - ney) * spacing;
    const double cz = (double)(iz - nez) * spacing;
    // (the far kernel's coordinates of the point and the tree build's FP32
    // query point are formed where they are used, in rare branches of the
    // candidate loop, so they hold no registers across it)
    const int bh = GENERAL ? p.bhtree : 0;
    const float ccf = (float)p.ccut, ccf2 = ccf * ccf;
    bool collide = false;

    // ---- stage candidate atoms: cells overlapping the brick box + radius ----
    int c0[3], c1[3];
    for (int d = 0; d < 3; d++) {
        c0[d] = (int)floor((bmin[d] - kNearRadius - cf.cell_origin[d]) / p.cell_size);
        c1[d] = (int)floor((bmax[d] + kNearRadius - cf.cell_origin[d]) / p.cell_size);
        c0[d] = max(c0[d], 0);
        c1[d] = min(c1[d], cf.nc[d] - 1);
    }
    const int ncx = c1[0] - c0[0] + 1, ncy = c1[1] - c0[1] + 1, ncz = c1[2] - c0[2] + 1;
    const int ncell = (ncx > 0 && ncy > 0 && ncz > 0) ? ncx * ncy * ncz : 0;
    // Staging pass: the brick's candidates (atoms within 8.03 A of the box),
    // in cell order and slot order within a cell, are numbered from 0 as
    // found, and the donor H among them likewise.  Candidates numbered
    // cpage .. cpage + kCandCap - 1 go to cand[] and donor H numbered hpage ..
    // hpage + kHCap - 1 to hcand[] (a negative page stores none); ncand_s
    // and nh_s end as the totals.  A brick with more than fits is processed
    // page by page, each page staged by another pass, in the same order, so
    // its sums are those one unbounded list would give.
    const float r2cut = kNearRadius * kNearRadius;
    auto stage = [&](int cpage, int hpage) {
        if (tid == 0) { ncand_s = 0; nh_s = 0; }
        __syncthreads();
        for (int cbase = 0; cbase < ncell; cbase += kNearThreads) {
            const int ci = cbase + tid;
            int cstart = 0, ccount = 0;
            if (ci < ncell) {
                const int gx = c0[0] + ci % ncx, gy = c0[1] + (ci / ncx) % ncy, gz = c0[2] + ci / (ncx * ncy);
                const int g = cf.cell_base + gx + cf.nc[0] * (gy + cf.nc[1] * gz);
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
                const int cj = base + pos - cpage;
                if (cpage >= 0 && ok && cj >= 0 && cj < kCandCap) cand[cj] = slot * (int)sizeof(NearAtom);
                int htot;
                const int hpos_i = block_compact(is_h, wave_tot, htot);
                const int hbase = nh_s;
                const int hj = hbase + hpos_i - hpage;
                if (hpage >= 0 && is_h && hj >= 0 && hj < kHCap) hcand[hj] = slot;
                __syncthreads();
                if (tid == 0) { ncand_s = base + tot; nh_s = hbase + htot; }
                __syncthreads();
            }
        }
    };
    stage(0, 0);
    if constexpr (!PAGED) {
        if (ncand_s > kCandCap || nh_s > kHCap) {   // uniform per workgroup
            if (tid == 0) atomicOr(p.status, 1);
            return;
        }
    }

    const int ncand = ncand_s, nh = nh_s;

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
    // The reference compares r = sqrt(d2), which is correctly rounded and so
    // monotone in d2: squared distances more than 2^-49 apart (relative)
    // order the same way, strictly.  Within that band the two rounded
    // distances may tie, so there they are formed and compared as the
    // reference does.  The start, 999999^2, is exact and its root is 999999.
    double best_d2 = 999999.0 * 999999.0;
    int best_idx = 0x7fffffff;
    float hrv[3] = {0.f, 0.f, 0.f};   // bond vector of the closest H so far
    for (int h0 = 0; h0 < nh; h0 += kHCap) {   // pages of donor H
        if constexpr (PAGED) {
            if (h0 > 0) {
                __syncthreads();
                stage(-1, h0);
            }
        }
        const int hn = min(nh - h0, kHCap);
        for (int h = tid; h < hn; h += kNearThreads) {
            const NearAtom &a = p.atoms[hcand[h]];
            hpos[h][0] = a.x; hpos[h][1] = a.y; hpos[h][2] = a.z;
            hidx[h] = a.index;
            hrvs[h][0] = a.rv[0]; hrvs[h][1] = a.rv[1]; hrvs[h][2] = a.rv[2];
        }
        __syncthreads();
        for (int h = 0; h < hn; h++) {
#pragma clang fp contract(off)
            const double dx = hpos[h][0] - cx, dy = hpos[h][1] - cy, dz = hpos[h][2] - cz;
            const double d2 = dx * dx + dy * dy + dz * dz;   // as ref_dist forms it
            const int ix_h = hidx[h];
            bool take;
            if (d2 < best_d2 * (1.0 - 0x1p-49)) {
                take = true;
            } else if (d2 > best_d2 * (1.0 + 0x1p-49)) {
                take = false;
            } else {
                const double r = sqrt(d2), rb = sqrt(best_d2);
                take = r < rb || (r == rb && ix_h < best_idx);
            }
            if (take) {
                best_d2 = d2;
                best_idx = ix_h;
                hrv[0] = hrvs[h][0]; hrv[1] = hrvs[h][1]; hrv[2] = hrvs[h][2];
            }
        }
        if constexpr (!PAGED) break;
    }

    // The point indices are formed again from the thread index where they
    // are needed after this point (a rare branch of the candidate loop and
    // the stores), so three indices do not stay live across the loop: one
    // register instead of three, and no spill.
    const auto point_index = [&](int &qx, int &qy, int &qz) {
        int t = tid;
        asm volatile("" : "+v"(t));   // opaque at each use: not hoisted out of the loop
        brick_lane<CUBE>(t, qx, qy, qz);
        qx += bx * kBX;
        qy += by * kBY;
        qz += bz * kBZ;
    };
    float e_near = 0.f, d_acc = 0.f, dv_acc = 0.f, dq_acc = 0.f;
    int ccount = 0;
    const bool want_cmap = GENERAL && p.cmap != nullptr;
    const float *tab = p.tab;
    const bool want_e = p.e_slot >= 0;
    // cos and sin of 1.24791 rad, the disordered-hydroxyl offset angle
    const float cdis = 0.31936680f, sdis = 0.94763168f;

    // H-bond codes of the pass, one word per atom class (lane hb holds
    // hcode[k][hb] in byte k), read with readlane instead of a load per atom.
    // Cross-lane reads (this one and the candidate index) happen only at the
    // top of a candidate iteration, where every lane of the wave is active:
    // a register reloaded from a spill under a partial exec mask would hold
    // stale values in the inactive lanes.
    const int wlane = tid & 63;
    const int hcw = wlane < 7 ? (p.hcode[0][wlane] | p.hcode[1][wlane] << 8 | p.hcode[2][wlane] << 16 |
                                 p.hcode[3][wlane] << 24)
                              : 0;
    for (int cpage = 0; cpage < ncand; cpage += kCandCap) {   // pages of candidates
        if constexpr (PAGED) {
            if (cpage > 0) {
                __syncthreads();
                stage(cpage, -1);
            }
        }
        const int cn = min(ncand - cpage, kCandCap);
        // candidate records go straight into scalar registers through the
        // constant address space (the atom array is read-only here).  The
        // wave holds 64 candidate indices in one register and loads the
        // next record's coordinates, class and index one candidate ahead,
        // which also brings both of its cache lines into the scalar cache.
        // The list holds byte offsets of the records (below 4 GiB, checked
        // when the plan is built), so a record's address is the array base
        // plus an unsigned 32-bit scalar, which the scalar loads take as
        // their offset operand: no 64-bit multiply-add per record.
        const AG_CONST char *cbase = reinterpret_cast<const AG_CONST char *>(as_const(p.atoms));
        const auto rec = [&](unsigned off) {
            return reinterpret_cast<const AG_CONST NearAtom *>(cbase + off);
        };
        for (int j0 = 0; j0 < cn; j0 += 64) {
            const int vcand = cand[min(j0 + wlane, cn - 1)];
            const int jn = min(64, cn - j0);
            unsigned off = (unsigned)__builtin_amdgcn_readlane(vcand, 0);
            double nx = rec(off)->x, ny = rec(off)->y, nz = rec(off)->z;
            int nmeta = rec(off)->meta, nindex = rec(off)->index;
            for (int jj = 0; jj < jn; jj++) {
                // this record's address, formed once: its fields load from it
                // with immediate offsets in whichever branch needs them
                const AG_CONST NearAtom *ap = rec(off);
                asm volatile("" : "+s"(ap));
                const AG_CONST NearAtom &a = *ap;
                const double ax = nx, ay = ny, az = nz;
                const int ameta = nmeta, aindex = nindex;
                const unsigned hw = (unsigned)__builtin_amdgcn_readlane(hcw, (ameta >> 8) & 63);
                unsigned noff = (unsigned)__builtin_amdgcn_readlane(vcand, min(jj + 1, jn - 1));
                // the next loads may issue only once this record has arrived:
                // scalar loads complete out of order, so the wait for this
                // record would otherwise also wait for the next one
                asm volatile("" : "+s"(noff) : "s"(ameta), "s"(aindex), "s"(ax), "s"(ay), "s"(az));
                nx = rec(noff)->x;
                ny = rec(noff)->y;
                nz = rec(noff)->z;
                nmeta = rec(noff)->meta;
                nindex = rec(noff)->index;
                off = noff;   // a still names this record
                // geometry: FP64 difference as the reference forms it, FP32 after
                const double dx64 = ax - cx, dy64 = ay - cy, dz64 = az - cz;
                const float dxf = (float)dx64, dyf = (float)dy64, dzf = (float)dz64;
                const float r2 = __builtin_fmaf(dxf, dxf, __builtin_fmaf(dyf, dyf, dzf * dzf));
                if (r2 > 64.33f) continue;  // beyond 8.02 A: neither cutoff pass nor near e term
                // hardware rsq: 128 r is good to ~3e-4 bins here, inside the guard band;
                // the same rsq is 1/r for the near e term and the unit vector
                const float r2c = fmaxf(r2, 1.0e-12f);
                float inv_r = __builtin_amdgcn_rsqf(r2c);
                float r = r2c * inv_r;
                const float g = r * 128.f;
                int bin = (int)g;
                bool cut = r > 8.0f;
                if (fabsf(g - rintf(g)) < kGuardBins) {
                    double rdx, rdy, rdz;
                    const double r64 = ref_dist(ax, ay, az, cx, cy, cz, rdx, rdy, rdz);
                    bin = (int)(r64 * 128.0);
                    cut = r64 > 8.0;
                    r = (float)r64;
                    inv_r = __builtin_amdgcn_rcpf(fmaxf(r, 1.0e-6f));
                }
                // the tree build's radius queries (bhtree.cpp:341-362): FP32 point
                // and atom, each component within the radius, then the rounded
                // squared distance against the rounded squared radius
                bool in8 = false;
                float rB = 0.f;
                int bin_bh = 0;
                if (bh) {
                    // the tree build's query point: the grid point as float (main.cpp:2860)
                    const float e0 = (float)cx - (float)ax, e1 = (float)cy - (float)ay, e2 = (float)cz - (float)az;
                    const float a0 = fabsf(e0), a1 = fabsf(e1), a2 = fabsf(e2);
                    const float D = (e0 * e0 + e1 * e1) + e2 * e2;
                    in8 = a0 <= 8.f && a1 <= 8.f && a2 <= 8.f && D <= 64.f;
                    if (a0 <= 1.5f && a1 <= 1.5f && a2 <= 1.5f && D <= 2.25f) collide = true;
                    if (want_cmap) ccount += a0 <= ccf && a1 <= ccf && a2 <= ccf && D <= ccf2;
                    rB = sqrtf(D);   // correctly rounded, as the tree's sqrt of a float
                    bin_bh = (int)((double)rB * 128.0);
                }
                // The pair's vdW/H-bond table row for the pass's slots, loaded
                // here so that the loads are in flight during the constriction,
                // electrostatic and directional work below.  The type's rows
                // start at a scalar base, the bin's row is a 32-bit byte offset
                // from it; the bin is the one Loop 2 uses (the tree build's FP32
                // distance in that mode).  Every lane here has r < 8.03 A, so
                // the row index stays inside the table.
                const int indx_n = min(bh ? bin_bh : bin, AG_NEINT - 1);
                float E[kMapSlots];
                {
                    const char *trow =
                        reinterpret_cast<const char *>(tab + (size_t)(ameta & 0xff) * AG_NEINT * kMapSlots);
                    const unsigned roff = (unsigned)indx_n * (unsigned)(kMapSlots * sizeof(float));
    #pragma unroll
                    for (int q4 = 0; q4 < kMapSlots / 4; q4++) {
                        if (q4 < NPLAIN / 4 || q4 == kMapSlots / 4 - 1) {
                            const float4 v = *reinterpret_cast<const float4 *>(trow + roff + 16u * q4);
                            E[4 * q4] = v.x; E[4 * q4 + 1] = v.y; E[4 * q4 + 2] = v.z; E[4 * q4 + 3] = v.w;
                        } else {
                            E[4 * q4] = E[4 * q4 + 1] = E[4 * q4 + 2] = E[4 * q4 + 3] = 0.f;
                        }
                    }
                }
                // electrostatic factor and sol_fn of the bin, one LDS read
                const float2 rsv = rs_s[min(bin, kNearRepsLen - 1)];
                // constriction grid: atoms within the cutoff (main.cpp:2969-2973)
                if (want_cmap && !bh) {
                    if (fabsf(r - (float)p.ccut) < 1.0e-5f) {
                        double rdx, rdy, rdz;
                        ccount += ref_dist(ax, ay, az, cx, cy, cz, rdx, rdy, rdz) <= p.ccut;
                    } else {
                        ccount += r <= (float)p.ccut;
                    }
                }
                // electrostatics of pairs the far kernel leaves out
                if (want_e) {
                    bool near_e = g < 1022.f;
                    if (g >= 1022.f && g <= 1027.f) {
                        // the far kernel's point coordinates: run centre, offset in the run, y, z
                        int qx, qy, qz;
                        point_index(qx, qy, qz);
                        const int kxE = qx % kFarK;
                        near_e = far_bin4(a.xE, a.yE, a.zE, far_cx(qx - kxE, nex, s128), kxE, s128,
                                          (float)(qy - ney) * s128, (float)(qz - nez) * s128, far_bias(s128)) <
                                 4u * kFarBinMin;
                    }
                    // 1/max(r, 0.5) with r clamped at APPROX_ZERO; adding 0 leaves
                    // the sum as it is (it is never -0)
                    const float inv_rmax = fminf(inv_r, 2.f);
                    const float te = unif(a.qe) * inv_rmax * rsv.x;
                    e_near += near_e ? te : 0.f;
                }
                // Loop 2 of the tree build takes the atoms its 8 A query returns,
                // at their FP32 distance (main.cpp:3019-3028)
                if (bh) {
                    cut = !in8;
                    r = rB;
                    bin = bin_bh;
                    inv_r = __builtin_amdgcn_rcpf(fmaxf(rB, 1.0e-6f));
                }
                if (cut || !p.has_affinity) continue;
                const int meta = ameta;
                const int type = meta & 0xff, hb = (meta >> 8) & 0xff, rexp = (meta >> 16) & 0xff;
                const bool dis = GENERAL && ((meta >> 24) & 1);
                if (p.disorder_h && dis && type == p.hydrogen_type) continue;
                float racc = 1.f, rdon = 1.f, Hramp = 1.f;
                // directional classes only (D1, A1, A2): the unit vector to the point
                constexpr unsigned kDirClasses = (1u << AG_HB_D1) | (1u << AG_HB_A1) | (1u << AG_HB_A2);
                if ((kDirClasses >> hb) & 1u) {
                    const float dn0 = dxf * inv_r, dn1 = dyf * inv_r, dn2 = dzf * inv_r;
                    const float rv0 = unif(a.rv[0]), rv1 = unif(a.rv[1]), rv2 = unif(a.rv[2]);
                    const int aidx = aindex;
                    if (hb == AG_HB_D1 || (hb == AG_HB_A2 && !dis)) {
                        // cos(theta) of the D1 / ordered-A2 branches; FP64 near 0
                        float cos_t = -(dn0 * rv0) - dn1 * rv1 - dn2 * rv2;
                        bool pos;
                        if (fabsf(cos_t) < 1.0e-5f) {
                            double rdx, rdy, rdz;
                            double r64 = ref_dist(ax, ay, az, cx, cy, cz, rdx, rdy, rdz);
                            if (bh) r64 = (double)rB;   // the tree build's r (main.cpp:3026)
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
                                const float sin2ti =
                                    2.f * fabsf(ti) * __builtin_amdgcn_sqrtf(fmaxf(1.f - ti * ti, 0.f));
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
                }  // directional classes
                const float S = bh ? rs_s[min(bin, AG_NEINT)].y : rsv.y;   // sol_fn[min(bin, 1024)]
                // desolvation (main.cpp:3375-3390): slot m gains
                // (sp[m] vol + volp[m] solq) S per pair, i.e. sp[m] sum(vol S) +
                // volp[m] sum(solq S) over the point's pairs: two sums per point
                // instead of a term per slot, combined after the loop
                dv_acc = __builtin_fmaf(unif(a.vol), S, dv_acc);
                dq_acc = __builtin_fmaf(unif(a.solq), S, dq_acc);
                // plain slots: the table value
    #pragma unroll
                for (int m = 0; m < NPLAIN; m++) acc[m] += E[m];
                // H-bonding slots (main.cpp:3286-3374): the (probe, atom) case is a
                // set of 0/1 coefficients chosen per atom class, uniform across
                // the wave.  The plain table add is fma(c, E, acc) with c = 1 or
                // 0 (bit for bit acc + E, or acc); the directional mix runs in a
                // scalar branch for the classes that have one, with every slot's
                // update written without a branch of its own, so the
                // accumulators and min/max keep their registers through it.  A
                // disordered partner uses the hydrogen-type row at r - 1.10 A
                // (main.cpp:3313-3349), FP64 like the reference.
                const int dsh = dis ? 4 : 0;
                bool any_mix = false;
    #pragma unroll
                for (int k = 0; k < kHbSlots; k++) {
                    const int code = (hw >> (8 * k + dsh)) & 15;
                    any_mix |= (code & ~HC_PLAIN) != 0;
                    const float kP = __int_as_float((code & HC_PLAIN) ? 0x3f800000 : 0);
                    acc[kPlainSlots + k] = __builtin_fmaf(kP, E[kPlainSlots + k], acc[kPlainSlots + k]);
                }
                if (any_mix) {
                    float Eh[kHbSlots] = {0.f, 0.f, 0.f, 0.f};
                    if (dis && p.hydrogen_type >= 0) {
                        double rdx, rdy, rdz;
                        const double r64 = bh ? (double)rB : ref_dist(ax, ay, az, cx, cy, cz, rdx, rdy, rdz);
                        const int hb_bin = min(max(0, (int)((r64 - 1.10) * 128.0)), AG_NEINT - 1);
                        const float4 v = reinterpret_cast<const float4 *>(
                            tab + ((size_t)p.hydrogen_type * AG_NEINT + hb_bin) * kMapSlots)[kPlainSlots / 4];
                        Eh[0] = v.x; Eh[1] = v.y; Eh[2] = v.z; Eh[3] = v.w;
                    }
                    const float dfac = racc - rdon;
    #pragma unroll
                    for (int k = 0; k < kHbSlots; k++) {
                        const int m = kPlainSlots + k;
                        const int code = (hw >> (8 * k + dsh)) & 15;
                        const float e = E[m];
                        // 0/1 coefficients as scalar float bit patterns
                        const float kF = __int_as_float((code & HC_RACC) ? 0x3f800000 : 0);
                        const float kH = __int_as_float((code & HC_HRAMP) ? 0x3f800000 : 0);
                        const bool mm = (code & HC_MINMAX) != 0;
                        const float rsph = __builtin_amdgcn_fmed3f(e * 0.01f, 0.f, 1.f);
                        const float f = __builtin_fmaf(kF, dfac, rdon);
                        const float mix = __builtin_fmaf(1.f - f, rsph, f);
                        acc[m] = __builtin_fmaf((kH * (dis ? Eh[k] : e)) * Hramp, mix, acc[m]);
                        const float v = e * mix;
                        hmin[k] = fminf(hmin[k], mm ? v : __builtin_inff());
                        hmax[k] = fmaxf(hmax[k], mm ? v : -__builtin_inff());
                    }
                }
                if (p.d_slot >= 0) d_acc += unif(a.dvol) * S;
            }
        }
        if constexpr (!PAGED) break;
    }

    int qx, qy, qz;
    point_index(qx, qy, qz);
    if (qx < n1x && qy < n1y && qz < n1z) {
        // output pointers are formed here, after the candidate loop, so they
        // hold no registers during it
        int *const cmap = p.cmap ? p.cmap + cf.pt_off : nullptr;
        float *const maps = p.maps + cf.map_off;
        const long long P = cf.P;
        const long long pt = (long long)qx + (long long)n1x * ((long long)qy + (long long)n1y * qz);
#pragma unroll
        for (int m = 0; m < kMapSlots; m++) {
            if ((m >= NPLAIN && m < kPlainSlots) || p.out_slot[m] < 0) continue;
            float v = acc[m] + __builtin_fmaf(p.sp[m], dv_acc, p.volp[m] * dq_acc);
            if (m >= kPlainSlots && hmin[m - kPlainSlots] < 999999.f)
                v += hmin[m - kPlainSlots] + fmaxf(hmax[m - kPlainSlots], 0.f);
            maps[(long long)p.out_slot[m] * P + pt] = v;
        }
        if (want_e) maps[(long long)p.e_slot * P + pt] = e_near;
        if (cmap) cmap[pt] = ccount;
        if (p.collide) p.collide[cf.pt_off + pt] = collide ? 1 : 0;
        if (p.d_slot >= 0) maps[(long long)p.d_slot * P + pt] = d_acc;
    }
}

// The tree build's collision rule (main.cpp:2906-2927): a point within
// 1.5 A of a receptor atom has e = 0 and floating value 1.5 (Loop 1 is
// skipped there).  Runs after the near, far and floating-grid kernels.
__global__ void k_collide(long long P, const unsigned char *mask, float *emap, float *fmap) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P || !mask[i]) return;
    if (emap) emap[i] = 0.f;
    if (fmap) fmap[i] = 1.5f;
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
        const int kx = ix % kFarK;
        const float cxE = far_cx(ix - kx, p.ne[0], p.s128);
        const float pyE = (float)(iy - p.ne[1]) * p.s128, pzE = (float)(iz - p.ne[2]) * p.s128;
        const float biasE = far_bias(p.s128);
        for (int a = 0; a < p.n_atoms; a++) {
            const float4 at = p.far_atoms[a];
            const int fb = (int)(far_bin4(at.x, at.y, at.z, cxE, kx, p.s128, pyE, pzE, biasE) >> 2);
            double dx, dy, dz;
            const double r64 = ref_dist(p.coord[3 * a], p.coord[3 * a + 1], p.coord[3 * a + 2], cx, cy, cz, dx, dy, dz);
            const int b64 = min((int)(r64 * 128.0), AG_NDIEL - 1);
            if (fb >= kFarBinMin) {
                far_pairs++;
                far_flips += min(fb, AG_NDIEL - 1) != b64;
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

// Bins at and beyond kFarLastBin all read r_eps[8191]: the reference clamps
// the index there (main.cpp:2956).  An atom at least kFarConstBins from
// every point of a wave's box reaches only such bins (FP32 rounding moves a
// bin by ~1e-3, the margin is one bin), so its pairs need no table.
constexpr int kFarLastBin = AG_NDIEL - 1;
constexpr float kFarConstBins = (float)(AG_NDIEL + 1);

// One pair of an atom with the 16 points of a lane's run, accumulated into
// part[]: the squared distance in bins^2 by the run-centre expansion, rsq,
// and either the table (bin offset in one fma and one and, the LDS read;
// CLAMP caps the offset at the table's last entry) or, for an atom beyond
// the table's last bin from the whole wave (TABLE false), the constant
// r_eps[8191] folded into the atom's weight.
template <bool TABLE, bool CLAMP>
__device__ __forceinline__ void far_pair(const PlainF4 &at, float wc, float py, float pz, float cx, float bias,
                                         float s128, float s2, const float *T, unsigned tmax4, float *part) {
    const float dy = at.y - py, dz = at.z - pz;
    const float dyz2 = __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, bias));
    const float u = at.x - cx;
    const float U = __builtin_fmaf(u, u, dyz2);
    const float us = u * s128;
    float W[kFarK / 2];
#pragma unroll
    for (int j = 0; j < kFarK / 2; j++) W[j] = __builtin_fmaf(s2, far_c2(kFarK / 2 + j), U);
    float rs[kFarK];
    unsigned off[kFarK];
#pragma unroll
    for (int k = 0; k < kFarK; k++) {
        const int j = k >= kFarK / 2 ? k - kFarK / 2 : kFarK / 2 - 1 - k;
        const float rb2 = __builtin_fmaf(us, far_m(k), W[j]);
        rs[k] = __builtin_amdgcn_rsqf(rb2);
        if (TABLE) {
            off[k] = bin4_of(rb2, rs[k]);
            if (CLAMP) off[k] = min(off[k], tmax4);
        }
    }
    if (TABLE) {
        float tv[kFarK];
#pragma unroll
        for (int k = 0; k < kFarK; k++)
            tv[k] = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(T) + off[k]);
#pragma unroll
        for (int k = 0; k < kFarK; k++) part[k] = __builtin_fmaf(at.w, rs[k] * tv[k], part[k]);
    } else {
#pragma unroll
        for (int k = 0; k < kFarK; k++) part[k] = __builtin_fmaf(wc, rs[k], part[k]);
    }
}

// A wave-uniform float moved to a scalar register.
__device__ __forceinline__ float sreg(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}


__device__ __forceinline__ int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// CAP: LDS table capacity in entries (>= tab_len).  The table is the first
// thing in the kernel's only LDS array, so it starts at address 0 and a
// bin's byte offset is its LDS address; after it one int broadcasts the work
// counter and kFarRuns * kFarK floats carry the second half's sums.  The
// table runs to the largest bin a pair can reach (at most kFarTabMax
// entries), every entry past 8191 holding the reference's last value.
//
// An item is kFarRuns runs; the workgroup's two halves (four waves each)
// take the same runs, the first half the atoms [0, a_mid), the second
// [a_mid, n), a_mid = ceil(n / 2).  The second half leaves its
// compensated sums in LDS and the first adds them in a fixed order, so the
// result does not depend on scheduling.  Splitting the atoms keeps four
// waves per SIMD on grids that give only two per SIMD at one wave per run.
// Run order: a wave takes one x-run of a patch of 64 rows (8 x 8 in y and
// z, far_row), so its 64 runs lie in a small box; the patch's x-runs go to
// consecutive waves, so the four waves of a half cover consecutive 64-byte
// pieces of the same 64 rows, and their loads and stores of the e map meet
// in the same cache lines at about the same time.
//
// Per conformation (ConfDev::far_const, so a conformation's maps are the
// same in every plan; MODE kFarTable and kFarTfree serve launches whose
// items all take one path, kFarMixed decides per item, uniformly per
// workgroup, at the price of more spills around the item loop): with few
// pairs past bin 8192, every atom goes through the table, in index order,
// one atom ahead, and the table covers every bin.  For large boxes, for each block of 64 atoms the wave tests every
// atom's distance to its box at once (one atom per lane, one ballot).
// Atoms at least kFarConstBins from the whole box reach only bins past
// 8191, whose value is r_eps[8191], and take the table-free path; within a
// block they go first, then the table's atoms, each in index order (the
// order depends only on the geometry, so results repeat).  When the
// conformation's pairs may pass the table's end (ConfDev::far_clamp), a
// wave whose box could carry a table-path pair there caps its table
// indices; past bin 8191 every entry is r_eps[8191], so the cap changes no
// value.  Lanes past the conformation's runs
// compute at a point of their wave's box and store nothing.
template <int CAP, int MODE>
__global__ void __launch_bounds__(kFarThreads) __attribute__((amdgpu_waves_per_eu(4))) k_far(FarParams p) {
    __shared__ float T[CAP + 1 + kFarRuns * kFarK];
    int &item_s = reinterpret_cast<int *>(T)[CAP];
    float *const R = T + CAP + 1;
    for (int i = threadIdx.x; i < p.tab_len; i += kFarThreads) T[i] = p.tab[i];
    const int half = __builtin_amdgcn_readfirstlane(threadIdx.x / kFarRuns);   // uniform per wave
    const int lane = threadIdx.x % kFarRuns;
    const int lane64 = threadIdx.x % 64;
    int item;
    if (p.queue) {
        if (threadIdx.x == 0) item_s = p.item_base + atomicAdd(p.queue, 1);
        __syncthreads();
        item = item_s;
    } else {
        __syncthreads();
        item = p.item_base + xcd_chunked<16>(blockIdx.x, gridDim.x);
    }
    item = __builtin_amdgcn_readfirstlane(item);
    while (item < p.item_end) {
        const int ci = find_conf(as_const(p.tile_prefix), p.nconf, item);
        const AG_CONST ConfDev &cf = as_const(p.confs)[ci];
        const int n1x = cf.n1[0], n1y = cf.n1[1];
        // wave slot (uniform): row patch pr, x-run xr; lane: row pr * 64 + l
        const int ws = __builtin_amdgcn_readfirstlane((item - as_const(p.tile_prefix)[ci]) * (kFarRuns / 64) +
                                                      lane / 64);
        const int pr = ws / cf.runs_per_row;
        const int q = pr * 64 + lane % 64;
        const bool active = q < n1y * cf.n1[2];
        int ix0 = 0, iy = 0, iz = 0;
        if (active) {
            ix0 = (ws - pr * cf.runs_per_row) * kFarK;
            far_row(q, n1y, cf.n1[2], iy, iz);
        }
        // the wave's box of points, in grid indices
        const int bx0 = __builtin_amdgcn_readfirstlane(wave_min(active ? ix0 : INT_MAX));
        const int bx1 = __builtin_amdgcn_readfirstlane(wave_max(active ? ix0 : INT_MIN));
        const int by0 = __builtin_amdgcn_readfirstlane(wave_min(active ? iy : INT_MAX));
        const int by1 = __builtin_amdgcn_readfirstlane(wave_max(active ? iy : INT_MIN));
        const int bz0 = __builtin_amdgcn_readfirstlane(wave_min(active ? iz : INT_MAX));
        const int bz1 = __builtin_amdgcn_readfirstlane(wave_max(active ? iz : INT_MIN));
        float *const emap = p.maps + cf.map_off + (long long)p.e_slot * cf.P;
        float s[kFarK], c[kFarK];
        if (bx0 <= bx1) {   // the wave has runs (uniform)
            if (!active) {  // compute at a point of the box
                ix0 = bx0;
                iy = by0;
                iz = bz0;
            }
            const float s128 = cf.s128;
            const float py = (float)(iy - cf.ne[1]) * s128;
            const float pz = (float)(iz - cf.ne[2]) * s128;
            const float cx = far_cx(ix0, cf.ne[0], s128);
            const float bias = far_bias(s128);
            const float s2 = s128 * s128;
            const long long base = (long long)ix0 + (long long)n1x * ((long long)iy + (long long)n1y * iz);
#pragma unroll
            for (int k = 0; k < kFarK; k++) {
                s[k] = (half == 0 && active && ix0 + k < n1x) ? emap[base + k] : 0.f;
                c[k] = 0.f;
            }
            // atoms are read-only here: scalar loads, one atom ahead
            const float4 *const vatoms = p.atoms + cf.atom_base;
            const AG_CONST PlainF4 *atoms = as_const(reinterpret_cast<const PlainF4 *>(vatoms));
            const int n_atoms = cf.n_atoms;
            const int a_mid = (n_atoms + 1) / 2;   // Kahan blocks of each half start at its first atom
            const int a_beg = half ? a_mid : 0;
            const int a_end = half ? n_atoms : a_mid;
            const bool tfree = MODE == kFarMixed ? cf.far_const != 0 : MODE == kFarTfree;
            if (!tfree) {   // uniform per workgroup
                PlainF4 nxt = ld4(atoms, min(a_beg, max(a_end - 1, 0)));
                for (int a0 = a_beg; a0 < a_end; a0 += kFarAtomChunk) {
                    const int a1 = min(a0 + kFarAtomChunk, a_end);
                    float part[kFarK];
#pragma unroll
                    for (int k = 0; k < kFarK; k++) part[k] = 0.f;
                    for (int a = a0; a < a1; a++) {
                        const PlainF4 at = nxt;
                        nxt = ld4(atoms, min(a + 1, a_end - 1));
                        far_pair<true, false>(at, 0.f, py, pz, cx, bias, s128, s2, T, 0u, part);
                    }
#pragma unroll
                    for (int k = 0; k < kFarK; k++) {
                        const float y = part[k] - c[k];
                        const float t = s[k] + y;
                        c[k] = (t - s[k]) - y;
                        s[k] = t;
                    }
                }
            } else {
                // box corners in the far kernel's frame (bins); a run spans
                // kFarK points in x.  Uniform: held in scalar registers,
                // which the atom test reads directly.
                const float xlo = sreg((float)(bx0 - cf.ne[0]) * s128);
                const float xhi = sreg((float)(bx1 + kFarK - 1 - cf.ne[0]) * s128);
                const float ylo = sreg((float)(by0 - cf.ne[1]) * s128), yhi = sreg((float)(by1 - cf.ne[1]) * s128);
                const float zlo = sreg((float)(bz0 - cf.ne[2]) * s128), zhi = sreg((float)(bz1 - cf.ne[2]) * s128);
                const float thr2 = kFarConstBins * kFarConstBins;
                // a table-path atom is within kFarConstBins of the box, so its
                // pairs stay below that plus the box diagonal
                const float diag = __builtin_sqrtf((xhi - xlo) * (xhi - xlo) + (yhi - ylo) * (yhi - ylo) +
                                                   (zhi - zlo) * (zhi - zlo));
                const bool clamp = cf.far_clamp && kFarConstBins + diag + 4.f >= (float)(p.tab_len - 1);
                const float c_last = p.c_last;
                for (int a0 = a_beg; a0 < a_end; a0 += kFarAtomChunk) {
                    const int a1 = min(a0 + kFarAtomChunk, a_end);
                    float part[kFarK];
#pragma unroll
                    for (int k = 0; k < kFarK; k++) part[k] = 0.f;
                    for (int b0 = a0; b0 < a1; b0 += 64) {
                        const int nb = min(64, a1 - b0);
                        bool beyond = false;
                        if (lane64 < nb) {
                            const float4 A = vatoms[b0 + lane64];
                            const float ex = fmaxf(fmaxf(xlo - A.x, A.x - xhi), 0.f);
                            const float ey = fmaxf(fmaxf(ylo - A.y, A.y - yhi), 0.f);
                            const float ez = fmaxf(fmaxf(zlo - A.z, A.z - zhi), 0.f);
                            beyond = __builtin_fmaf(ex, ex, __builtin_fmaf(ey, ey, ez * ez)) >= thr2;
                        }
                        const unsigned long long all = nb == 64 ? ~0ull : ((1ull << nb) - 1);
                        unsigned long long far_set = __ballot(beyond) & all;
                        unsigned long long tab_set = ~far_set & all;
                        if (far_set) {
                            PlainF4 nxt = ld4(atoms, b0 + __builtin_ctzll(far_set));
                            far_set &= far_set - 1;
                            while (true) {
                                const PlainF4 at = nxt;
                                const bool more = far_set != 0;
                                if (more) {
                                    nxt = ld4(atoms, b0 + __builtin_ctzll(far_set));
                                    far_set &= far_set - 1;
                                }
                                far_pair<false, false>(at, at.w * c_last, py, pz, cx, bias, s128, s2, T, 0u, part);
                                if (!more) break;
                            }
                        }
                        if (tab_set) {
                            PlainF4 nxt = ld4(atoms, b0 + __builtin_ctzll(tab_set));
                            tab_set &= tab_set - 1;
                            while (true) {
                                const PlainF4 at = nxt;
                                const bool more = tab_set != 0;
                                if (more) {
                                    nxt = ld4(atoms, b0 + __builtin_ctzll(tab_set));
                                    tab_set &= tab_set - 1;
                                }
                                if (clamp) far_pair<true, true>(at, 0.f, py, pz, cx, bias, s128, s2, T, p.tmax4, part);
                                else far_pair<true, false>(at, 0.f, py, pz, cx, bias, s128, s2, T, 0u, part);
                                if (!more) break;
                            }
                        }
                    }
#pragma unroll
                    for (int k = 0; k < kFarK; k++) {
                        const float y = part[k] - c[k];
                        const float t = s[k] + y;
                        c[k] = (t - s[k]) - y;
                        s[k] = t;
                    }
                }
            }
            if (half && active) {
#pragma unroll
                for (int k = 0; k < kFarK; k++) R[k * kFarRuns + lane] = s[k] - c[k];
            }
        }
        __syncthreads();
        if (half == 0 && active) {
            const long long base = (long long)ix0 + (long long)n1x * ((long long)iy + (long long)n1y * iz);
#pragma unroll
            for (int k = 0; k < kFarK; k++) {
                const float y = R[k * kFarRuns + lane] - c[k];
                const float t = s[k] + y;
                if (ix0 + k < n1x) emap[base + k] = t;
            }
        }
        if (!p.queue) break;
        __syncthreads();
        if (threadIdx.x == 0) item_s = p.item_base + atomicAdd(p.queue, 1);
        __syncthreads();
        item = __builtin_amdgcn_readfirstlane(item_s);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// context, plans and host orchestration
// ---------------------------------------------------------------------------


// Near-kernel layout chosen by the autotuner.
struct Tune {
    double cell_size = 4.0;     // cell edge of the cell list (A)
    bool cube = true;           // lanes of a wave on a 4x4x4 cube instead of an 8x8x1 slab
                                // (the cube measures faster on the 5,000-atom, 127^3 map set)
};


// Receptor conformations (one for a plan, many for a batch) on their grids,
// with a shared map set and tables, resident on one device.
struct ag_plan {
    explicit ag_plan(std::shared_ptr<DevPool> pl) : pool(std::move(pl)) {}
    std::shared_ptr<DevPool> pool;     // first: the buffers below draw on it
    ag_ctx *ctx = nullptr;
    int nconf = 0, N = 0, M = 0, T = 0, n_maps = 0, groups = 0;
    bool vina = false, persistent = false;
    long long P_total = 0;
    std::vector<ConfDev> confs;
    std::vector<int> brick_prefix, tile_prefix;
    int total_cells = 0, total_bricks = 0, total_tiles = 0;
    Tune tune;
    double tune_ms = 0.0;              // time the autotuner spent
    DevBuf confs_d{pool}, brick_prefix_d{pool}, tile_prefix_d{pool}, atom_conf_d{pool}, queue_d{pool};
    DevBuf coord{pool}, charge{pool}, vol{pool}, solpar{pool}, rvec{pool}, rvec2{pool}, type{pool}, hbond{pool},
        rexp{pool}, disorder{pool};
    DevBuf atom_cell{pool}, cell_count{pool}, cell_start{pool}, cell_cursor{pool}, slot_atom{pool},
        near_atoms{pool}, far_atoms{pool}, fmin_atoms{pool};
    DevBuf tab{pool}, sol{pool}, reps{pool}, fartab{pool}, status{pool};
    DevBuf collide{pool};              // tree build: collision mask per point
    bool want_collide = false;
    int e_index = -1;                  // e map slot, -1 if none
    AtomSrc src{};
    std::vector<NearParams> near;      // one per pass of 16 map slots
    int nplain = 12;                   // plain slots in use (4, 8 or 12)
    std::vector<int> cov_slots;
    double cov_pos[3] = {0, 0, 0}, cov_half_width = 1.0, cov_barrier = 0.0;
    bool want_fmap = false, want_cmap = false;
    std::vector<int> zero_slots;       // output maps that are all zero (Vina e/d)
    FarParams far{};
    bool run_far = false;
    bool far_atoms_ready = false;
    double coeff_estat = 0.0;
    int far_grid = 0;                  // persistent far-kernel grid
    bool paged = false;                // near kernel in its paged form (a brick overflowed one LDS list)
    bool any_disorder = false;         // some atom flagged as a disordered hydroxyl
    // captured map set (persistent mode): the cell list and near field, then
    // the far field of conformations [0, far_split) and of [far_split,
    // nconf), as graphs launched back to back (the second far graph is
    // empty when far_split == nconf)
    hipGraphExec_t graph = nullptr, graph_far = nullptr, graph_far2 = nullptr;
    int far_split = 0;
    float *g_maps = nullptr, *g_fmap = nullptr;
    int32_t *g_cmap = nullptr;
    void drop_graphs() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (graph_far) (void)hipGraphExecDestroy(graph_far);
        if (graph_far2) (void)hipGraphExecDestroy(graph_far2);
        graph = graph_far = graph_far2 = nullptr;
    }
    ~ag_plan() { drop_graphs(); }
};

// The far-kernel mode of conformations [c0, c1): one path for all, or mixed.
static int far_mode(const ag_plan &p, int c0, int c1) {
    int n_tfree = 0;
    for (int ci = c0; ci < c1; ci++) n_tfree += p.confs[ci].far_const != 0;
    return n_tfree == 0 ? kFarTable : n_tfree == c1 - c0 ? kFarTfree : kFarMixed;
}

namespace agdev {

// The device address of `scratch` on `device` (current), looked up once
// per device and kept: every GPU holds its own copy of the array.
static void *scratch_address(int device) {
    static std::mutex mu;
    static std::vector<void *> addr;
    std::lock_guard<std::mutex> lk(mu);
    if ((int)addr.size() <= device) addr.resize(device + 1, nullptr);
    if (!addr[device]) {
        void *p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        addr[device] = p;
    }
    return addr[device];
}

ag_status context_create(ag_ctx **out, int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        ag::fail(AG_ERR_DEVICE, "no HIP device is available (this engine has no CPU fallback)");
    if (device < 0 || device >= n) ag::fail(AG_ERR_DEVICE, "HIP device index out of range");
    HIP_CHECK(hipSetDevice(device));
    ag_ctx *c = new ag_ctx();
    c->device = device;
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_CHECK(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreateWithFlags(&c->cells_ev, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&c->near_ev, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&c->far_ev, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&c->far2_ev, hipEventDisableTiming));
    for (auto &e : c->ev) HIP_CHECK(hipEventCreate(&e));
    for (auto &e : c->chunk_ev) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_CHECK(hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device));
    c->scratch = scratch_address(device);
    *out = c;
    return AG_OK;
}

void *host_alloc_pinned(ag_ctx *c, size_t bytes) {
    HIP_CHECK(hipSetDevice(c->device));
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}

void host_free_pinned(void *p) {
    if (p) (void)hipHostFree(p);
}

void context_destroy(ag_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    if (c->cells_ev) (void)hipEventDestroy(c->cells_ev);
    if (c->near_ev) (void)hipEventDestroy(c->near_ev);
    if (c->far_ev) (void)hipEventDestroy(c->far_ev);
    if (c->far2_ev) (void)hipEventDestroy(c->far2_ev);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : c->chunk_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
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
    if (N > 0 && (!rec->coord || !rec->charge || !rec->vol || !rec->solpar || !rec->type || !rec->hbond ||
                  !rec->rexp || !rec->rvector || !rec->rvector2 || !rec->disorder))
        ag::fail(AG_ERR_BAD_INPUT, "receptor arrays missing");
    for (int i = 0; i < N; i++) {
        if (rec->type[i] < 0 || rec->type[i] >= T) ag::fail(AG_ERR_BAD_INPUT, "atom type index out of range");
        if (rec->hbond[i] < 0 || rec->hbond[i] > 6) ag::fail(AG_ERR_BAD_INPUT, "atom hbond class out of range");
    }
    if (vina && ms->separate_desolvation)
        ag::fail(AG_ERR_UNSUPPORTED, "Vina-potential maps do not combine with separate_desolvation_maps");
    if (ms->bhtree < 0 || ms->bhtree > 2) ag::fail(AG_ERR_BAD_INPUT, "bhtree mode must be 0, 1 or 2");
    if (vina && ms->bhtree) ag::fail(AG_ERR_UNSUPPORTED, "Vina-potential maps have no tree-build mode");
    if (M > 0 && !vina) {
        bool any_h = false;
        for (int i = 0; i < N && !any_h; i++) any_h = rec->hbond[i] == AG_HB_DS || rec->hbond[i] == AG_HB_D1;
        if (!any_h) ag::fail(AG_ERR_NO_DONOR_H, "no closestH atom was found");
    }
    if (ms->hydrogen_type < -1 || ms->hydrogen_type >= T) ag::fail(AG_ERR_BAD_INPUT, "hydrogen type index out of range");
    if (M > 0 && !ms->maps) ag::fail(AG_ERR_BAD_INPUT, "map descriptors missing");
    if (M > 0 && T > 0 && !tab->e_vdw_hb) ag::fail(AG_ERR_BAD_INPUT, "vdW/H-bond tables missing");
    if (!tab->sol_fn || (ms->dddiel && !tab->r_eps)) ag::fail(AG_ERR_BAD_INPUT, "desolvation/dielectric tables missing");
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

static void launch_near(const ag_plan &p, int grid, const NearParams &np, hipStream_t st) {
#define AG_NEAR(NP, CB)                                                                 \
    do {                                                                                \
        if (p.paged) k_near<NP, CB, kNearPaged><<<grid, kNearThreads, 0, st>>>(np);     \
        else if (np.bhtree || np.cmap || p.any_disorder)                                \
            k_near<NP, CB, kNearGeneral><<<grid, kNearThreads, 0, st>>>(np);            \
        else k_near<NP, CB, kNearPlain><<<grid, kNearThreads, 0, st>>>(np);             \
    } while (0)
    if (p.tune.cube) {
        if (p.nplain == 4) AG_NEAR(4, true);
        else if (p.nplain == 8) AG_NEAR(8, true);
        else AG_NEAR(12, true);
    } else {
        if (p.nplain == 4) AG_NEAR(4, false);
        else if (p.nplain == 8) AG_NEAR(8, false);
        else AG_NEAR(12, false);
    }
#undef AG_NEAR
    HIP_CHECK(hipGetLastError());
}

// Variants by LDS table capacity (12288 entries, 48 KB, when the table
// fits, else kFarTabMax, 64 KB) and by the paths the launch's items take.
template <int MODE>
static const void *far_kernel_mode(const FarParams &fp) {
    return fp.tab_len <= 12288 ? reinterpret_cast<const void *>(&k_far<12288, MODE>)
                               : reinterpret_cast<const void *>(&k_far<kFarTabMax, MODE>);
}
static const void *far_kernel(const FarParams &fp, int mode) {
    return mode == kFarTable ? far_kernel_mode<kFarTable>(fp)
         : mode == kFarTfree ? far_kernel_mode<kFarTfree>(fp)
                             : far_kernel_mode<kFarMixed>(fp);
}

static void launch_far(int blocks, const FarParams &fp, int mode, hipStream_t st) {
    void *args[] = {const_cast<FarParams *>(&fp)};
    HIP_CHECK(hipLaunchKernel(far_kernel(fp, mode), dim3(blocks), dim3(kFarThreads), args, 0, st));
}

// Per-conformation geometry, work prefixes and cell arrays for the plan's
// cell size.
// far_flags[c] (or all zero when null): bit 0 the far kernel's table-free
// path for conformation c, bit 1 its table-index cap.
static void plan_geometry(ag_plan &p, const ag_grid *grids, const int *natoms, const int *far_flags) {
    const int n = p.nconf;
    const double cs = p.tune.cell_size;
    p.confs.assign(n, ConfDev{});
    p.brick_prefix.assign(n + 1, 0);
    p.tile_prefix.assign(n + 1, 0);
    long long brick = 0, tile = 0, moff = 0, poff = 0;
    int cell_base = 0, atom_base = 0;
    for (int c = 0; c < n; c++) {
        ConfDev &cf = p.confs[c];
        memset(&cf, 0, sizeof cf);
        const ag_grid &g = grids[c];
        for (int d = 0; d < 3; d++) {
            cf.n1[d] = g.n1[d];
            cf.ne[d] = g.ne[d];
            const double lo = -g.ne[d] * g.spacing - 8.1;
            const double hi = (g.n1[d] - 1 - g.ne[d]) * g.spacing + 8.1;
            cf.cell_origin[d] = lo;
            cf.nc[d] = (int)ceil((hi - lo) / cs) + 1;
        }
        cf.spacing = g.spacing;
        cf.s128 = (float)(g.spacing * 128.0);
        cf.nb[0] = (g.n1[0] + kBX - 1) / kBX;
        cf.nb[1] = (g.n1[1] + kBY - 1) / kBY;
        cf.nb[2] = (g.n1[2] + kBZ - 1) / kBZ;
        cf.cell_base = cell_base;
        cf.atom_base = atom_base;
        cf.n_atoms = natoms[c];
        cf.P = (long long)g.n1[0] * g.n1[1] * g.n1[2];
        cf.map_off = moff;
        cf.pt_off = poff;
        cf.runs_per_row = (g.n1[0] + kFarK - 1) / kFarK;
        cf.nruns = cf.runs_per_row * g.n1[1] * g.n1[2];
        cf.far_const = far_flags ? far_flags[c] & 1 : 0;
        cf.far_clamp = far_flags ? far_flags[c] >> 1 & 1 : 0;
        p.brick_prefix[c] = (int)brick;
        p.tile_prefix[c] = (int)tile;
        brick += (long long)cf.nb[0] * cf.nb[1] * cf.nb[2];
        // far-kernel items: patches of 64 rows x x-runs, kFarRuns / 64 wave slots each
        const long long far_waves = (long long)((g.n1[1] * g.n1[2] + 63) / 64) * cf.runs_per_row;
        tile += (far_waves + kFarRuns / 64 - 1) / (kFarRuns / 64);
        const long long cells = (long long)cf.nc[0] * cf.nc[1] * cf.nc[2];
        if (cell_base + cells > INT32_MAX / 2 || brick > INT32_MAX / 2 || tile > INT32_MAX / 2)
            ag::fail(AG_ERR_UNSUPPORTED, "batch too large for one plan");
        cell_base += (int)cells;
        atom_base += natoms[c];
        moff += (long long)p.n_maps * cf.P;
        poff += cf.P;
    }
    p.brick_prefix[n] = (int)brick;
    p.tile_prefix[n] = (int)tile;
    p.total_bricks = (int)brick;
    p.total_tiles = (int)tile;
    p.total_cells = cell_base;
    p.P_total = poff;
    if (p.want_collide) {
        p.near[0].collide = (unsigned char *)p.collide.get((size_t)p.P_total);
    }
    hipStream_t st = p.ctx->stream;
    upload(st, p.confs_d, p.confs.data(), p.confs.size());
    upload(st, p.brick_prefix_d, p.brick_prefix.data(), p.brick_prefix.size());
    upload(st, p.tile_prefix_d, p.tile_prefix.data(), p.tile_prefix.size());
    p.cell_count.get(sizeof(int) * (size_t)p.total_cells);
    p.cell_start.get(sizeof(int) * ((size_t)p.total_cells + 1));
    p.cell_cursor.get(sizeof(int) * (size_t)p.total_cells);
    const ConfDev *d_confs = (const ConfDev *)p.confs_d.ptr;
    for (auto &q : p.near) {
        q.confs = d_confs;
        q.brick_prefix = (const int *)p.brick_prefix_d.ptr;
        q.nconf = n;
        q.total_bricks = p.total_bricks;
        q.cell_size = cs;
        q.cell_start = (const int *)p.cell_start.ptr;
    }
    p.far.confs = d_confs;
    p.far.tile_prefix = (const int *)p.tile_prefix_d.ptr;
    p.far.nconf = n;
    p.far.total_tiles = p.total_tiles;
    int occ = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, far_kernel(p.far, kFarMixed),
                                                           kFarThreads, 0));
    p.far_grid = std::max(1, std::min(p.total_tiles, std::max(occ, 1) * p.ctx->n_cu));
}

// Build a plan over n conformations (a single plan is n = 1).
// An arena pool (the allocation-free entry point) also keeps the host
// staging in pageable memory instead of the context's pinned buffer.
static ag_plan *plan_build(ag_ctx *c, int n, const ag_receptor *recs, const ag_grid *grids, const ag_mapset *ms,
                           const ag_tables *tab, int potential, bool persistent, const Tune &tune,
                           std::shared_ptr<DevPool> pool = nullptr) {
    const bool vina = potential == AG_POTENTIAL_VINA;
    if (potential != AG_POTENTIAL_AD4 && !vina) ag::fail(AG_ERR_BAD_INPUT, "unknown potential");
    if (n < 1) ag::fail(AG_ERR_BAD_INPUT, "a plan needs at least one conformation");
    for (int i = 0; i < n; i++) validate(&recs[i], &grids[i], ms, tab, vina);
    HIP_CHECK(hipSetDevice(c->device));
    if (!pool) pool = c->pool;
    const bool no_alloc = pool->arena != nullptr;
    std::unique_ptr<ag_plan> pl(new ag_plan(pool));
    ag_plan &p = *pl;
    p.ctx = c;
    p.nconf = n;
    p.vina = vina;
    p.persistent = persistent;
    p.tune = tune;
    const int M = p.M = ms->n_atom_maps;
    const int T = p.T = tab->n_rec_types;
    p.n_maps = ms->n_maps;
    p.coeff_estat = ms->coeff_estat;
    hipStream_t st = c->stream;

    // concatenated atoms
    std::vector<int> natoms(n);
    long long Nl = 0;
    for (int i = 0; i < n; i++) Nl += (natoms[i] = recs[i].n_atoms);
    if (Nl > INT32_MAX / 4) ag::fail(AG_ERR_UNSUPPORTED, "too many atoms in one batch");
    const int N = p.N = (int)Nl;
    // concatenated receptor arrays, staged in the context's pinned buffer
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~size_t(255);
        return o;
    };
    const size_t n3 = 3 * (size_t)N, nn = (size_t)N;
    const size_t o_coord = carve(8 * n3), o_charge = carve(8 * nn), o_vol = carve(8 * nn), o_solpar = carve(8 * nn);
    const size_t o_rvec = carve(8 * n3), o_rvec2 = carve(8 * n3), o_type = carve(4 * nn), o_hbond = carve(4 * nn);
    const size_t o_rexp = carve(4 * nn), o_conf = carve(4 * nn), o_dis = carve(nn);
    std::vector<char> pageable(no_alloc ? off + 256 : 0);
    char *const h = no_alloc ? pageable.data() : (char *)c->stage.get(off + 256);
    double *const coord = (double *)(h + o_coord), *const charge = (double *)(h + o_charge);
    double *const vol = (double *)(h + o_vol), *const solpar = (double *)(h + o_solpar);
    double *const rvec = (double *)(h + o_rvec), *const rvec2 = (double *)(h + o_rvec2);
    int *const type = (int *)(h + o_type), *const hbond = (int *)(h + o_hbond), *const rexp = (int *)(h + o_rexp);
    int *const atom_conf = (int *)(h + o_conf);
    unsigned char *const dis = (unsigned char *)(h + o_dis);
    std::vector<int> first_atom(n + 1, 0);
    for (int i = 0; i < n; i++) first_atom[i + 1] = first_atom[i] + natoms[i];
    auto stage_conf = [&](int i) {
        const int o = first_atom[i];
        const ag_receptor &r = recs[i];
        const size_t k = natoms[i];
        if (!k) return;
        std::copy(r.coord, r.coord + 3 * k, coord + 3 * (size_t)o);
        std::copy(r.charge, r.charge + k, charge + o);
        std::copy(r.vol, r.vol + k, vol + o);
        std::copy(r.solpar, r.solpar + k, solpar + o);
        std::copy(r.rvector, r.rvector + 3 * k, rvec + 3 * (size_t)o);
        std::copy(r.rvector2, r.rvector2 + 3 * k, rvec2 + 3 * (size_t)o);
        std::copy(r.type, r.type + k, type + o);
        std::copy(r.hbond, r.hbond + k, hbond + o);
        std::copy(r.rexp, r.rexp + k, rexp + o);
        std::copy(r.disorder, r.disorder + k, dis + o);
        std::fill(atom_conf + o, atom_conf + o + k, i);
    };
    // large batches copy on several host threads (at most 16), one
    // conformation at a time
    const int n_thr = N > 200000 ? std::max(1, std::min({16, n, (int)std::thread::hardware_concurrency()})) : 1;
    if (n_thr > 1) {
        std::atomic<int> next{0};
        std::vector<std::thread> th;
        for (int t = 0; t < n_thr; t++)
            th.emplace_back([&] {
                for (int i; (i = next.fetch_add(1)) < n;) stage_conf(i);
            });
        for (auto &x : th) x.join();
    } else {
        for (int i = 0; i < n; i++) stage_conf(i);
    }
    double *d_coord = upload(st, p.coord, coord, n3);
    double *d_charge = upload(st, p.charge, charge, nn);
    double *d_vol = upload(st, p.vol, vol, nn);
    double *d_solpar = upload(st, p.solpar, solpar, nn);
    double *d_rvec = upload(st, p.rvec, rvec, n3);
    double *d_rvec2 = upload(st, p.rvec2, rvec2, n3);
    int *d_type = upload(st, p.type, type, nn);
    int *d_hbond = upload(st, p.hbond, hbond, nn);
    int *d_rexp = upload(st, p.rexp, rexp, nn);
    unsigned char *d_dis = upload(st, p.disorder, dis, nn);
    p.any_disorder = std::any_of(dis, dis + N, [](unsigned char v) { return v != 0; });
    if (n > 1) upload(st, p.atom_conf_d, atom_conf, nn);
    p.src = AtomSrc{d_coord, d_charge, d_vol, d_solpar, d_rvec, d_rvec2, d_type, d_hbond, d_rexp, d_dis,
                    ms->coeff_estat, ms->solpar_q};
    p.atom_cell.get(sizeof(int) * (size_t)N);
    p.slot_atom.get(sizeof(int) * (size_t)N);
    // the near kernel's candidate lists hold 32-bit byte offsets of records
    if ((unsigned long long)N * sizeof(NearAtom) > UINT32_MAX)
        ag::fail(AG_ERR_UNSUPPORTED, "too many atoms in one plan for the near kernel (at most 44,739,242)");
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
    for (int m = 0; m < M; m++) {
        const int role = map_role(ms, m);
        if (role == ROLE_COV) p.cov_slots.push_back(m);
        else if (!vina && (role & ROLE_AFF) && ms->maps[m].is_hbonder) hb_maps.push_back(m);
        else plain_maps.push_back(m);
    }
    // A pass has kPlainSlots plain slots and kHbSlots H-bond slots; plain
    // maps beyond the plain slots take free H-bond slots, whose code is then
    // the plain table add for every atom class (13 Vina maps: one pass).
    const int n_plain = (int)plain_maps.size(), n_hb = (int)hb_maps.size();
    p.groups = std::max({1, (n_hb + kHbSlots - 1) / kHbSlots, (n_plain + n_hb + kMapSlots - 1) / kMapSlots});
    const size_t gstride = (size_t)(T > 0 ? T : 1) * AG_NEINT * kMapSlots;
    std::vector<float> gtab(gstride * p.groups, 0.f);
    std::vector<std::array<int, kMapSlots>> slot_map(p.groups);
    std::vector<std::array<bool, kHbSlots>> hb_plain(p.groups);
    int nplain_max = 0;
    for (int g = 0; g < p.groups; g++) {
        slot_map[g].fill(-1);
        hb_plain[g].fill(false);
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
    }
    for (size_t i = (size_t)p.groups * kPlainSlots; i < plain_maps.size(); i++) {
        bool placed = false;
        for (int g = 0; g < p.groups && !placed; g++)
            for (int k = 0; k < kHbSlots && !placed; k++)
                if (slot_map[g][kPlainSlots + k] < 0) {
                    slot_map[g][kPlainSlots + k] = plain_maps[i];
                    hb_plain[g][k] = true;
                    placed = true;
                }
        if (!placed) ag::fail(AG_ERR_UNSUPPORTED, "near-kernel slot assignment left a map without a slot");
    }
    for (int g = 0; g < p.groups; g++) {
        for (int t = 0; t < T; t++)
            for (int j = 0; j < kMapSlots; j++) {
                const int m = slot_map[g][j];
                // a slot without an affinity term (desolvation only) keeps
                // zeros, so the kernel adds every slot's table value
                if (m < 0 || !(map_role(ms, m) & ROLE_AFF)) continue;
                const float *srcrow = tab->e_vdw_hb + ((size_t)t * M + m) * AG_NEINT;
                float *dst = gtab.data() + g * gstride + (size_t)t * AG_NEINT * kMapSlots + j;
                for (int k = 0; k < AG_NEINT; k++) dst[(size_t)k * kMapSlots] = srcrow[k];
            }
    }
    p.nplain = nplain_max <= 4 ? 4 : (nplain_max <= 8 ? 8 : 12);
    float *d_tab = upload(st, p.tab, gtab.data(), gtab.size());
    p.queue_d.get(sizeof(int) * kMaxChunks);

    NearParams np;
    memset(&np, 0, sizeof np);
    np.atoms = (const NearAtom *)p.near_atoms.ptr;
    np.coord64 = d_coord;
    np.rvec64 = d_rvec;
    np.rvec2_64 = d_rvec2;
    np.sol = d_sol;
    np.reps = d_reps;
    np.disorder_h = ms->disorder_h;
    np.hydrogen_type = ms->hydrogen_type;
    np.status = (int *)p.status.ptr;
    for (int g = 0; g < p.groups; g++) {
        NearParams q = np;
        bool used = false;
        for (int j = 0; j < kMapSlots; j++) {
            const int m = slot_map[g][j];
            q.out_slot[j] = m;
            q.sp[j] = 0.f;
            q.volp[j] = 0.f;
            if (m < 0) continue;
            used = true;
            const int role = map_role(ms, m);
            const ag_map_desc &md = ms->maps[m];
            if ((role & ROLE_DES) && !vina) {
                q.sp[j] = (float)md.solpar_probe;
                q.volp[j] = (float)md.vol_probe;
            }
            if (j >= kPlainSlots) {
                const int k = j - kPlainSlots;
                for (int ah = 0; ah <= 6; ah++) {
                    if (hb_plain[g][k]) {   // a plain map in an H-bond slot
                        q.hcode[k][ah] = HC_PLAIN | (HC_PLAIN << 4);
                        continue;
                    }
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
    // the tree build's semantics (every pass: Loop 2; pass 0: cmap, collisions)
    for (auto &q : p.near) q.bhtree = ms->bhtree;
    p.want_collide = ms->bhtree == 1 && (ms->elec_index >= 0 || ms->floating_grid);
    p.e_index = vina ? -1 : ms->elec_index;
    if (p.want_fmap && N > 0) {
        float4 *d_fa = (float4 *)p.fmin_atoms.get(sizeof(float4) * (size_t)N);
        k_fmin_atoms<<<(N + 255) / 256, 256, 0, st>>>(N, d_coord, d_fa);
        HIP_CHECK(hipGetLastError());
    }
    if (p.cov_slots.size() > 16) ag::fail(AG_ERR_UNSUPPORTED, "more than 16 covalent maps");
    for (int d = 0; d < 3; d++) p.cov_pos[d] = ms->cov_pos[d];
    p.cov_half_width = ms->cov_half_width;
    p.cov_barrier = ms->cov_barrier;
    // Vina: the e and d maps stay zero (main.cpp:2968, 3391)
    if (vina) {
        if (ms->elec_index >= 0) p.zero_slots.push_back(ms->elec_index);
        if (ms->dsolv_index >= 0) p.zero_slots.push_back(ms->dsolv_index);
    }

    // far kernel (e map): per-atom float4 and the table with bins < 1025
    // zeroed, long enough for the largest bin any (point, atom) pair of any
    // conformation reaches
    p.run_far = !vina && ms->elec_index >= 0 && N > 0;
    std::vector<int> far_flags;
    if (p.run_far) {
        float4 *d_far = (float4 *)p.far_atoms.get(sizeof(float4) * (size_t)N);
        k_far_atoms<<<(N + 255) / 256, 256, 0, st>>>(N, d_coord, d_charge, ms->coeff_estat, d_far);
        HIP_CHECK(hipGetLastError());
        p.far_atoms_ready = true;
        // Per conformation: the largest bin any of its pairs reaches
        // (need) and the share of its pairs past bin 8192 (r >= 64 A),
        // from a 4 x 4 x 4 lattice of its box against about 128 of its
        // atoms.  The table-free path saves 3 of 6.875 instructions on each
        // such pair but its per-block atom test costs about 0.8 per pair, so
        // it pays from a share of about 30 %; it is also the one path that
        // caps table indices, so a conformation that may pass the table's
        // end takes it.
        // Both depend on the conformation alone: its maps are the same in a
        // batch as in a plan of its own.
        long long need_max = 0;
        far_flags.assign(n, 0);
        for (int i = 0, o = 0; i < n; o += natoms[i], i++) {
            const ag_grid &grid = grids[i];
            const int na = natoms[i];
            if (na == 0) continue;
            double rmax = 0.0;
            for (int a = o; a < o + na; a++) {
                double q = 0.0;
                for (int d = 0; d < 3; d++) {
                    const double lo = -grid.ne[d] * grid.spacing, hi = (grid.n1[d] - 1 - grid.ne[d]) * grid.spacing;
                    const double x = coord[3 * (size_t)a + d];
                    const double e = std::max(std::fabs(x - lo), std::fabs(x - hi));
                    q += e * e;
                }
                rmax = std::max(rmax, q);
            }
            const long long need = (long long)(std::sqrt(rmax) * 128.0) + 8;
            need_max = std::max(need_max, need);
            const int stride = std::max(1, na / 128);
            long long far_n = 0, tot_n = 0;
            for (int qz = 0; qz < 4; qz++)
                for (int qy = 0; qy < 4; qy++)
                    for (int qx = 0; qx < 4; qx++) {
                        const int q[3] = {qx, qy, qz};
                        double pnt[3];
                        for (int d = 0; d < 3; d++)
                            pnt[d] = ((grid.n1[d] - 1) * (q[d] + 0.5) / 4.0 - grid.ne[d]) * grid.spacing;
                        for (int a = o; a < o + na; a += stride) {
                            double r2 = 0.0;
                            for (int d = 0; d < 3; d++) {
                                const double t = coord[3 * (size_t)a + d] - pnt[d];
                                r2 += t * t;
                            }
                            far_n += r2 >= 64.0 * 64.0;
                            tot_n++;
                        }
                    }
            const bool clamp = need > kFarTabMax;
            const bool tfree = clamp || (double)far_n > 0.30 * (double)tot_n;
            far_flags[i] = (tfree ? 1 : 0) | (clamp ? 2 : 0);
        }
        // The table runs to the largest bin any pair reaches, up to
        // kFarTabMax entries; entries past bin 8191 all hold r_eps[8191].
        const int tab_len = (int)std::min<long long>(std::max<long long>(need_max, kFarBinMin + 1), kFarTabMax);
        std::vector<float> ft(tab_len);
        for (int k = 0; k < tab_len; k++) ft[k] = k < kFarBinMin ? 0.f : reps[std::min(k, AG_NDIEL - 1)];
        float *d_ft = upload(st, p.fartab, ft.data(), ft.size());
        FarParams &fp = p.far;
        memset(&fp, 0, sizeof fp);
        fp.atoms = d_far;
        fp.tab = d_ft;
        fp.tab_len = tab_len;
        fp.tmax4 = 4u * (unsigned)(tab_len - 1);
        fp.c_last = reps[AG_NDIEL - 1];
        fp.e_slot = ms->elec_index;
    }
    plan_geometry(p, grids, natoms.data(), p.run_far ? far_flags.data() : nullptr);
    // A batch's far field is captured as two graphs: conformations [0,
    // far_split), then the rest, so that the first part's e maps can go to
    // the host while the second part computes.  The second part is sized so
    // that its far field (at about 5.5e12 pairs/s) lasts about as long as
    // that copy (at about 45 GB/s): what is left after the map set is the
    // second part's e maps.
    p.far_split = n;
    if (persistent && n > 1 && p.run_far) {
        double work = 0.0, ebytes = 0.0;
        for (const ConfDev &cf : p.confs) {
            work += (double)cf.P * cf.n_atoms;
            ebytes += 4.0 * (double)cf.P;
        }
        const double t_far = work / 5.5e12, t_copy = ebytes / 45e9;
        const double want_b = work * t_copy / (t_far + t_copy);
        double acc = 0.0;
        int split = n;
        while (split > 1 && acc < want_b) {
            --split;
            acc += (double)p.confs[split].P * p.confs[split].n_atoms;
        }
        p.far_split = split;
    }
    HIP_CHECK(hipStreamSynchronize(st));   // host staging vectors go out of scope
    return pl.release();
}


// Host destinations of a batch: conformation i's maps / fmap / cmap.
struct HostOut {
    float *const *maps;
    float *const *fmap;
    int32_t *const *cmap;
};

// Contiguous conformation ranges of about equal work (points x atoms).
// Chunks of consecutive conformations for a batch with host outputs: `want`
// chunks whose work (points x atoms) follows the weights 1, ..., 1, 1/2,
// 1/4, ..., 1/2^taper.  The last chunks shrink so that the copy left after
// the final kernel is short; fewer, larger chunks before them keep the
// number of chunk boundaries low.
static std::vector<std::pair<int, int>> plan_chunks(const ag_plan &p, int want, int taper) {
    std::vector<std::pair<int, int>> out;
    const int n = p.nconf;
    want = std::max(1, std::min({want, n, kMaxChunks}));
    taper = std::max(0, std::min(taper, want - 1));
    std::vector<double> bound(want);   // cumulative work fraction at the end of each chunk
    {
        double sum = 0.0;
        for (int k = 0; k < want; k++) sum += k < want - taper ? 1.0 : std::ldexp(1.0, -(k - (want - taper) + 1));
        double acc = 0.0;
        for (int k = 0; k < want; k++) {
            acc += k < want - taper ? 1.0 : std::ldexp(1.0, -(k - (want - taper) + 1));
            bound[k] = acc / sum;
        }
    }
    double total = 0.0;
    for (const ConfDev &cf : p.confs) total += (double)cf.P * (cf.n_atoms + 64);
    int c0 = 0;
    double acc = 0.0;
    for (int c = 0; c < n; c++) {
        acc += (double)p.confs[c].P * (p.confs[c].n_atoms + 64);
        const int k = (int)out.size();
        if (c == n - 1 || (k < want - 1 && acc >= total * bound[k])) {
            out.emplace_back(c0, c + 1);
            c0 = c + 1;
        }
    }
    return out;
}

enum : int { kPartAll = 0, kPartNear = 1, kPartFar = 2 };

// Enqueue one map set on the context's stream, conformation chunk by chunk.
// With events, HIP events bracket the cell list, the near-field and the
// far-field phases (single chunk).  With host outputs, each chunk's maps
// are copied to the host on the context's copy stream as soon as the chunk
// is done, while the next chunk computes.  `part` (single chunk) enqueues
// only the cell list and near field (kPartNear) or only the far field
// (kPartFar); with kPartAll, mark_near records the context's near_ev once the
// near field of every conformation is enqueued.
static void enqueue(ag_plan *p, float *d_maps, float *d_fmap, int32_t *d_cmap, bool events, bool skip_far,
                    const std::vector<std::pair<int, int>> &chunks, const HostOut *host, bool mark_near = false,
                    int part = kPartAll, int queue_base = 0) {
    ag_ctx *c = p->ctx;
    hipStream_t st = c->stream;
    const int N = p->N;
    const int tpb = 256;
    int *d_atom_cell = (int *)p->atom_cell.ptr, *d_count = (int *)p->cell_count.ptr;
    int *d_start = (int *)p->cell_start.ptr, *d_cursor = (int *)p->cell_cursor.ptr;
    int *d_slot_atom = (int *)p->slot_atom.ptr;
    int *d_status = (int *)p->status.ptr;
    int *d_queue = (int *)p->queue_d.ptr;
    const int nchunks = (int)chunks.size();
    if (nchunks < 1 || nchunks > kMaxChunks) ag::fail(AG_ERR_BAD_INPUT, "bad chunking of the batch");
    if (part != kPartAll && (nchunks != 1 || host)) ag::fail(AG_ERR_BAD_INPUT, "bad chunking of the batch");
    const unsigned cgrid = (unsigned)((p->total_cells + tpb - 1) / tpb);
    if (events) HIP_CHECK(hipEventRecord(c->ev[0], st));
    if (part != kPartFar) {
    k_zero_cells<<<std::max(cgrid, (unsigned)((kMaxChunks + tpb - 1) / tpb)), tpb, 0, st>>>(
        d_count, d_cursor, p->total_cells, d_status, d_queue, kMaxChunks);
    if (N > 0)
        k_cell_count<<<(N + tpb - 1) / tpb, tpb, 0, st>>>((const ConfDev *)p->confs_d.ptr,
                                                          p->nconf > 1 ? (const int *)p->atom_conf_d.ptr : nullptr, N,
                                                          p->tune.cell_size, (const double *)p->coord.ptr,
                                                          d_atom_cell, d_count);
    k_scan<<<1, 1024, 0, st>>>(d_count, d_start, p->total_cells);
    if (N > 0) k_cell_scatter<<<(N + tpb - 1) / tpb, tpb, 0, st>>>(N, d_atom_cell, d_start, d_cursor, d_slot_atom);
    k_cell_sort<<<cgrid, tpb, 0, st>>>(p->total_cells, d_start, d_slot_atom);
    if (N > 0)
        k_gather<<<(N + tpb - 1) / tpb, tpb, 0, st>>>(d_start + p->total_cells, d_slot_atom, p->src,
                                                      (NearAtom *)p->near_atoms.ptr);
    HIP_CHECK(hipGetLastError());
    }
    if (events) HIP_CHECK(hipEventRecord(c->ev[1], st));
    // The near field (near kernel, floating grid, covalent and zero maps) of
    // conformations [c0, c1) on stream s.
    auto near_pass = [&](int c0, int c1, hipStream_t s) {
        const int b0 = p->brick_prefix[c0], b1 = p->brick_prefix[c1];
        for (int g = 0; g < p->groups; g++) {
            NearParams np = p->near[g];
            np.maps = d_maps;
            np.cmap = (g == 0 && p->want_cmap) ? d_cmap : nullptr;
            np.item_base = b0;
            if (!(np.has_affinity || np.e_slot >= 0 || np.cmap) || b1 <= b0) continue;
            launch_near(*p, b1 - b0, np, s);
        }
        for (int ci = c0; ci < c1; ci++) {
            const ConfDev &cf = p->confs[ci];
            const unsigned pgrid = (unsigned)((cf.P + 255) / 256);
            if (p->want_fmap) {
                FminParams fm;
                memset(&fm, 0, sizeof fm);
                for (int d = 0; d < 3; d++) {
                    fm.n1[d] = cf.n1[d];
                    fm.ne[d] = cf.ne[d];
                }
                fm.spacing = (float)cf.spacing;
                fm.runs_per_row = cf.runs_per_row;
                fm.nruns = cf.nruns;
                fm.atoms = (const float4 *)p->fmin_atoms.ptr + cf.atom_base;
                fm.n_atoms = cf.n_atoms;
                fm.fmap = d_fmap + cf.pt_off;
                k_fmin<<<(fm.nruns + 255) / 256, 256, 0, s>>>(fm);
            }
            if (!p->cov_slots.empty()) {
                CovParams cv;
                memset(&cv, 0, sizeof cv);
                for (int d = 0; d < 3; d++) {
                    cv.n1[d] = cf.n1[d];
                    cv.ne[d] = cf.ne[d];
                    cv.pos[d] = p->cov_pos[d];
                }
                cv.spacing = cf.spacing;
                cv.half_width = p->cov_half_width;
                cv.barrier = p->cov_barrier;
                cv.n_cov = (int)p->cov_slots.size();
                for (int q = 0; q < cv.n_cov; q++) cv.slot[q] = p->cov_slots[q];
                cv.P = cf.P;
                cv.maps = d_maps + cf.map_off;
                k_covalent<<<pgrid, 256, 0, s>>>(cv);
            }
            for (int m : p->zero_slots)
                k_fill<<<pgrid, 256, 0, s>>>(d_maps + cf.map_off + (size_t)m * cf.P, cf.P, 0.f);
        }
        HIP_CHECK(hipGetLastError());
    };
    // The far field of chunk k ([c0, c1), work counter k) and the tree
    // build's collision rule after it, on stream s.
    auto far_pass = [&](int k, int c0, int c1, hipStream_t s) {
        const int t0 = p->tile_prefix[c0], t1 = p->tile_prefix[c1];
        if (p->run_far && !skip_far && t1 > t0) {
            FarParams fp = p->far;
            fp.maps = d_maps;
            fp.item_base = t0;
            fp.item_end = t1;
            fp.queue = p->persistent ? d_queue + queue_base + k : nullptr;
            const int blocks = p->persistent ? std::min(p->far_grid, t1 - t0) : t1 - t0;
            launch_far(blocks, fp, far_mode(*p, c0, c1), s);
        }
        if (p->want_collide)
            for (int ci = c0; ci < c1; ci++) {
                const ConfDev &cf = p->confs[ci];
                k_collide<<<(unsigned)((cf.P + 255) / 256), 256, 0, s>>>(
                    cf.P, (const unsigned char *)p->collide.ptr + cf.pt_off,
                    p->e_index >= 0 ? d_maps + cf.map_off + (size_t)p->e_index * cf.P : nullptr,
                    d_fmap ? d_fmap + cf.pt_off : nullptr);
            }
        HIP_CHECK(hipGetLastError());
    };
    // Chunk k's outputs to the host on the copy stream, once s has them.
    auto copy_out = [&](int k, int c0, int c1, hipStream_t s) {
        HIP_CHECK(hipEventRecord(c->chunk_ev[k], s));
        HIP_CHECK(hipStreamWaitEvent(c->copy_stream, c->chunk_ev[k], 0));
        for (int ci = c0; ci < c1; ci++) {
            const ConfDev &cf = p->confs[ci];
            HIP_CHECK(hipMemcpyAsync(host->maps[ci], d_maps + cf.map_off, sizeof(float) * (size_t)p->n_maps * cf.P,
                                     hipMemcpyDeviceToHost, c->copy_stream));
            if (d_fmap)
                HIP_CHECK(hipMemcpyAsync(host->fmap[ci], d_fmap + cf.pt_off, sizeof(float) * (size_t)cf.P,
                                         hipMemcpyDeviceToHost, c->copy_stream));
            if (d_cmap)
                HIP_CHECK(hipMemcpyAsync(host->cmap[ci], d_cmap + cf.pt_off, sizeof(int32_t) * (size_t)cf.P,
                                         hipMemcpyDeviceToHost, c->copy_stream));
        }
    };
    const hipStream_t st_main = st;
    if (host && nchunks > 1) {
        // Pipelined batch: the whole batch's near field first, on the main
        // stream, then the chunks' far fields alternating between the two
        // compute streams (chunks touch disjoint conformations), each chunk's
        // copies behind its far kernel.  Copies start after the first chunk's
        // far field and overlap the rest; two far launches in flight fill
        // each other's tails.  (Near and far passes interleaved chunk by
        // chunk measured 8 ms slower on the 64-conformation ensemble: a near
        // pass beside the other stream's persistent far kernel waits for CU
        // slots.)
        near_pass(0, p->nconf, st_main);
        HIP_CHECK(hipEventRecord(c->cells_ev, st_main));
        HIP_CHECK(hipStreamWaitEvent(c->stream2, c->cells_ev, 0));
        for (int k = 0; k < nchunks; k++) {
            const hipStream_t s = (k & 1) ? c->stream2 : st_main;
            far_pass(k, chunks[k].first, chunks[k].second, s);
            copy_out(k, chunks[k].first, chunks[k].second, s);
        }
    } else {
        for (int k = 0; k < nchunks; k++) {
            if (part != kPartFar) near_pass(chunks[k].first, chunks[k].second, st_main);
            if (events && k == nchunks - 1) HIP_CHECK(hipEventRecord(c->ev[2], st_main));
            if (mark_near && k == nchunks - 1) HIP_CHECK(hipEventRecord(c->near_ev, st_main));
            if (part != kPartNear) far_pass(k, chunks[k].first, chunks[k].second, st_main);
            if (host) copy_out(k, chunks[k].first, chunks[k].second, st_main);
        }
    }
    st = st_main;
    if (host) {
        // the copy stream has waited for every chunk, so this joins both
        // compute streams as well
        HIP_CHECK(hipEventRecord(c->chunk_ev[kMaxChunks], c->copy_stream));
        HIP_CHECK(hipStreamWaitEvent(st, c->chunk_ev[kMaxChunks], 0));
    }
    if (events) HIP_CHECK(hipEventRecord(c->ev[3], st));
}

static void enqueue(ag_plan *p, float *d_maps, float *d_fmap, int32_t *d_cmap, bool events, bool skip_far,
                    bool mark_near = false, int part = kPartAll) {
    enqueue(p, d_maps, d_fmap, d_cmap, events, skip_far, {{0, p->nconf}}, nullptr, mark_near, part);
}

// Whether a brick of the last map set overflowed the near kernel's LDS lists
// (syncs the context's stream).
static bool overflowed(ag_plan *p) {
    // into pinned memory: a direct copy, not a staged one (a plan on an
    // arena pool allocates no pinned word and copies into the stack)
    int local = 0;
    int *status = p->pool->arena ? &local : (int *)p->ctx->status_word.get(sizeof(int));
    HIP_CHECK(hipMemcpyAsync(status, p->status.ptr, sizeof(int), hipMemcpyDeviceToHost, p->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(p->ctx->stream));
    return (*status & 1) != 0;
}

// Runs one map set of p through `run` and checks the overflow word: when a
// brick had more than 2048 candidate atoms or 256 donor H within 8.03 A, the
// plan switches to the paged near kernel for good (its graph, if any, is
// captured again) and runs the map set once more.
template <class F>
static void run_checked(ag_plan *p, F &&run) {
    run();
    if (!overflowed(p)) return;
    if (p->paged) ag::fail(AG_ERR_DEVICE, "the paged near kernel reported an overflow");
    p->paged = true;
    p->drop_graphs();
    run();
    if (overflowed(p)) ag::fail(AG_ERR_DEVICE, "the paged near kernel reported an overflow");
}

// Autotuner: times the near-field pass of every candidate layout on a
// sample of the batch (up to 8 conformations spread over its sizes) and
// keeps the fastest.  Results differ between cell sizes only in the order
// of the FP32 sums; for a fixed choice every run is bitwise repeatable.
static Tune autotune(ag_ctx *c, int n, const ag_receptor *recs, const ag_grid *grids, const ag_mapset *ms,
                     const ag_tables *tab, int potential, double *spent_ms) {
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        const long long pa = (long long)grids[a].n1[0] * grids[a].n1[1] * grids[a].n1[2] * std::max(recs[a].n_atoms, 1);
        const long long pb = (long long)grids[b].n1[0] * grids[b].n1[1] * grids[b].n1[2] * std::max(recs[b].n_atoms, 1);
        return pa != pb ? pa < pb : a < b;
    });
    const int ns = std::min(n, 8);
    std::vector<ag_receptor> srec(ns);
    std::vector<ag_grid> sgrid(ns);
    for (int k = 0; k < ns; k++) {
        const int i = order[ns == 1 ? n - 1 : (int)((long long)k * (n - 1) / (ns - 1))];
        srec[k] = recs[i];
        sgrid[k] = grids[i];
    }
    long long sp = 0;
    for (int k = 0; k < ns; k++) sp += (long long)sgrid[k].n1[0] * sgrid[k].n1[1] * sgrid[k].n1[2];
    DevBuf out;
    float *d_maps = (float *)out.get(sizeof(float) * (size_t)sp * ms->n_maps + 256);
    DevBuf extra;
    char *x = (char *)extra.get(2 * sizeof(float) * (size_t)sp + 256);
    float *d_fmap = ms->floating_grid ? (float *)x : nullptr;
    int32_t *d_cmap = ms->constriction_grid ? (int32_t *)(x + sizeof(float) * (size_t)sp) : nullptr;
    const double cells[3] = {3.0, 4.0, 5.0};
    Tune best;
    float best_ms = 1e30f;
    for (int cube = 0; cube < 2; cube++)
        for (double cs : cells) {
            Tune t;
            t.cell_size = cs;
            t.cube = cube != 0;
            std::unique_ptr<ag_plan> p(plan_build(c, ns, srec.data(), sgrid.data(), ms, tab, potential, true, t));
            float ms_near = 0.f;
            for (int rep = 0; rep < 2; rep++) {
                run_checked(p.get(), [&] { enqueue(p.get(), d_maps, d_fmap, d_cmap, true, true); });
                HIP_CHECK(hipEventElapsedTime(&ms_near, c->ev[1], c->ev[2]));
            }
            if (ms_near < best_ms) {
                best_ms = ms_near;
                best = t;
            }
        }
    *spent_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return best;
}

ag_plan *plan_create(ag_ctx *c, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                     const ag_tables *tab, int potential) {
    return plan_build(c, 1, rec, grid, ms, tab, potential, false, Tune());
}

ag_plan *batch_create(ag_ctx *c, int n, const ag_receptor *recs, const ag_grid *grids, const ag_mapset *ms,
                      const ag_tables *tab, int potential, const ag_tune *tune) {
    Tune t;
    double spent = 0.0;
    if (tune) {
        if (!(tune->cell_size >= 2.0 && tune->cell_size <= 8.0))
            ag::fail(AG_ERR_BAD_INPUT, "cell size must lie in [2, 8] A");
        t.cell_size = tune->cell_size;
        t.cube = tune->cube != 0;
    } else {
        for (int i = 0; i < n; i++) validate(&recs[i], &grids[i], ms, tab, potential == AG_POTENTIAL_VINA);
        t = autotune(c, n, recs, grids, ms, tab, potential, &spent);
    }
    ag_plan *p = plan_build(c, n, recs, grids, ms, tab, potential, true, t);
    p->tune_ms = spent;
    return p;
}

void plan_tune(const ag_plan *p, ag_tune *out) {
    out->cell_size = p->tune.cell_size;
    out->cube = p->tune.cube ? 1 : 0;
    out->reserved = 0;
    out->tune_ms = p->tune_ms;
}

void plan_layout(const ag_plan *p, long long *map_off, long long *pt_off, long long *map_floats, long long *points) {
    for (int i = 0; i < p->nconf; i++) {
        if (map_off) map_off[i] = p->confs[i].map_off;
        if (pt_off) pt_off[i] = p->confs[i].pt_off;
    }
    if (map_floats) *map_floats = p->P_total * p->n_maps;
    if (points) *points = p->P_total;
}

int plan_count(const ag_plan *p) { return p->nconf; }

void plan_destroy(ag_plan *p) {
    if (!p) return;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    (void)hipStreamSynchronize(p->ctx->stream2);
    (void)hipStreamSynchronize(p->ctx->copy_stream);
    delete p;
}

// One map set (all conformations of the plan) into d_maps
// ([sum n_maps * P] float, device) and, when asked for, d_fmap / d_cmap
// ([sum P]), enqueued on the context's stream.  A single plan runs its
// kernels directly with phase timing; a batch replays captured hipGraphs
// (re-captured when the outputs move): the cell list and near field, then
// the persistent far field in two parts (conformations split at
// far_split).  Either way the context's near_ev is recorded once the near
// field is done; a batch records far_ev after its first far-field part.
// The second part goes on stream2 behind near_ev, submitted after the
// first: the first part's persistent grid fills the GPU, and the second
// part's workgroups take the CUs as the first part's drain, instead of
// starting after its last one.  The parts write different conformations
// and take separate work counters; the main stream joins the second part
// before ev[3].
static hipGraphExec_t capture(ag_plan *p, float *d_maps, float *d_fmap, int32_t *d_cmap, int part, int c0, int c1,
                              int queue_base) {
    hipStream_t st = p->ctx->stream;
    hipGraph_t g = nullptr;
    HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        enqueue(p, d_maps, d_fmap, d_cmap, false, false, {{c0, c1}}, nullptr, false, part, queue_base);
    } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(st, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
    }
    HIP_CHECK(hipStreamEndCapture(st, &g));
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess)
        ag::fail(AG_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e) + " at hipGraphInstantiate");
    return exec;
}

static void launch_mapset(ag_plan *p, float *d_maps, float *d_fmap, int32_t *d_cmap) {
    ag_ctx *c = p->ctx;
    hipStream_t st = c->stream;
    if (!p->persistent) {
        enqueue(p, d_maps, d_fmap, d_cmap, true, false, true);
        return;
    }
    if (!p->graph || p->g_maps != d_maps || p->g_fmap != d_fmap || p->g_cmap != d_cmap) {
        p->drop_graphs();
        p->graph = capture(p, d_maps, d_fmap, d_cmap, kPartNear, 0, p->nconf, 0);
        p->graph_far = capture(p, d_maps, d_fmap, d_cmap, kPartFar, 0, p->far_split, 0);
        if (p->far_split < p->nconf)
            p->graph_far2 = capture(p, d_maps, d_fmap, d_cmap, kPartFar, p->far_split, p->nconf, 1);
        p->g_maps = d_maps;
        p->g_fmap = d_fmap;
        p->g_cmap = d_cmap;
    }
    HIP_CHECK(hipEventRecord(c->ev[0], st));
    HIP_CHECK(hipGraphLaunch(p->graph, st));
    HIP_CHECK(hipEventRecord(c->near_ev, st));
    HIP_CHECK(hipGraphLaunch(p->graph_far, st));
    HIP_CHECK(hipEventRecord(c->far_ev, st));
    if (p->graph_far2) {
        HIP_CHECK(hipStreamWaitEvent(c->stream2, c->near_ev, 0));
        HIP_CHECK(hipGraphLaunch(p->graph_far2, c->stream2));
        HIP_CHECK(hipEventRecord(c->far2_ev, c->stream2));
        HIP_CHECK(hipStreamWaitEvent(st, c->far2_ev, 0));
    }
    HIP_CHECK(hipEventRecord(c->ev[3], st));
}

static void read_timing(ag_plan *p) {
    ag_ctx *c = p->ctx;
    if (!p->persistent) {
        HIP_CHECK(hipEventElapsedTime(&c->last_ms[0], c->ev[0], c->ev[1]));
        HIP_CHECK(hipEventElapsedTime(&c->last_ms[1], c->ev[1], c->ev[2]));
        HIP_CHECK(hipEventElapsedTime(&c->last_ms[2], c->ev[2], c->ev[3]));
    } else {
        c->last_ms[0] = c->last_ms[1] = c->last_ms[2] = -1.f;
    }
    HIP_CHECK(hipEventElapsedTime(&c->last_ms[3], c->ev[0], c->ev[3]));
}

void plan_execute(ag_plan *p, float *d_maps, float *d_fmap, int32_t *d_cmap) {
    if ((p->want_fmap && !d_fmap) || (p->want_cmap && !d_cmap))
        ag::fail(AG_ERR_BAD_INPUT, "the map set asks for a floating/constriction grid but no output was given");
    HIP_CHECK(hipSetDevice(p->ctx->device));
    run_checked(p, [&] { launch_mapset(p, d_maps, d_fmap, d_cmap); });
    read_timing(p);
}

// plan_execute, then every conformation's outputs into host memory
// (maps_out[i]: [n_maps][P_i] floats; fmap_out[i] / cmap_out[i]: [P_i]).
// Only the e map is written after the near field (by the far kernel and the
// tree build's collision rule, which also sets fmap), so the copy stream
// takes the other maps, cmap and -- without the collision rule -- fmap as
// soon as near_ev is reached, while the far kernel runs; the e maps (and
// fmap) of a batch's first far-field graph follow far_ev, during its second
// far-field graph, and the rest the end of the map set.  Host memory pinned
// with hipHostMalloc lets the copies overlap the kernels.
void plan_execute_host(ag_plan *p, float *d_maps, float *d_fmap, int32_t *d_cmap, float *const *maps_out,
                       float *const *fmap_out, int32_t *const *cmap_out) {
    if ((p->want_fmap && !d_fmap) || (p->want_cmap && !d_cmap))
        ag::fail(AG_ERR_BAD_INPUT, "the map set asks for a floating/constriction grid but no output was given");
    if (!maps_out) ag::fail(AG_ERR_BAD_INPUT, "no map outputs");
    for (int i = 0; i < p->nconf; i++)
        if (!maps_out[i] || (p->want_fmap && (!fmap_out || !fmap_out[i])) ||
            (p->want_cmap && (!cmap_out || !cmap_out[i])))
            ag::fail(AG_ERR_BAD_INPUT, "missing output for a conformation");
    ag_ctx *c = p->ctx;
    HIP_CHECK(hipSetDevice(c->device));
    hipStream_t st = c->stream, cs = c->copy_stream;
    const int e = p->e_index;
    const bool fmap_late = p->want_collide;
    auto copy = [&](void *dst, const void *src, size_t bytes) {
        if (bytes) HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, cs));
    };
    run_checked(p, [&] {
        launch_mapset(p, d_maps, d_fmap, d_cmap);
        HIP_CHECK(hipStreamWaitEvent(cs, c->near_ev, 0));
        for (int i = 0; i < p->nconf; i++) {
            const ConfDev &cf = p->confs[i];
            const size_t pb = sizeof(float) * (size_t)cf.P;
            const float *src = d_maps + cf.map_off;
            if (e < 0) {
                copy(maps_out[i], src, pb * p->n_maps);
            } else {
                copy(maps_out[i], src, pb * e);
                copy(maps_out[i] + (size_t)(e + 1) * cf.P, src + (size_t)(e + 1) * cf.P, pb * (p->n_maps - e - 1));
            }
            if (d_cmap) copy(cmap_out[i], d_cmap + cf.pt_off, sizeof(int32_t) * (size_t)cf.P);
            if (d_fmap && !fmap_late) copy(fmap_out[i], d_fmap + cf.pt_off, pb);
        }
        // the e maps (and late fmaps) of a batch's first far-field graph,
        // then of the rest
        const auto late = [&](int c0, int c1) {
            for (int i = c0; i < c1; i++) {
                const ConfDev &cf = p->confs[i];
                const size_t pb = sizeof(float) * (size_t)cf.P;
                if (e >= 0) copy(maps_out[i] + (size_t)e * cf.P, d_maps + cf.map_off + (size_t)e * cf.P, pb);
                if (d_fmap && fmap_late) copy(fmap_out[i], d_fmap + cf.pt_off, pb);
            }
        };
        const int split = p->persistent ? p->far_split : p->nconf;
        if (split < p->nconf) {
            HIP_CHECK(hipStreamWaitEvent(cs, c->far_ev, 0));
            late(0, split);
        }
        HIP_CHECK(hipEventRecord(c->chunk_ev[0], st));
        HIP_CHECK(hipStreamWaitEvent(cs, c->chunk_ev[0], 0));
        late(split < p->nconf ? split : 0, p->nconf);
        HIP_CHECK(hipEventRecord(c->chunk_ev[kMaxChunks], cs));
        HIP_CHECK(hipStreamWaitEvent(st, c->chunk_ev[kMaxChunks], 0));
    });
    read_timing(p);
}

// One map set, handed to the host map by map: each map is copied into host
// memory as soon as the kernels are done and on_ready(j) is called for it
// (j = n_maps for fmap, n_maps + 1 for cmap), so the caller can format and
// write map j while later maps are still in flight.
// ag_maps_plain with device outputs, every device buffer drawn from this
// device's scratch array: a plan on an arena pool over it, built, run and
// dropped.  Nothing is allocated on the device or pinned on the host.
void maps_plain_noalloc(ag_ctx *c, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                        const ag_tables *tab, float *maps, float *fmap, int32_t *cmap) {
    HIP_CHECK(hipSetDevice(c->device));
    if (!c->scratch) ag::fail(AG_ERR_DEVICE, "no scratch address for this device");
    auto arena = std::make_shared<DevPool>(c->scratch, (size_t)AG_NOALLOC_SCRATCH_BYTES);
    std::unique_ptr<ag_plan, void (*)(ag_plan *)> p(
        plan_build(c, 1, rec, grid, ms, tab, AG_POTENTIAL_AD4, false, Tune(), arena), plan_destroy);
    plan_execute(p.get(), maps, ms->floating_grid ? fmap : nullptr, ms->constriction_grid ? cmap : nullptr);
}

void maps_plain_streamed(ag_ctx *c, const ag_receptor *rec, const ag_grid *grid, const ag_mapset *ms,
                         const ag_tables *tab, float *maps_out, float *fmap, int32_t *cmap,
                         const std::function<void(int)> &on_ready, int potential) {
    std::unique_ptr<ag_plan, void (*)(ag_plan *)> p(plan_create(c, rec, grid, ms, tab, potential), plan_destroy);
    const long long P = p->P_total;
    const size_t pbytes = sizeof(float) * (size_t)P;
    const size_t bytes = pbytes * (size_t)p->n_maps;
    float *d_maps = (float *)c->maps.get(bytes + 2 * pbytes + 256);
    char *extra = (char *)d_maps + bytes;
    float *d_fmap = ms->floating_grid ? (float *)extra : nullptr;
    int32_t *d_cmap = ms->constriction_grid ? (int32_t *)(extra + pbytes) : nullptr;
    plan_execute(p.get(), d_maps, d_fmap, d_cmap);
    for (int j = 0; j < p->n_maps; j++) {
        HIP_CHECK(hipMemcpyAsync(maps_out + (size_t)j * P, d_maps + (size_t)j * P, pbytes,
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
    if (p->nconf != 1) ag::fail(AG_ERR_UNSUPPORTED, "the bin audit runs on single-conformation plans");
    if (!p->far_atoms_ready) {
        float4 *d_far = (float4 *)p->far_atoms.get(sizeof(float4) * (size_t)std::max(p->N, 1));
        if (p->N > 0)
            k_far_atoms<<<(p->N + 255) / 256, 256, 0, st>>>(p->N, (const double *)p->coord.ptr,
                                                            (const double *)p->charge.ptr, p->coeff_estat, d_far);
        p->far_atoms_ready = true;
    }
    DevBuf acc;
    unsigned long long *d_out = (unsigned long long *)acc.get(4 * sizeof(unsigned long long));
    HIP_CHECK(hipMemsetAsync(d_out, 0, 4 * sizeof(unsigned long long), st));
    const ConfDev &cf = p->confs[0];
    AuditParams ap;
    memset(&ap, 0, sizeof ap);
    for (int d = 0; d < 3; d++) {
        ap.n1[d] = cf.n1[d];
        ap.ne[d] = cf.ne[d];
    }
    ap.spacing = cf.spacing;
    ap.s128 = cf.s128;
    ap.far_atoms = (const float4 *)p->far_atoms.ptr;
    ap.coord = (const double *)p->coord.ptr;
    ap.n_atoms = p->N;
    ap.P = cf.P;
    ap.out = d_out;
    k_audit<<<(unsigned)((cf.P + 255) / 256), 256, 0, st>>>(ap);
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
    const long long P = p->P_total;
    const size_t bytes = sizeof(float) * (size_t)p->n_maps * P;
    const size_t pbytes = sizeof(float) * (size_t)P;
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
    if (outputs_on_device) {
        plan_execute(p.get(), d_maps, d_fmap, d_cmap);
        return;
    }
    // host outputs: the maps other than e go to the host while the far
    // kernel runs, the e map after it
    float *const mo[1] = {maps_out};
    float *const fo[1] = {fmap};
    int32_t *const co[1] = {cmap};
    plan_execute_host(p.get(), d_maps, d_fmap, d_cmap, mo, d_fmap ? fo : nullptr, d_cmap ? co : nullptr);
}

static bool is_pinned(const void *ptr) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, ptr) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

// Batch with host outputs: maps_out[i] receives conformation i's
// [n_maps][P_i] floats (fmap_out / cmap_out its [P_i] values when asked).
// When every destination is pinned host memory the conformations run in
// chunks of about equal work, launched directly and alternating between
// the two compute streams, and each chunk's maps are copied to the host on
// the copy stream behind an event while later chunks compute.  (A captured graph would run its
// copy nodes in line with its kernels, serialising the two; a one-shot
// batch has few launches to save, so it is not captured.)  Otherwise the
// plan's graph computes everything and the copies follow it.
void maps_batch(ag_ctx *c, int n, const ag_receptor *recs, const ag_grid *grids, const ag_mapset *ms,
                const ag_tables *tab, float *const *maps_out, float *const *fmap_out, int32_t *const *cmap_out,
                int potential, const ag_tune *tune) {
    if (!maps_out) ag::fail(AG_ERR_BAD_INPUT, "no map outputs");
    for (int i = 0; i < n; i++)
        if (!maps_out[i] || (ms->floating_grid && (!fmap_out || !fmap_out[i])) ||
            (ms->constriction_grid && (!cmap_out || !cmap_out[i])))
            ag::fail(AG_ERR_BAD_INPUT, "missing output for a conformation");
    std::unique_ptr<ag_plan, void (*)(ag_plan *)> p(batch_create(c, n, recs, grids, ms, tab, potential, tune),
                                                   plan_destroy);
    const long long P = p->P_total;
    const size_t bytes = sizeof(float) * (size_t)p->n_maps * P;
    const size_t pbytes = sizeof(float) * (size_t)P;
    float *d_maps = (float *)c->maps.get(bytes + 2 * pbytes + 256);
    char *base = (char *)d_maps + bytes;
    float *d_fmap = ms->floating_grid ? (float *)base : nullptr;
    int32_t *d_cmap = ms->constriction_grid ? (int32_t *)(base + pbytes) : nullptr;
    bool pinned = true;
    for (int i = 0; i < n && pinned; i++)
        pinned = is_pinned(maps_out[i]) && (!d_fmap || is_pinned(fmap_out[i])) && (!d_cmap || is_pinned(cmap_out[i]));
    hipStream_t st = c->stream;
    if (pinned) {
        const HostOut host{maps_out, fmap_out, cmap_out};
        // 10 chunks, the last 3 halving: on the 64-conformation ensemble this
        // measured 112-114 ms of GPU time against 116-117 ms for 16 equal chunks
        const auto chunks = plan_chunks(*p, kBatchChunks, kBatchTaper);
        run_checked(p.get(), [&] {
            HIP_CHECK(hipEventRecord(c->ev[0], st));
            enqueue(p.get(), d_maps, d_fmap, d_cmap, false, false, chunks, &host);
            HIP_CHECK(hipEventRecord(c->ev[3], st));
        });
        HIP_CHECK(hipEventElapsedTime(&c->last_ms[3], c->ev[0], c->ev[3]));
        c->last_ms[0] = c->last_ms[1] = c->last_ms[2] = -1.f;
        return;
    }
    plan_execute(p.get(), d_maps, d_fmap, d_cmap);
    for (int i = 0; i < n; i++) {
        const ConfDev &cf = p->confs[i];
        HIP_CHECK(hipMemcpyAsync(maps_out[i], d_maps + cf.map_off, sizeof(float) * (size_t)p->n_maps * cf.P,
                                 hipMemcpyDeviceToHost, st));
        if (d_fmap)
            HIP_CHECK(hipMemcpyAsync(fmap_out[i], d_fmap + cf.pt_off, sizeof(float) * (size_t)cf.P,
                                     hipMemcpyDeviceToHost, st));
        if (d_cmap)
            HIP_CHECK(hipMemcpyAsync(cmap_out[i], d_cmap + cf.pt_off, sizeof(int32_t) * (size_t)cf.P,
                                     hipMemcpyDeviceToHost, st));
    }
    HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace agdev
