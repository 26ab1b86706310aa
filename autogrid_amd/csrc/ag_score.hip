// Ligand scoring against a map set on gfx950: the consumer side of the maps.
//
// AutoDock reads AutoGrid maps by trilinear interpolation.  The maps are
// [n_maps][n1z][n1y][n1x] floats, x fastest (main.cpp:322), on the grid
// described by the .fld file (main.cpp:1903-1956): point (i, j, k) sits at
// ((i - ne_x) s, (j - ne_y) s, (k - ne_z) s) in the grid frame.  A ligand
// atom of map index t and charge q at x scores
//     E(x) = sum over the 8 corners c of w_c(x) (M_t[c] + q M_e[c] + |q| M_d[c]),
// w_c the trilinear weights of x in its cell, and a pose scores the sum of
// its atoms.  An atom outside the grid scores a fixed penalty with zero
// gradient.
//
// Kernels:
//   k_score      one workgroup per pose, lanes over its atoms: per-atom
//                energy and gradient dE/dx, the pose's sum reduced in FP64
//                in a fixed order;
//   k_score_adj  dL/dmaps = S^T dL/dscores for the scores' sparse matrix S,
//                one lane per touched map entry over a host-built,
//                index-sorted list of its contributions: a fixed-order sum,
//                no atomics.

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "ag_devcommon.h"
#include "ag_internal.h"

namespace {

constexpr int kST = 256;

struct __align__(16) ScoreAtom {   // 32 bytes
    float gx, gy, gz;   // position in grid units: (x + ne s) / s
    float q;
    int map;            // map index of the atom's type, -1 for none
    int pose;
    int inside;         // 1 when the 8 corners lie on the grid
    int pad;
};

struct ScoreParams {
    const ScoreAtom *atoms;
    const int *pose_start;      // [n_poses + 1]
    const float *maps;
    long long P;
    int n1x, n1y, n1z;
    int e_slot, d_slot;
    float inv_spacing;
    float outside;
    float *scores;              // [n_poses]
    float *grad;                // [n_atoms][3] or null
};

__global__ void __launch_bounds__(kST) k_score(ScoreParams p) {
    __shared__ double red[kST];
    const int pose = blockIdx.x;
    const int a0 = p.pose_start[pose], a1 = p.pose_start[pose + 1];
    double part = 0.0;
    for (int a = a0 + (int)threadIdx.x; a < a1; a += kST) {
        const ScoreAtom at = p.atoms[a];
        float e = p.outside, g0 = 0.f, g1 = 0.f, g2 = 0.f;
        if (at.inside) {
            const int i0 = (int)floorf(at.gx), j0 = (int)floorf(at.gy), k0 = (int)floorf(at.gz);
            const float fx = at.gx - (float)i0, fy = at.gy - (float)j0, fz = at.gz - (float)k0;
            const float aq = fabsf(at.q);
            e = 0.f;
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
                const long long idx =
                    (long long)(i0 + dx) + (long long)p.n1x * ((long long)(j0 + dy) + (long long)p.n1y * (k0 + dz));
                float v = 0.f;
                if (at.map >= 0) v = p.maps[(long long)at.map * p.P + idx];
                if (p.e_slot >= 0) v = __builtin_fmaf(at.q, p.maps[(long long)p.e_slot * p.P + idx], v);
                if (p.d_slot >= 0) v = __builtin_fmaf(aq, p.maps[(long long)p.d_slot * p.P + idx], v);
                const float wx = dx ? fx : 1.f - fx, wy = dy ? fy : 1.f - fy, wz = dz ? fz : 1.f - fz;
                const float sx = dx ? 1.f : -1.f, sy = dy ? 1.f : -1.f, sz = dz ? 1.f : -1.f;
                e = __builtin_fmaf(wx * wy * wz, v, e);
                g0 = __builtin_fmaf(sx * wy * wz, v, g0);
                g1 = __builtin_fmaf(wx * sy * wz, v, g1);
                g2 = __builtin_fmaf(wx * wy * sz, v, g2);
            }
        }
        if (p.grad) {
            p.grad[3 * (size_t)a] = g0 * p.inv_spacing;
            p.grad[3 * (size_t)a + 1] = g1 * p.inv_spacing;
            p.grad[3 * (size_t)a + 2] = g2 * p.inv_spacing;
        }
        part += (double)e;
    }
    red[threadIdx.x] = part;
    __syncthreads();
    for (int s = kST / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) p.scores[pose] = (float)red[0];
}

__global__ void k_score_zero(float *out, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = 0.f;
}

// row r: dmaps[target[r]] = sum over k in [row_start[r], row_start[r+1]) of
// val[k] * dscore[pose[k]], k in the order the host listed them
__global__ void k_score_adj(int n_rows, const int *row_start, const long long *target, const int *pose,
                            const float *val, const float *dscore, float *dmaps) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    double s = 0.0;
    for (int k = row_start[r]; k < row_start[r + 1]; k++) s += (double)val[k] * (double)dscore[pose[k]];
    dmaps[target[r]] = (float)s;
}

}  // namespace

struct ag_scorer {
    ag_ctx *ctx = nullptr;
    int n1[3] = {0, 0, 0}, ne[3] = {0, 0, 0};
    double spacing = 0.0;
    long long P = 0;
    int n_maps = 0, e_slot = -1, d_slot = -1;
    int n_poses = 0, n_atoms = 0;
    float outside = 0.f;
    std::vector<int> pose_start;
    std::vector<int> map_index;
    std::vector<double> charge;
    int n_rows = 0;
    DevBuf atoms, pose_start_d, row_start, target, pose_of, val;
};

namespace agdev {

static void scorer_positions(ag_scorer *s, const double *xyz) {
    const int n = s->n_atoms;
    std::vector<ScoreAtom> h(std::max(n, 1));
    struct Entry {
        long long target;
        int pose;
        float val;
    };
    std::vector<Entry> ent;
    ent.reserve((size_t)n * 24);
    for (int p = 0; p < s->n_poses; p++)
        for (int a = s->pose_start[p]; a < s->pose_start[p + 1]; a++) {
            ScoreAtom &at = h[a];
            double g[3];
            bool inside = true;
            for (int d = 0; d < 3; d++) {
                if (!std::isfinite(xyz[3 * (size_t)a + d])) ag::fail(AG_ERR_BAD_INPUT, "ligand position is not finite");
                g[d] = (xyz[3 * (size_t)a + d] + s->ne[d] * s->spacing) / s->spacing;
                // the 8 corners must lie on the grid: 0 <= g < n1 - 1
                inside = inside && g[d] >= 0.0 && g[d] < (double)(s->n1[d] - 1);
            }
            at.gx = (float)g[0];
            at.gy = (float)g[1];
            at.gz = (float)g[2];
            // FP32 rounding may carry g onto n1 - 1: keep such atoms outside
            inside = inside && at.gx < (float)(s->n1[0] - 1) && at.gy < (float)(s->n1[1] - 1) &&
                     at.gz < (float)(s->n1[2] - 1);
            at.q = (float)s->charge[a];
            at.map = s->map_index[a];
            at.pose = p;
            at.inside = inside ? 1 : 0;
            at.pad = 0;
            if (!inside) continue;
            // contributions to dL/dmaps, from the same FP32 weights k_score uses
            const int i0 = (int)floorf(at.gx), j0 = (int)floorf(at.gy), k0 = (int)floorf(at.gz);
            const float fx = at.gx - (float)i0, fy = at.gy - (float)j0, fz = at.gz - (float)k0;
            for (int c = 0; c < 8; c++) {
                const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
                const long long idx = (long long)(i0 + dx) +
                                      (long long)s->n1[0] * ((long long)(j0 + dy) + (long long)s->n1[1] * (k0 + dz));
                const float w = (dx ? fx : 1.f - fx) * (dy ? fy : 1.f - fy) * (dz ? fz : 1.f - fz);
                if (at.map >= 0) ent.push_back({(long long)at.map * s->P + idx, p, w});
                if (s->e_slot >= 0) ent.push_back({(long long)s->e_slot * s->P + idx, p, w * at.q});
                if (s->d_slot >= 0) ent.push_back({(long long)s->d_slot * s->P + idx, p, w * std::fabs(at.q)});
            }
        }
    std::stable_sort(ent.begin(), ent.end(), [](const Entry &x, const Entry &y) { return x.target < y.target; });
    std::vector<int> rs, pose_of;
    std::vector<long long> tg;
    std::vector<float> val;
    rs.reserve(ent.size() + 1);
    for (size_t k = 0; k < ent.size(); k++) {
        if (k == 0 || ent[k].target != ent[k - 1].target) {
            rs.push_back((int)k);
            tg.push_back(ent[k].target);
        }
        pose_of.push_back(ent[k].pose);
        val.push_back(ent[k].val);
    }
    rs.push_back((int)ent.size());
    if (ent.size() > (size_t)INT32_MAX) ag::fail(AG_ERR_UNSUPPORTED, "too many ligand atoms for one scorer");
    s->n_rows = (int)tg.size();
    hipStream_t st = s->ctx->stream;
    auto up = [&](DevBuf &b, const void *src, size_t bytes) {
        void *d = b.get(std::max<size_t>(bytes, 16));
        if (bytes) HIP_CHECK(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, st));
    };
    up(s->atoms, h.data(), sizeof(ScoreAtom) * (size_t)n);
    up(s->row_start, rs.data(), sizeof(int) * rs.size());
    up(s->target, tg.data(), sizeof(long long) * tg.size());
    up(s->pose_of, pose_of.data(), sizeof(int) * pose_of.size());
    up(s->val, val.data(), sizeof(float) * val.size());
    HIP_CHECK(hipStreamSynchronize(st));   // host vectors go out of scope
}

ag_scorer *scorer_create(ag_ctx *c, const ag_grid *grid, int n_maps, int elec_index, int dsolv_index, int n_poses,
                         const int32_t *pose_start, const double *xyz, const int32_t *map_index,
                         const double *charge, double outside) {
    for (int d = 0; d < 3; d++)
        if (grid->n1[d] < 2 || grid->n1[d] > AG_MAX_GRID_PTS || grid->ne[d] < 0)
            ag::fail(AG_ERR_BAD_INPUT, "grid dimensions out of range (at least 2 points per axis)");
    if (!(grid->spacing > 0)) ag::fail(AG_ERR_BAD_INPUT, "grid spacing must be positive");
    if (n_maps < 1 || n_maps > AG_MAX_MAPS) ag::fail(AG_ERR_BAD_INPUT, "map count out of range");
    if (elec_index >= n_maps || dsolv_index >= n_maps) ag::fail(AG_ERR_BAD_INPUT, "e/d map index out of range");
    if (n_poses < 1) ag::fail(AG_ERR_BAD_INPUT, "no poses");
    if (pose_start[0] != 0) ag::fail(AG_ERR_BAD_INPUT, "pose_start[0] must be 0");
    for (int p = 0; p < n_poses; p++)
        if (pose_start[p + 1] < pose_start[p]) ag::fail(AG_ERR_BAD_INPUT, "pose_start must not decrease");
    const int n = pose_start[n_poses];
    if (n > 0 && (!xyz || !map_index || !charge)) ag::fail(AG_ERR_BAD_INPUT, "ligand arrays missing");
    for (int a = 0; a < n; a++) {
        if (map_index[a] < -1 || map_index[a] >= n_maps) ag::fail(AG_ERR_BAD_INPUT, "ligand map index out of range");
        if (!std::isfinite(charge[a])) ag::fail(AG_ERR_BAD_INPUT, "ligand charge is not finite");
    }
    HIP_CHECK(hipSetDevice(c->device));
    std::unique_ptr<ag_scorer> s(new ag_scorer());
    s->ctx = c;
    for (int d = 0; d < 3; d++) {
        s->n1[d] = grid->n1[d];
        s->ne[d] = grid->ne[d];
    }
    s->spacing = grid->spacing;
    s->P = (long long)grid->n1[0] * grid->n1[1] * grid->n1[2];
    s->n_maps = n_maps;
    s->e_slot = elec_index;
    s->d_slot = dsolv_index;
    s->n_poses = n_poses;
    s->n_atoms = n;
    s->outside = (float)outside;
    s->pose_start.assign(pose_start, pose_start + n_poses + 1);
    s->map_index.assign(map_index, map_index + n);
    s->charge.assign(charge, charge + n);
    void *d = s->pose_start_d.get(sizeof(int) * (size_t)(n_poses + 1));
    HIP_CHECK(hipMemcpyAsync(d, s->pose_start.data(), sizeof(int) * (size_t)(n_poses + 1), hipMemcpyHostToDevice,
                             c->stream));
    scorer_positions(s.get(), xyz);
    return s.release();
}

void scorer_set_positions(ag_scorer *s, const double *xyz) {
    HIP_CHECK(hipSetDevice(s->ctx->device));
    if (s->n_atoms > 0 && !xyz) ag::fail(AG_ERR_BAD_INPUT, "ligand positions missing");
    scorer_positions(s, xyz);
}

void scorer_score(ag_scorer *s, const float *d_maps, float *d_scores, float *d_grad) {
    HIP_CHECK(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    ScoreParams p;
    memset(&p, 0, sizeof p);
    p.atoms = (const ScoreAtom *)s->atoms.ptr;
    p.pose_start = (const int *)s->pose_start_d.ptr;
    p.maps = d_maps;
    p.P = s->P;
    p.n1x = s->n1[0];
    p.n1y = s->n1[1];
    p.n1z = s->n1[2];
    p.e_slot = s->e_slot;
    p.d_slot = s->d_slot;
    p.inv_spacing = (float)(1.0 / s->spacing);
    p.outside = s->outside;
    p.scores = d_scores;
    p.grad = d_grad;
    k_score<<<s->n_poses, kST, 0, st>>>(p);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(st));
}

void scorer_adjoint(ag_scorer *s, const float *d_dscore, float *d_dmaps) {
    HIP_CHECK(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const long long total = s->P * s->n_maps;
    const unsigned zb = (unsigned)std::min<long long>((total + 255) / 256, 65536);
    k_score_zero<<<zb, 256, 0, st>>>(d_dmaps, total);
    if (s->n_rows > 0)
        k_score_adj<<<(s->n_rows + 255) / 256, 256, 0, st>>>(s->n_rows, (const int *)s->row_start.ptr,
                                                             (const long long *)s->target.ptr,
                                                             (const int *)s->pose_of.ptr, (const float *)s->val.ptr,
                                                             d_dscore, d_dmaps);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(st));
}

void scorer_destroy(ag_scorer *s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    delete s;
}

}  // namespace agdev
