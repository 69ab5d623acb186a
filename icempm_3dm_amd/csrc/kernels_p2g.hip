// kernels_p2g.hip — particle-to-grid transfer (partition_kernel_p2g,
// mpm_kernels.cuh:20-100) for brick-sorted particles.
//
// One workgroup handles one brick segment (the particles whose cell lay in
// one 4x4x4-cell brick at the last sort) and owns an 8x8x8-node LDS tile:
// the brick's 6^3-node stencil footprint plus one node of margin for
// particles that drifted since the sort. Per 256-particle sub-chunk:
//   phase 1  one lane per particle: stress, APIC affine matrix, weights ->
//            LDS record (22 doubles);
//   phase 2  the sub-chunk is cut into 28 slices, one per 9-lane group
//            (7 groups per wave); lane (ox,oy) of a group owns the column of
//            3 stencil nodes (ox,oy,0..2) and sums their mass and momentum
//            over consecutive particles sharing a base node in registers,
//            flushing to the tile with ds_add_f64 when the base changes.
//            Each lane reads 18 doubles per particle for 3 nodes;
//   phase 2b a particle outside the tile (drifted more than one cell since
//            the sort) is taken by a half-wave alone and its 27 nodes go
//            straight to HBM with FP64 atomics;
//   flush    non-zero tile nodes go to HBM with one FP64 atomic per value
//            and mark their grid bricks active.
// Grid dimensions are limited to 1024 nodes per axis (10-bit packing).
#include "constitutive.h"
#include "kernels.h"

namespace icempm {

namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 256;
constexpr int kTile = 8;
constexpr int kTileNodes = kTile * kTile * kTile;
constexpr int kRec = 22; // m | wx[3] wy[3] wz[3] | a[3] | Mh[9]

// Appends brick b to the active list once (its flag guards the append).
struct BrickList {
    int32_t *flag;
    int32_t *list;
    int32_t *count;
};

__device__ __forceinline__ void mark_brick(const BrickList &bl, long long b) {
    if (bl.flag[b] == 0 && atomicCAS(&bl.flag[b], 0, 1) == 0) {
        int slot = atomicAdd(bl.count, 1);
        bl.list[slot] = (int32_t)b;
    }
}

// Column accumulator of one group lane: mass and momentum of the stencil
// nodes (ox, oy, 0..2) of the current base node.
struct Column {
    double m[3], v0[3], v1[3], v2[3];
};

__device__ __forceinline__ void column_zero(Column &c) {
#pragma unroll
    for (int k = 0; k < 3; k++) c.m[k] = c.v0[k] = c.v1[k] = c.v2[k] = 0.0;
}

// Adds particle j of the LDS records: W (m, a + Mh o) for o = (ox, oy, k).
__device__ __forceinline__ void column_add(const double (*rec)[kChunk], int j, int ox, int oy, double dox,
                                           double doy, Column &c) {
    const double wxy = rec[1 + ox][j] * rec[4 + oy][j];
    const double m = rec[0][j];
    double base[3];
#pragma unroll
    for (int r = 0; r < 3; r++) base[r] = rec[10 + r][j] + rec[13 + 3 * r][j] * dox + rec[14 + 3 * r][j] * doy;
    const double mz[3] = {rec[15][j], rec[18][j], rec[21][j]}; // Mh(r, 2)
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double W = wxy * rec[7 + k][j];
        c.m[k] += W * m;
        const double u0 = (k == 0) ? base[0] : base[0] + mz[0] * (double)k;
        const double u1 = (k == 0) ? base[1] : base[1] + mz[1] * (double)k;
        const double u2 = (k == 0) ? base[2] : base[2] + mz[2] * (double)k;
        c.v0[k] += W * u0;
        c.v1[k] += W * u1;
        c.v2[k] += W * u2;
    }
}

__global__ __launch_bounds__(kThreads) void k_p2g(KParams kp, PtsView pts, SegInfo seg, GridState gs, int cur,
                                                  int *err) {
    __shared__ double s_tile[4][kTileNodes];
    __shared__ double s_rec[kRec][kChunk];
    __shared__ int s_base[kChunk];   // tile-relative base (packed), -1 none, -2 overflow
    __shared__ int s_abs[kChunk];    // absolute base of overflow particles (packed 10:10:10 bits)
    __shared__ int s_over[kChunk];
    __shared__ int s_wave_count[kThreads / 64];
    __shared__ int s_nover;
    __shared__ unsigned s_brickmask;

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    const int nseg = *seg.nseg;
    // select, not index: a runtime index into the by-value struct would spill it
    const BrickList bl{gs.brick_flag, cur ? gs.brick_list[1] : gs.brick_list[0], gs.brick_count + (cur ? 1 : 0)};
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    // a contiguous range of segments per logical block
    const long long s_lo = (long long)nseg * lb / gridDim.x;
    const long long s_hi = (long long)nseg * (lb + 1) / gridDim.x;

    // phase-2 lane roles: 7 groups of 9 lanes per wave; lane (ox,oy) owns a node column
    const int gq = lane % 9;
    const bool group_lane = lane < 63;
    const int group = wave * 7 + lane / 9; // 0..27 for group lanes
    const int ox = gq / 3, oy = gq % 3;
    const double dox = ox, doy = oy;
    constexpr int kGroups = (kThreads / 64) * 7;

    for (long long s = s_lo; s < s_hi; s++) {
        const int begin = seg.seg_start[s];
        const int end = seg.seg_start[s + 1];
        const int brick = seg.seg_brick[s];
        const int bz = brick % kp.nbz;
        const int by = (brick / kp.nbz) % kp.nby;
        const int bx = brick / (kp.nbz * kp.nby);
        const int org[3] = {bx * 4 - 2, by * 4 - 2, bz * 4 - 2};

        for (int k = t; k < 4 * kTileNodes; k += kThreads) (&s_tile[0][0])[k] = 0.0;
        if (t == 0) s_brickmask = 0u;
        __syncthreads();

        for (int c0 = begin; c0 < end; c0 += kChunk) {
            // ---------------- phase 1: one lane per particle -> LDS record
            const int i = c0 + t;
            int packed = -1;
            if (i < end) {
                const long long util = double_to_bits(pts.at(ICEMPM_IDX_UTILITY, i));
                if (!(util & ICEMPM_STATUS_DISABLED)) {
                    const bool liquid = util & ICEMPM_STATUS_LIQUID;
                    const double m = liquid ? kp.mass_liquid : kp.mass_solid;
                    double pos[3];
#pragma unroll
                    for (int a = 0; a < 3; a++) pos[a] = pts.at(ICEMPM_POSX + a, i);
                    int base[3];
                    double fx[3];
                    const int code = locate(kp, pos, base, fx);
                    if (code) {
                        raise_error(err, code);
                    } else {
                        double sub[9];
                        {
                            double PFt[9];
                            if (liquid) {
                                water_stress(kp, pts.at(ICEMPM_IDX_JP_INV, i), PFt);
                            } else {
                                double Fe[9];
#pragma unroll
                                for (int k = 0; k < 9; k++) Fe[k] = pts.at(ICEMPM_FE00 + k, i);
                                kirchhoff_stress(kp, Fe, PFt);
                            }
#pragma unroll
                            for (int k = 0; k < 9; k++) sub[k] = m * pts.at(ICEMPM_BP00 + k, i) - kp.dt_vol_Dpinv * PFt[k];
                        }
                        // W (m v + M (o - f) h) = W (a + Mh o), a = m v - M f h, Mh = M h
                        const double h = kp.h;
                        const double fh[3] = {fx[0] * h, fx[1] * h, fx[2] * h};
                        s_rec[0][t] = m;
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            double wd[3];
                            bspline3(fx[d], wd);
                            s_rec[1 + 3 * d][t] = wd[0];
                            s_rec[2 + 3 * d][t] = wd[1];
                            s_rec[3 + 3 * d][t] = wd[2];
                        }
#pragma unroll
                        for (int r = 0; r < 3; r++) {
                            const double vr = pts.at(ICEMPM_VELX + r, i);
                            s_rec[10 + r][t] =
                                vr * m - (sub[3 * r] * fh[0] + sub[3 * r + 1] * fh[1] + sub[3 * r + 2] * fh[2]);
#pragma unroll
                            for (int c = 0; c < 3; c++) s_rec[13 + 3 * r + c][t] = sub[3 * r + c] * h;
                        }
                        const int rx = base[0] - org[0], ry = base[1] - org[1], rz = base[2] - org[2];
                        const bool fits = (unsigned)rx <= kTile - 3 && (unsigned)ry <= kTile - 3 &&
                                          (unsigned)rz <= kTile - 3;
                        if (fits) {
                            packed = rx | (ry << 3) | (rz << 6);
                        } else {
                            packed = -2;
                            s_abs[t] = base[0] | (base[1] << 10) | (base[2] << 20);
                        }
                    }
                }
            }
            s_base[t] = packed;
            __syncthreads();

            // ---------------- list of particles outside the tile
            const bool over = packed == -2;
            const unsigned long long om = __ballot(over);
            if (lane == 0) s_wave_count[wave] = __popcll(om);
            __syncthreads();
            {
                int offo = 0;
                for (int w = 0; w < wave; w++) offo += s_wave_count[w];
                if (over) s_over[offo + __popcll(om & ((1ull << lane) - 1ull))] = t;
                if (t == 0) {
                    int to = 0;
                    for (int w = 0; w < kThreads / 64; w++) to += s_wave_count[w];
                    s_nover = to;
                }
            }
            __syncthreads();

            // ---------------- phase 2: group slices, lane per node column
            {
                const int nchunk = min(kChunk, end - c0);
                const int per = (nchunk + kGroups - 1) / kGroups;
                const int j_lo = group_lane ? min(group * per, nchunk) : nchunk;
                const int j_hi = min(j_lo + per, nchunk);
                Column col;
                column_zero(col);
                int cur_pk = -1;
                for (int j = j_lo; j < j_hi; j++) {
                    const int pk = s_base[j];
                    if (pk < 0) continue;
                    if (pk != cur_pk) {
                        if (cur_pk >= 0) {
                            const int tn = ((((cur_pk & 7) + ox) * kTile + ((cur_pk >> 3) & 7) + oy) * kTile) + ((cur_pk >> 6) & 7);
#pragma unroll
                            for (int k = 0; k < 3; k++) {
                                atomicAdd(&s_tile[0][tn + k], col.m[k]);
                                atomicAdd(&s_tile[1][tn + k], col.v0[k]);
                                atomicAdd(&s_tile[2][tn + k], col.v1[k]);
                                atomicAdd(&s_tile[3][tn + k], col.v2[k]);
                            }
                            column_zero(col);
                        }
                        cur_pk = pk;
                    }
                    column_add(s_rec, j, ox, oy, dox, doy, col);
                }
                if (cur_pk >= 0) {
                    const int tn = ((((cur_pk & 7) + ox) * kTile + ((cur_pk >> 3) & 7) + oy) * kTile) + ((cur_pk >> 6) & 7);
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        atomicAdd(&s_tile[0][tn + k], col.m[k]);
                        atomicAdd(&s_tile[1][tn + k], col.v0[k]);
                        atomicAdd(&s_tile[2][tn + k], col.v1[k]);
                        atomicAdd(&s_tile[3][tn + k], col.v2[k]);
                    }
                }
            }
            // ---------------- phase 2b: particles outside the tile go straight to HBM
            const int nover = s_nover;
            for (int r = group_lane ? group : nover; r < nover; r += kGroups) {
                const int j = s_over[r];
                const int ab = s_abs[j];
                Column col;
                column_zero(col);
                column_add(s_rec, j, ox, oy, dox, doy, col);
                const int x = (ab & 1023) + ox, y = ((ab >> 10) & 1023) + oy, z0 = (ab >> 20) & 1023;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const long long idx = node_index(kp, x, y, z0 + k);
                    unsafeAtomicAdd(&gs.grid[idx], col.m[k]);
                    unsafeAtomicAdd(&gs.grid[kp.grid_pitch + idx], col.v0[k]);
                    unsafeAtomicAdd(&gs.grid[2 * kp.grid_pitch + idx], col.v1[k]);
                    unsafeAtomicAdd(&gs.grid[3 * kp.grid_pitch + idx], col.v2[k]);
                }
                // the corner columns reach every brick the stencil touches
                if (ox != 1 && oy != 1) {
                    mark_brick(bl, brick_of(kp, x, y, z0));
                    mark_brick(bl, brick_of(kp, x, y, z0 + 2));
                }
            }
            __syncthreads();
        }

        // ---------------- flush the tile to HBM
        for (int k = t; k < kTileNodes; k += kThreads) {
            const double m = s_tile[0][k];
            if (m != 0.0) {
                const int tx = k >> 6, ty = (k >> 3) & 7, tz = k & 7;
                const long long idx = node_index(kp, org[0] + tx, org[1] + ty, org[2] + tz);
                unsafeAtomicAdd(&gs.grid[idx], m);
                unsafeAtomicAdd(&gs.grid[kp.grid_pitch + idx], s_tile[1][k]);
                unsafeAtomicAdd(&gs.grid[2 * kp.grid_pitch + idx], s_tile[2][k]);
                unsafeAtomicAdd(&gs.grid[3 * kp.grid_pitch + idx], s_tile[3][k]);
                // tile node t lies in brick (bx - 1) + ((t + 2) >> 2) along each axis
                const int dbx = (tx + 2) >> 2, dby = (ty + 2) >> 2, dbz = (tz + 2) >> 2;
                atomicOr(&s_brickmask, 1u << ((dbx * 3 + dby) * 3 + dbz));
            }
        }
        __syncthreads();
        if (t < 27 && ((s_brickmask >> t) & 1u)) {
            const int dbx = t / 9, dby = (t / 3) % 3, dbz = t % 3;
            const long long b = ((long long)(bx - 1 + dbx) * kp.nby + (by - 1 + dby)) * kp.nbz + (bz - 1 + dbz);
            mark_brick(bl, b);
        }
        __syncthreads();
    }
}

} // namespace

hipError_t launch_p2g(const KParams &kp, PtsView pts, int n, SegInfo seg, GridState gs, int cur, int *err,
                      int n_blocks, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    n_blocks = (n_blocks + kXcd - 1) / kXcd * kXcd;
    k_p2g<<<n_blocks, kThreads, 0, stream>>>(kp, pts, seg, gs, cur, err);
    return hipGetLastError();
}

} // namespace icempm
