// p2g_tile.h — the LDS-tiled particle-to-grid scatter shared by the P2G
// kernel and the fused G2P+P2G kernel (partition_kernel_p2g,
// mpm_kernels.cuh:20-100).
//
// A workgroup (256 threads) owns one brick segment (the points whose cell
// lay in one 4x4x4-cell brick at the last sort) and an 8x8x8-node LDS tile:
// the brick's 6^3-node stencil footprint plus one node of margin for points
// that drifted since the sort. Per 256-point sub-chunk:
//   record   one lane per point: stress, APIC affine matrix M = m C - dt V
//            (4/h^2) tau, weights -> LDS record (22 doubles);
//   scatter  the sub-chunk is cut into 28 slices, one per 9-lane group
//            (7 groups per wave); lane (ox,oy) of a group owns the column of
//            3 stencil nodes (ox,oy,0..2) and sums their mass and momentum
//            over consecutive points sharing a base node in registers,
//            flushing to the tile with ds_add_f64 when the base changes
//            (only the lowest node when the base moves one node up in z,
//            the other two shift down);
//            a point outside the tile is taken by a group alone and its 27
//            nodes go straight to HBM with FP64 atomics;
//   flush    non-zero tile nodes go to HBM with one FP64 atomic per value
//            and mark their grid bricks active.
// W (m v + M (x_i - x_p)) is evaluated as W (a + Mh o), a = m v - M f h,
// Mh = M h, o the node offset. Grid dimensions are limited to 1024 nodes
// per axis (10-bit packing of an absolute base node).
#pragma once

#include "constitutive.h"
#include "kernels.h"

namespace icempm {

// Points per chunk (= threads per workgroup) of the segment kernels. The
// P2G machinery below is written for any multiple of 64 up to 256 (template
// parameter N); the fused kernel also runs with 128-point chunks, whose
// smaller record block lets three workgroups share a CU.
constexpr int kP2GThreads = 256;
constexpr int kP2GChunk = 256;
constexpr int kTile = 8;
constexpr int kTileNodes = kTile * kTile * kTile;
// LDS tile strides in doubles: x stride 67 (= 3 mod 32) and y stride 8 put
// the 9 node columns of a lane group on distinct banks.
constexpr int kTileSY = kTile;
constexpr int kTileSX = 67;
constexpr int kTileAlloc = (kTile - 1) * kTileSX + (kTile - 1) * kTileSY + kTile;
// 9-lane column groups of an N-thread workgroup (7 per wavefront).
template <int N> constexpr int groups_of() { return (N / 64) * 7; }

__device__ __forceinline__ int tile_index(int tx, int ty, int tz) { return tx * kTileSX + ty * kTileSY + tz; }

template <int N>
struct P2GSharedT {
    static_assert(N % 64 == 0 && N >= 64 && N <= 256, "whole wavefronts, at most 256 points per chunk");
    double tile[4][kTileAlloc];
    // The 22 record values of each point slot, m | wx[3] wy[3] wz[3] | a[3] |
    // Mh[9]: the weights wx, wy (each lane of a column group reads its own
    // one) as rows of doubles, the 16 values every lane of the group reads
    // as rows of pairs (m wz0) (wz1 wz2) (a0 a1) (a2 Mh0) ... (Mh7 Mh8), so a
    // lane reads a point with 2 + 8 LDS instructions instead of 18. Odd row
    // length: consecutive rows start on different banks.
    double w[6][N + 1];
    double2 rq[8][N + 1];
    // value r (0..21, in the order above) of slot t
    __device__ __forceinline__ double &rec(int r, int t) {
        if (r >= 1 && r <= 6) return w[r - 1][t];
        if (r == 0) return rq[0][t].x;
        if (r == 7) return rq[0][t].y;
        const int p = (r - 8) / 2 + 1;
        return ((r - 8) & 1) ? rq[p][t].y : rq[p][t].x;
    }
    int base[N];             // tile-relative base node (packed 3:3:3), -1 none, -2 outside the tile
    int absb[N];             // absolute base node of points outside the tile (packed 10:10:10)
    int over[N];             // sub-chunk indices of points outside the tile
    int wave_count[N / 64];
    int nover;
    unsigned brickmask;
};
using P2GShared = P2GSharedT<kP2GChunk>;

// Appends brick b to the active list once (its flag guards the append).
struct BrickList {
    int32_t *flag;
    int32_t *list;
    int32_t *count;
};

// `cur` picks one of the two lists of GridState by select, not by runtime
// index (indexing the by-value struct would spill it to scratch).
__device__ __forceinline__ BrickList brick_list(const GridState &gs, int cur) {
    return BrickList{gs.brick_flag, cur ? gs.brick_list[1] : gs.brick_list[0], gs.brick_count + (cur ? 1 : 0)};
}

__device__ __forceinline__ void mark_brick(const BrickList &bl, long long b) {
    if (bl.flag[b] == 0 && atomicCAS(&bl.flag[b], 0, 1) == 0) {
        int slot = atomicAdd(bl.count, 1);
        bl.list[slot] = (int32_t)b;
    }
}

// Origin (node coordinates) of the tile of the brick with packed coordinates
// `coord` (bx | by << 8 | bz << 16), and its brick coordinates.
__device__ __forceinline__ void tile_origin_packed(int coord, int org[3], int bc[3]) {
    bc[0] = coord & 255;
    bc[1] = (coord >> 8) & 255;
    bc[2] = coord >> 16;
    for (int d = 0; d < 3; d++) org[d] = bc[d] * 4 - 2;
}

// Clears the tile and the brick mask. The mask is cleared by thread 0, a
// lane of wavefront 0 like every reader of the mask (p2g_mark_tile_bricks).
template <int N>
__device__ __forceinline__ void p2g_zero_tile(P2GSharedT<N> &sh, int t) {
    for (int k = t; k < 4 * kTileAlloc; k += N) (&sh.tile[0][0])[k] = 0.0;
    if (t == 0) sh.brickmask = 0u;
}

// Builds the LDS record of point slot t from its state (the body of
// partition_kernel_p2g up to the scatter, mpm_kernels.cuh:27-81); Je = det Fe
// and rc = rcbrt(Je^2) are passed in for solid points. Points with a stencil
// outside the grid raise the error code and get no record.
template <int N>
__device__ __forceinline__ void p2g_record(P2GSharedT<N> &sh, int t, const KParams &kp, bool valid, long long util,
                                           const double pos[3], const double vel[3], const double Fe[9],
                                           double Je, double rc, const double Bp[9], double Jp_inv,
                                           const int org[3], int *err) {
    int packed = -1;
    if (valid && !(util & ICEMPM_STATUS_DISABLED)) {
        const bool liquid = util & ICEMPM_STATUS_LIQUID;
        const double m = liquid ? kp.mass_liquid : kp.mass_solid;
        int base[3];
        double fx[3];
        const int code = locate(kp, pos, base, fx);
        if (code) {
            raise_error(err, code);
        } else {
            double sub[9];
            {
                double PFt[9];
                if (liquid) water_stress(kp, Jp_inv, PFt);
                else kirchhoff_stress_j(kp, Fe, Je, rc, PFt);
#pragma unroll
                for (int k = 0; k < 9; k++) sub[k] = m * Bp[k] - kp.dt_vol_Dpinv * PFt[k];
            }
            const double h = kp.h;
            const double fh[3] = {fx[0] * h, fx[1] * h, fx[2] * h};
            sh.rec(0, t) = m;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                double wd[3];
                bspline3(fx[d], wd);
                sh.rec(1 + 3 * d, t) = wd[0];
                sh.rec(2 + 3 * d, t) = wd[1];
                sh.rec(3 + 3 * d, t) = wd[2];
            }
#pragma unroll
            for (int r = 0; r < 3; r++) {
                sh.rec(10 + r, t) = vel[r] * m - (sub[3 * r] * fh[0] + sub[3 * r + 1] * fh[1] + sub[3 * r + 2] * fh[2]);
#pragma unroll
                for (int c = 0; c < 3; c++) sh.rec(13 + 3 * r + c, t) = sub[3 * r + c] * h;
            }
            const int rx = base[0] - org[0], ry = base[1] - org[1], rz = base[2] - org[2];
            if ((unsigned)rx <= kTile - 3 && (unsigned)ry <= kTile - 3 && (unsigned)rz <= kTile - 3) {
                packed = rx | (ry << 3) | (rz << 6);
            } else {
                packed = -2;
                sh.absb[t] = base[0] | (base[1] << 10) | (base[2] << 20);
            }
        }
    }
    sh.base[t] = packed;
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

// The 18 record values a column lane uses of point j.
struct ColRec {
    double wx, wy, m, wz[3], a[3], M[9];
};

template <int N>
__device__ __forceinline__ void column_load(const P2GSharedT<N> &sh, int j, int ox, int oy, ColRec &r) {
    r.wx = sh.w[ox][j];
    r.wy = sh.w[3 + oy][j];
    const double2 q0 = sh.rq[0][j], q1 = sh.rq[1][j], q2 = sh.rq[2][j], q3 = sh.rq[3][j];
    const double2 q4 = sh.rq[4][j], q5 = sh.rq[5][j], q6 = sh.rq[6][j], q7 = sh.rq[7][j];
    r.m = q0.x;
    r.wz[0] = q0.y;
    r.wz[1] = q1.x;
    r.wz[2] = q1.y;
    r.a[0] = q2.x;
    r.a[1] = q2.y;
    r.a[2] = q3.x;
    r.M[0] = q3.y;
    r.M[1] = q4.x;
    r.M[2] = q4.y;
    r.M[3] = q5.x;
    r.M[4] = q5.y;
    r.M[5] = q6.x;
    r.M[6] = q6.y;
    r.M[7] = q7.x;
    r.M[8] = q7.y;
}

// Adds record r: W (m, a + Mh o) for o = (ox, oy, k).
__device__ __forceinline__ void column_accum(const ColRec &r, double dox, double doy, Column &c) {
    const double m = r.m;
    const double *wz = r.wz, *a = r.a, *M = r.M;
    const double wxy = r.wx * r.wy;
    double base[3];
#pragma unroll
    for (int q = 0; q < 3; q++) base[q] = a[q] + M[3 * q] * dox + M[3 * q + 1] * doy;
    const double mz[3] = {M[2], M[5], M[8]}; // Mh(r, 2)
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double W = wxy * wz[k];
        c.m[k] += W * m;
        const double u0 = (k == 0) ? base[0] : base[0] + mz[0] * (double)k;
        const double u1 = (k == 0) ? base[1] : base[1] + mz[1] * (double)k;
        const double u2 = (k == 0) ? base[2] : base[2] + mz[2] * (double)k;
        c.v0[k] += W * u0;
        c.v1[k] += W * u1;
        c.v2[k] += W * u2;
    }
}

// Adds record j (all 18 values requested before any is used, so one LDS
// round trip serves the point).
template <int N>
__device__ __forceinline__ void column_add(const P2GSharedT<N> &sh, int j, int ox, int oy, double dox, double doy,
                                           Column &c) {
    ColRec r;
    column_load(sh, j, ox, oy, r);
    __builtin_amdgcn_sched_barrier(0);
    column_accum(r, dox, doy, c);
}

template <int N>
__device__ __forceinline__ void column_to_tile(P2GSharedT<N> &sh, int packed, int ox, int oy, const Column &c) {
    const int tn = tile_index((packed & 7) + ox, ((packed >> 3) & 7) + oy, (packed >> 6) & 7);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        atomicAdd(&sh.tile[0][tn + k], c.m[k]);
        atomicAdd(&sh.tile[1][tn + k], c.v0[k]);
        atomicAdd(&sh.tile[2][tn + k], c.v1[k]);
        atomicAdd(&sh.tile[3][tn + k], c.v2[k]);
    }
}

// Adds the column's lowest node (z offset 0) to the tile.
template <int N>
__device__ __forceinline__ void column_node_to_tile(P2GSharedT<N> &sh, int packed, int ox, int oy, const Column &c) {
    const int tn = tile_index((packed & 7) + ox, ((packed >> 3) & 7) + oy, (packed >> 6) & 7);
    atomicAdd(&sh.tile[0][tn], c.m[0]);
    atomicAdd(&sh.tile[1][tn], c.v0[0]);
    atomicAdd(&sh.tile[2][tn], c.v1[0]);
    atomicAdd(&sh.tile[3][tn], c.v2[0]);
}

// Moves the column one node down (for a base one node higher in z).
__device__ __forceinline__ void column_shift(Column &c) {
    c.m[0] = c.m[1];
    c.m[1] = c.m[2];
    c.m[2] = 0.0;
    c.v0[0] = c.v0[1];
    c.v0[1] = c.v0[2];
    c.v0[2] = 0.0;
    c.v1[0] = c.v1[1];
    c.v1[1] = c.v1[2];
    c.v1[2] = 0.0;
    c.v2[0] = c.v2[1];
    c.v2[1] = c.v2[2];
    c.v2[2] = 0.0;
}

// Per wavefront, the number of its points outside the tile (after the
// thread's own p2g_record; the barrier that completes the records also
// publishes the counts).
template <int N>
__device__ __forceinline__ void p2g_count_over(P2GSharedT<N> &sh, int t) {
    const unsigned long long om = __ballot(sh.base[t] == -2);
    if ((t & 63) == 0) sh.wave_count[t >> 6] = __popcll(om);
}

// Scatter of a sub-chunk whose records are complete (call after a barrier;
// ends with a barrier). nchunk: number of record slots used. kCounted: the
// caller ran p2g_count_over before that barrier.
template <bool kCounted = false, int N>
__device__ __forceinline__ void p2g_scatter(P2GSharedT<N> &sh, int t, int nchunk, const KParams &kp,
                                            const GridState &gs, const BrickList &bl) {
    const int lane = t & 63;
    const int wave = t >> 6;
    // list of points outside the tile
    const bool over = sh.base[t] == -2;
    const unsigned long long om = __ballot(over);
    int nover_counted = 0;
    if constexpr (kCounted) {
        // every thread totals the published counts; the list and its barrier
        // are needed only when some point is outside the tile (the same
        // total everywhere, so the branch is uniform)
        int offo = 0;
        for (int w = 0; w < N / 64; w++) {
            const int c = sh.wave_count[w];
            nover_counted += c;
            if (w < wave) offo += c;
        }
        if (nover_counted > 0) {
            if (over) sh.over[offo + __popcll(om & ((1ull << lane) - 1ull))] = t;
            if (t == 0) count_stat(kp, kStatP2GOutOfTile, (unsigned long long)nover_counted);
            __syncthreads();
        }
    } else {
        if (lane == 0) sh.wave_count[wave] = __popcll(om);
        __syncthreads();
        int offo = 0;
        for (int w = 0; w < wave; w++) offo += sh.wave_count[w];
        if (over) sh.over[offo + __popcll(om & ((1ull << lane) - 1ull))] = t;
        if (t == 0) {
            int to = 0;
            for (int w = 0; w < N / 64; w++) to += sh.wave_count[w];
            sh.nover = to;
            count_stat(kp, kStatP2GOutOfTile, (unsigned long long)to);
        }
        __syncthreads();
    }

    // lane roles: 7 groups of 9 lanes per wave; lane (ox,oy) owns a node column
    const int gq = lane % 9;
    const bool group_lane = lane < 63;
    const int group = wave * 7 + lane / 9;
    const int ox = gq / 3, oy = gq % 3;
    const double dox = ox, doy = oy;
    {
        constexpr int kGroups = groups_of<N>();
        const int per = (nchunk + kGroups - 1) / kGroups;
        const int j_lo = group_lane ? min(group * per, nchunk) : nchunk;
        const int j_hi = min(j_lo + per, nchunk);
        Column col;
        column_zero(col);
        int cur_pk = -1;
        for (int j = j_lo; j < j_hi; j++) {
            // the base node and the record are requested together: one LDS
            // round trip per point (the record of a slot without a point is
            // read and not used)
            const int pk = sh.base[j];
            ColRec rj;
            column_load(sh, j, ox, oy, rj);
            __builtin_amdgcn_sched_barrier(0);
            if (pk < 0) continue;
            if (pk != cur_pk) {
                if (cur_pk >= 0) {
                    if (pk == cur_pk + (1 << 6)) {
                        // base moved one node up in z (the common case in
                        // cell order): the column's lowest node is final
                        column_node_to_tile(sh, cur_pk, ox, oy, col);
                        column_shift(col);
                    } else {
                        column_to_tile(sh, cur_pk, ox, oy, col);
                        column_zero(col);
                    }
                }
                cur_pk = pk;
            }
            column_accum(rj, dox, doy, col);
        }
        if (cur_pk >= 0) column_to_tile(sh, cur_pk, ox, oy, col);
    }
    const int nover = kCounted ? nover_counted : sh.nover;
    for (int r = group_lane ? group : nover; r < nover; r += groups_of<N>()) {
        const int j = sh.over[r];
        const int ab = sh.absb[j];
        Column col;
        column_zero(col);
        column_add(sh, j, ox, oy, dox, doy, col);
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

// Tile node of flush slot k, in the address order of the grid's bricked
// layout: the tile spans 2 + 4 + 2 nodes of three bricks per axis; slots run
// over the 27 brick pieces (x, then y, then z) and within a piece in the
// brick's node order. A wavefront touches as many cache lines as in plain
// x-major slot order, but the flush measured about 2% faster this way.
__device__ __forceinline__ int piece_start(int p) { return p == 0 ? 0 : (p == 1 ? 2 : 6); }
__device__ __forceinline__ int piece_size(int p) { return p == 1 ? 4 : 2; }
__device__ __forceinline__ void flush_order(int k, int &tx, int &ty, int &tz) {
    const int px = k < 128 ? 0 : (k < 384 ? 1 : 2);
    const int sx = piece_size(px);
    int r = k - piece_start(px) * 64;
    const int py = r < sx * 16 ? 0 : (r < sx * 48 ? 1 : 2);
    const int sy = piece_size(py);
    r -= sx * piece_start(py) * 8;
    const int pz = r < sx * sy * 2 ? 0 : (r < sx * sy * 6 ? 1 : 2);
    const int sz = piece_size(pz);
    r -= sx * sy * piece_start(pz);
    const int syz = sy * sz;
    const int ix = r / syz, rem = r - ix * syz;
    const int iy = rem / sz, iz = rem - iy * sz;
    tx = piece_start(px) + ix;
    ty = piece_start(py) + iy;
    tz = piece_start(pz) + iz;
}

// Adds the tile to the grid and collects the touched bricks in
// sh.brickmask (call after a barrier; the mask is complete, and the tile
// free, after the next barrier).
template <int N>
__device__ __forceinline__ void p2g_flush_nodes(P2GSharedT<N> &sh, int t, const KParams &kp, const int org[3],
                                                const GridState &gs) {
    for (int k = t; k < kTileNodes; k += N) {
        int tx, ty, tz;
        flush_order(k, tx, ty, tz);
        const int tk = tile_index(tx, ty, tz);
        const double m = sh.tile[0][tk];
        if (m != 0.0) {
            const long long idx = node_index(kp, org[0] + tx, org[1] + ty, org[2] + tz);
            unsafeAtomicAdd(&gs.grid[idx], m);
            unsafeAtomicAdd(&gs.grid[kp.grid_pitch + idx], sh.tile[1][tk]);
            unsafeAtomicAdd(&gs.grid[2 * kp.grid_pitch + idx], sh.tile[2][tk]);
            unsafeAtomicAdd(&gs.grid[3 * kp.grid_pitch + idx], sh.tile[3][tk]);
            // tile node t lies in brick (b - 1) + ((t + 2) >> 2) along each axis
            const int dbx = (tx + 2) >> 2, dby = (ty + 2) >> 2, dbz = (tz + 2) >> 2;
            atomicOr(&sh.brickmask, 1u << ((dbx * 3 + dby) * 3 + dbz));
        }
    }
}

// End of a segment whose successor is the next brick up in z (its tile
// starts 4 nodes higher): for each node n = (tx, ty, q) = (n >> 5,
// (n >> 2) & 7, n & 3) of the tile's lower half it handles (n = t, t + N,
// ...), thread t adds the node to the grid and marks its brick, then moves
// its row's element q + 4 down to q and clears q + 4 (each thread touches
// only elements q and q + 4 of its rows, so no barrier is needed in
// between). The
// successor starts with the upper half carried over, and each node of a run
// of z-neighbour bricks goes to HBM once instead of once per brick reaching
// it. Call after a barrier; the mask is complete after the next barrier.
static_assert(kTile == 8, "the lower half of the tile is 256 nodes, 4 per row");
template <int N>
__device__ __forceinline__ void p2g_flush_lower_slide(P2GSharedT<N> &sh, int t, const KParams &kp, const int org[3],
                                                      const GridState &gs) {
#pragma unroll
    for (int n = t; n < kTileNodes / 2; n += N) {
        const int tx = n >> 5, ty = (n >> 2) & 7, q = n & 3;
        const int lo = tile_index(tx, ty, q);
        const double m = sh.tile[0][lo];
        if (m != 0.0) {
            const long long idx = node_index(kp, org[0] + tx, org[1] + ty, org[2] + q);
            unsafeAtomicAdd(&gs.grid[idx], m);
            unsafeAtomicAdd(&gs.grid[kp.grid_pitch + idx], sh.tile[1][lo]);
            unsafeAtomicAdd(&gs.grid[2 * kp.grid_pitch + idx], sh.tile[2][lo]);
            unsafeAtomicAdd(&gs.grid[3 * kp.grid_pitch + idx], sh.tile[3][lo]);
            const int dbx = (tx + 2) >> 2, dby = (ty + 2) >> 2, dbz = (q + 2) >> 2;
            atomicOr(&sh.brickmask, 1u << ((dbx * 3 + dby) * 3 + dbz));
        }
#pragma unroll
        for (int l = 0; l < 4; l++) {
            sh.tile[l][lo] = sh.tile[l][lo + 4];
            sh.tile[l][lo + 4] = 0.0;
        }
    }
}

// Lists the bricks of sh.brickmask (after p2g_flush_nodes and a barrier).
// Only lanes 0-26, all in wavefront 0, read the mask, so its clear by lane 0
// (p2g_zero_tile) later in the same wavefront's program needs no barrier in
// between. The fused kernel relies on this (no barrier after the marking):
// the asserts below fix the lane layout it needs, and building with
// ICEMPM_SEGMENT_END_BARRIER=1 puts the barrier back to compare against.
constexpr int kBrickMaskLanes = 27;
static_assert(kBrickMaskLanes <= 64, "the brick-mask readers must all lie in wavefront 0");
static_assert(kBrickMaskLanes <= 32, "the brick mask is one 32-bit word");
template <int N>
__device__ __forceinline__ void p2g_mark_tile_bricks(const P2GSharedT<N> &sh, int t, const KParams &kp, const int bc[3],
                                                     const BrickList &bl) {
    if (t < kBrickMaskLanes && ((sh.brickmask >> t) & 1u)) {
        const int dbx = t / 9, dby = (t / 3) % 3, dbz = t % 3;
        const long long b = ((long long)(bc[0] - 1 + dbx) * kp.nby + (bc[1] - 1 + dby)) * kp.nbz + (bc[2] - 1 + dbz);
        mark_brick(bl, b);
    }
}

// Adds the tile to the grid and marks the touched bricks (call after a
// barrier; ends with a barrier).
template <int N>
__device__ __forceinline__ void p2g_flush(P2GSharedT<N> &sh, int t, const KParams &kp, const int org[3],
                                          const int bc[3], const GridState &gs, const BrickList &bl) {
    p2g_flush_nodes(sh, t, kp, org, gs);
    __syncthreads();
    p2g_mark_tile_bricks(sh, t, kp, bc, bl);
    __syncthreads();
}

// Work items of a workgroup: the N-point sub-chunks of its contiguous range
// of segments, in order.
template <int N>
struct ChunkCursorT {
    long long s, s_hi;
    int begin, end, c0, brick, coord, slot;
    __device__ __forceinline__ bool valid() const { return s < s_hi; }
    __device__ __forceinline__ bool first() const { return c0 == begin; }
    __device__ __forceinline__ bool last() const { return c0 + N >= end; }
};
using ChunkCursor = ChunkCursorT<kP2GChunk>;

template <int N>
__device__ __forceinline__ void cursor_load_segment(ChunkCursorT<N> &c, const SegInfo &seg) {
    if (c.s < c.s_hi) {
        c.begin = seg.seg_start[c.s];
        c.end = seg.seg_start[c.s + 1];
        c.brick = seg.seg_brick[c.s];
        c.coord = seg.seg_coord[c.s];
        c.slot = seg.seg_slot[c.s];
        c.c0 = c.begin;
    }
}

// Cost of a segment in the work split, in points: its points plus a fixed
// kSegCost for the tile fetch, clear and flush (512 measured best among
// 128-1024 on both the 16M-point block and the 64-scene ensemble; an equal
// number of segments per workgroup was 2-5% slower, an equal number of
// points 25% slower on the block). The split is computed once per change
// of the segments or of the grid (k_work_split), not by every workgroup: a
// search per workgroup at kernel start adds ~10 us of dependent loads.
constexpr long long kSegCost = 512;

// First segment s in [0, nseg] whose cumulative cost seg_start[s] + kSegCost s
// reaches target (seg_start is non-decreasing).
__device__ __forceinline__ long long seg_at_cost(const int32_t *seg_start, int nseg, long long target) {
    long long lo = 0, hi = nseg;
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if ((long long)seg_start[mid] + kSegCost * mid < target) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The contiguous range of segments of workgroup lb (equal shares of the total
// cost, from the split launch_work_split computed for this grid).
template <int N = kP2GChunk>
__device__ __forceinline__ ChunkCursorT<N> cursor_first(const SegInfo &seg, int nseg, int lb, int nblocks) {
    (void)nseg;
    (void)nblocks;
    ChunkCursorT<N> c;
    c.s = seg.wg_start[lb];
    c.s_hi = seg.wg_start[lb + 1];
    c.begin = c.end = c.c0 = c.brick = c.coord = c.slot = 0;
    cursor_load_segment(c, seg);
    return c;
}

template <int N>
__device__ __forceinline__ ChunkCursorT<N> cursor_next(ChunkCursorT<N> c, const SegInfo &seg) {
    if (c.c0 + N < c.end) {
        c.c0 += N;
    } else {
        c.s++;
        cursor_load_segment(c, seg);
    }
    return c;
}

} // namespace icempm
