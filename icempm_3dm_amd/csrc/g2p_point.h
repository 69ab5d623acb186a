// g2p_point.h — grid-to-particle update of one point (partition_kernel_g2p,
// mpm_kernels.cuh:237-351) and the LDS velocity tile it gathers from,
// shared by the G2P kernel and the fused G2P+P2G kernel.
#pragma once

#include "p2g_tile.h"

namespace icempm {

// State of one point as the step carries it (Point3D, point_3d.h:14-27).
struct PointState {
    long long util;
    double pos[3], vel[3], Fe[9], Bp[9];
    double Jp;
    // J = det Fe and J^{-2/3} of the stored Fe, when g2p_finish formed them
    // (a solid point whose trial F the return mapping left as it was)
    double Je, rcJ;
    bool have_j;
};

// Loads what G2P needs before the gather: utility and position. Fe and
// Jp_inv are read by g2p_finish after the gather, so they are not live
// across it.
__device__ __forceinline__ void load_g2p_inputs(PtsView pts, int i, PointState &s) {
    s.util = double_to_bits(pts.at(ICEMPM_IDX_UTILITY, i));
#pragma unroll
    for (int a = 0; a < 3; a++) s.pos[a] = pts.at(ICEMPM_POSX + a, i);
}

// All 14 doubles G2P reads of a point, for a caller that loads them a chunk
// ahead (g2p_finish<..., true> then skips its own Fe and Jp_inv loads).
__device__ __forceinline__ void load_g2p_state(PtsView pts, int i, PointState &s) {
    load_g2p_inputs(pts, i, s);
#pragma unroll
    for (int k = 0; k < 9; k++) s.Fe[k] = pts.at(ICEMPM_FE00 + k, i);
    s.Jp = pts.at(ICEMPM_IDX_JP_INV, i);
}

// Node velocities of a tile: values thread t of an N-thread workgroup
// fetches for nodes t, t + N, ... (fetched ahead, stored to LDS later).
template <int N = kP2GThreads>
struct VelTileRegs {
    static_assert(kTileNodes % N == 0, "whole tile nodes per thread");
    static constexpr int kPer = kTileNodes / N;
    double v[kPer][3];
};

template <int N>
__device__ __forceinline__ void vel_tile_fetch(VelTileRegs<N> &r, int t, const KParams &kp, const int org[3],
                                               const double *grid) {
#pragma unroll
    for (int h = 0; h < VelTileRegs<N>::kPer; h++) {
        const int k = t + h * N;
        const int x = org[0] + (k >> 6), y = org[1] + ((k >> 3) & 7), z = org[2] + (k & 7);
        r.v[h][0] = r.v[h][1] = r.v[h][2] = 0.0;
        // nodes outside the allocated bricks are never weighted (their points raised an error)
        if ((unsigned)x < (unsigned)(kp.nbx * 4) && (unsigned)y < (unsigned)(kp.nby * 4) &&
            (unsigned)z < (unsigned)(kp.nbz * 4)) {
            const long long idx = node_index(kp, x, y, z);
#pragma unroll
            for (int c = 0; c < 3; c++) r.v[h][c] = grid[(1 + c) * kp.grid_pitch + idx];
        }
    }
}

template <int N>
__device__ __forceinline__ void vel_tile_store(double (*sv)[kTileAlloc], const VelTileRegs<N> &r, int t) {
#pragma unroll
    for (int h = 0; h < VelTileRegs<N>::kPer; h++) {
        const int k = t + h * N;
        const int tk = tile_index(k >> 6, (k >> 3) & 7, k & 7);
#pragma unroll
        for (int c = 0; c < 3; c++) sv[c][tk] = r.v[h][c];
    }
}

// Start of a segment that follows the next brick down in z: its tile is the
// previous one moved up 4 nodes. For each element n = (tx, ty, q) = (n >> 5,
// (n >> 2) & 7, n & 3) it handles (n = t, t + N, ...), thread t moves the
// previous tile's element q + 4 of row (tx, ty) down to q and reads element
// q + 4 (node (tx, ty, q + 4) of the new tile) from the grid; the reads are
// all issued before the LDS is touched. Call after the previous segment's
// gathers are finished.
template <int N>
__device__ __forceinline__ void vel_tile_slide(double (*sv)[kTileAlloc], int t, const KParams &kp,
                                               const int org[3], const double *grid) {
    constexpr int kPer = kTileNodes / 2 / N;
    static_assert(kPer * N == kTileNodes / 2, "whole elements per thread");
    double v[kPer][3];
#pragma unroll
    for (int h = 0; h < kPer; h++) {
        const int n = t + h * N;
        const int x = org[0] + (n >> 5), y = org[1] + ((n >> 2) & 7), z = org[2] + 4 + (n & 3);
        v[h][0] = v[h][1] = v[h][2] = 0.0;
        // nodes outside the allocated bricks are never weighted (their points raised an error)
        if ((unsigned)x < (unsigned)(kp.nbx * 4) && (unsigned)y < (unsigned)(kp.nby * 4) &&
            (unsigned)z < (unsigned)(kp.nbz * 4)) {
            const long long idx = node_index(kp, x, y, z);
#pragma unroll
            for (int c = 0; c < 3; c++) v[h][c] = grid[(1 + c) * kp.grid_pitch + idx];
        }
    }
#pragma unroll
    for (int h = 0; h < kPer; h++) {
        const int n = t + h * N;
        const int lo = tile_index(n >> 5, (n >> 2) & 7, n & 3);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            sv[c][lo] = sv[c][lo + 4];
            sv[c][lo + 4] = v[h][c];
        }
    }
}

// One wavefront loads the whole tile: lane l takes the x-planes' nodes
// (x, l >> 3, l & 7), all 24 loads in flight before the LDS stores.
__device__ __forceinline__ void vel_tile_load_wave(double (*sv)[kTileAlloc], int lane, const KParams &kp,
                                                   const int org[3], const double *grid) {
    static_assert(kTile * kTile == 64, "one tile x-plane per wavefront row");
    double v[kTile][3];
    const int y = org[1] + (lane >> 3), z = org[2] + (lane & 7);
    const bool yz_in = (unsigned)y < (unsigned)(kp.nby * 4) && (unsigned)z < (unsigned)(kp.nbz * 4);
#pragma unroll
    for (int tx = 0; tx < kTile; tx++) {
        const int x = org[0] + tx;
        v[tx][0] = v[tx][1] = v[tx][2] = 0.0;
        // nodes outside the allocated bricks are never weighted (their points raised an error)
        if (yz_in && (unsigned)x < (unsigned)(kp.nbx * 4)) {
            const long long idx = node_index(kp, x, y, z);
#pragma unroll
            for (int c = 0; c < 3; c++) v[tx][c] = grid[(1 + c) * kp.grid_pitch + idx];
        }
    }
    const int t0 = tile_index(0, lane >> 3, lane & 7);
#pragma unroll
    for (int tx = 0; tx < kTile; tx++)
#pragma unroll
        for (int c = 0; c < 3; c++) sv[c][t0 + tx * kTileSX] = v[tx][c];
}

// Velocity of the 27 stencil nodes -> point velocity and APIC matrix C
// (mpm_kernels.cuh:288-305): vel = sum W v, C = (4/h) sum W v (o - f)^T with
// W = wx[a] wy[b] wz[c], summed in separable order (over z, then y, then x),
// 6 FMAs per node instead of 18. `V(a, b, c, nv)` fetches node velocity.
template <typename Fetch>
__device__ __forceinline__ void gather(const double w[3][3], const double d[3][3], double c4h, Fetch V,
                                       double vel[3], double Bp[9]) {
    double wdz[3];
#pragma unroll
    for (int c = 0; c < 3; c++) wdz[c] = w[2][c] * d[2][c];
#pragma unroll
    for (int r = 0; r < 3; r++) vel[r] = 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) Bp[k] = 0.0;
#pragma unroll 1
    for (int a = 0; a < 3; a++) {
        double Ta[3] = {0.0, 0.0, 0.0}, Ty[3] = {0.0, 0.0, 0.0}, Tz[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int b = 0; b < 3; b++) {
            double S0[3] = {0.0, 0.0, 0.0}, S1[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double nv[3];
                V(a, b, c, nv);
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    S0[r] += w[2][c] * nv[r];
                    S1[r] += wdz[c] * nv[r];
                }
            }
            const double wy = w[1][b], wyd = w[1][b] * d[1][b];
#pragma unroll
            for (int r = 0; r < 3; r++) {
                Ta[r] += wy * S0[r];
                Ty[r] += wyd * S0[r];
                Tz[r] += wy * S1[r];
            }
        }
        const double wx = w[0][a], wxd = w[0][a] * d[0][a];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            vel[r] += wx * Ta[r];
            Bp[3 * r] += wxd * Ta[r];
            Bp[3 * r + 1] += wx * Ty[r];
            Bp[3 * r + 2] += wx * Tz[r];
        }
    }
#pragma unroll
    for (int k = 0; k < 9; k++) Bp[k] *= c4h;
}

__device__ __forceinline__ void stencil_weights(const double fx[3], double w[3][3], double d[3][3]) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
        bspline3(fx[a], w[a]);
#pragma unroll
        for (int o = 0; o < 3; o++) d[a][o] = o - fx[a];
    }
}

__device__ __forceinline__ void gather_tile(const double (*sv)[kTileAlloc], int t0, const double w[3][3],
                                            const double d[3][3], double c4h, double vel[3], double Bp[9]) {
    gather(w, d, c4h,
           [&](int a, int b, int c, double nv[3]) {
               const int tk = t0 + a * kTileSX + b * kTileSY + c;
               nv[0] = sv[0][tk];
               nv[1] = sv[1][tk];
               nv[2] = sv[2][tk];
           },
           vel, Bp);
}

__device__ __forceinline__ void gather_global(const KParams &kp, const double *grid, const int base[3],
                                              const double w[3][3], const double d[3][3], double c4h, double vel[3],
                                              double Bp[9]) {
    long long X[3], Y[3], Z[3];
#pragma unroll
    for (int o = 0; o < 3; o++) {
        const int x = base[0] + o, y = base[1] + o, z = base[2] + o;
        X[o] = (long long)(x >> 2) * kp.nby * kp.nbz * kBrickNodes + ((x & 3) << 4);
        Y[o] = (long long)(y >> 2) * kp.nbz * kBrickNodes + ((y & 3) << 2);
        Z[o] = (long long)(z >> 2) * kBrickNodes + (z & 3);
    }
    const double *gvx = grid + kp.grid_pitch;
    const double *gvy = grid + 2 * kp.grid_pitch;
    const double *gvz = grid + 3 * kp.grid_pitch;
    gather(w, d, c4h,
           [&](int a, int b, int c, double nv[3]) {
               const long long idx = X[a] + Y[b] + Z[c];
               nv[0] = gvx[idx];
               nv[1] = gvy[idx];
               nv[2] = gvz[idx];
           },
           vel, Bp);
}

// Everything after the gather (mpm_kernels.cuh:307-351): reads Fe and
// Jp_inv, then advection, F update, p/q, crush test; stores the 26 (+2 with
// record_pq) lines and leaves the new state in s. Without store_vc the
// velocity and C lines are not stored: a caller whose next reader of the
// point is another G2P (which recomputes both from the grid) saves the
// 96 B. With kInlineReturnMap the return mapping of a crushed point runs
// here and false is returned; otherwise the trial F is stored and true is
// returned for a crushed solid point, whose return mapping must follow
// before its F is used. With kPreloaded, s.Fe and s.Jp already hold the
// stored values. kDamage: the return mapping with the damage model.
template <bool kInlineReturnMap, bool kPreloaded = false, bool kDamage = false>
__device__ __forceinline__ bool g2p_finish(const KParams &kp, PtsView pts, int i, PointState &s, int record_pq,
                                           bool store_vc = true) {
    const double dt = kp.dt;
    const bool liquid = s.util & ICEMPM_STATUS_LIQUID;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        s.pos[a] += s.vel[a] * dt;
        pts.at(ICEMPM_POSX + a, i) = s.pos[a];
        if (store_vc) pts.at(ICEMPM_VELX + a, i) = s.vel[a];
    }
    if (!kPreloaded) {
#pragma unroll
        for (int k = 0; k < 9; k++) s.Fe[k] = pts.at(ICEMPM_FE00 + k, i);
        s.Jp = pts.at(ICEMPM_IDX_JP_INV, i);
    }
    double A[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        if (store_vc) pts.at(ICEMPM_BP00 + k, i) = s.Bp[k];
        A[k] = dt * s.Bp[k] + ((k % 4 == 0) ? 1.0 : 0.0);
    }
    PState p;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            p.Fe[3 * r + c] = A[3 * r] * s.Fe[c] + A[3 * r + 1] * s.Fe[3 + c] + A[3 * r + 2] * s.Fe[6 + c];
    p.Jp_inv = s.Jp;
    p.util = s.util;
    p.p_tr = 0.0;
    p.q_tr = 0.0;
    bool crushed = false;
    if (liquid) {
        p.Jp_inv *= det3(A);
    } else {
        compute_pq(kp, p);
        if (!(p.util & ICEMPM_STATUS_CRUSHED)) check_failure_surface(kp, p);
        crushed = (p.util & ICEMPM_STATUS_CRUSHED) != 0;
        bool kept = !crushed;
        if (kInlineReturnMap && crushed) {
            kept = !return_map<kDamage>(kp, p);
            crushed = false;
        }
        s.have_j = kept;
        s.Je = p.Je_tr;
        s.rcJ = p.rc_Jsq;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
        s.Fe[k] = p.Fe[k];
        pts.at(ICEMPM_FE00 + k, i) = p.Fe[k];
    }
    s.Jp = p.Jp_inv;
    s.util = p.util;
    pts.at(ICEMPM_IDX_JP_INV, i) = p.Jp_inv;
    pts.at(ICEMPM_IDX_UTILITY, i) = bits_to_double(p.util);
    if (record_pq) {
        pts.at(ICEMPM_IDX_P, i) = p.p_tr;
        pts.at(ICEMPM_IDX_Q, i) = p.q_tr;
    }
    return crushed;
}

// g2p_finish with the return mapping inline and the state preloaded, the
// damage model chosen at run time (kernels off the hot path).
__device__ __forceinline__ void g2p_finish_any(const KParams &kp, PtsView pts, int i, PointState &s, int record_pq,
                                               bool store_vc = true) {
    if (kp.dmg_on) g2p_finish<true, true, true>(kp, pts, i, s, record_pq, store_vc);
    else g2p_finish<true, true, false>(kp, pts, i, s, record_pq, store_vc);
}


// Workgroup-local list of the chunk's crushed points (built after a barrier
// from per-lane flags, no global atomics).
struct CrushList {
    int idx[kP2GThreads];
    int wave_count[kP2GThreads / 64];
    int n;
};

// All threads call; ends with a barrier after which cl.n / cl.idx are valid.
__device__ __forceinline__ void crush_list_build(CrushList &cl, int t, bool crushed, int i) {
    const int lane = t & 63, wave = t >> 6;
    const unsigned long long m = __ballot(crushed);
    if (lane == 0) cl.wave_count[wave] = __popcll(m);
    __syncthreads();
    int off = 0;
    for (int w = 0; w < wave; w++) off += cl.wave_count[w];
    if (crushed) cl.idx[off + __popcll(m & ((1ull << lane) - 1ull))] = i;
    if (t == 0) {
        int tot = 0;
        for (int w = 0; w < kP2GThreads / 64; w++) tot += cl.wave_count[w];
        cl.n = tot;
    }
    __syncthreads();
}

// Appends `i` to a list when `flag` holds, one atomic per wave.
__device__ __forceinline__ void wave_append(bool flag, int i, int *list, int *count) {
    const unsigned long long m = __ballot(flag);
    if (m) {
        const int lane = __lane_id();
        const int leader = __ffsll((long long)m) - 1;
        int base_slot = 0;
        if (lane == leader) base_slot = atomicAdd(count, __popcll(m));
        base_slot = __shfl(base_slot, leader, 64);
        if (flag) list[base_slot + __popcll(m & ((1ull << lane) - 1ull))] = i;
    }
}

// Reference-form scatter of one point straight to HBM (mpm_kernels.cuh:27-99),
// for points outside every tile. Returns the error code (0: scattered).
__device__ __forceinline__ int scatter_point_global(const KParams &kp, const GridState &gs, const BrickList &bl,
                                                    const PointState &s) {
    if (s.util & ICEMPM_STATUS_DISABLED) return 0;
    const bool liquid = s.util & ICEMPM_STATUS_LIQUID;
    const double m = liquid ? kp.mass_liquid : kp.mass_solid;
    int base[3];
    double fx[3];
    const int code = locate(kp, s.pos, base, fx);
    if (code) return code;
    double sub[9];
    {
        double PFt[9];
        if (liquid) water_stress(kp, s.Jp, PFt);
        else kirchhoff_stress(kp, s.Fe, PFt);
#pragma unroll
        for (int k = 0; k < 9; k++) sub[k] = m * s.Bp[k] - kp.dt_vol_Dpinv * PFt[k];
    }
    double w[3][3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double wd[3];
        bspline3(fx[d], wd);
        w[0][d] = wd[0];
        w[1][d] = wd[1];
        w[2][d] = wd[2];
    }
    const double h = kp.h;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double Wip = w[a][0] * w[b][1] * w[c][2];
                const double dpos[3] = {(a - fx[0]) * h, (b - fx[1]) * h, (c - fx[2]) * h};
                const long long idx = node_index(kp, base[0] + a, base[1] + b, base[2] + c);
                unsafeAtomicAdd(&gs.grid[idx], Wip * m);
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    const double sd = sub[3 * r] * dpos[0] + sub[3 * r + 1] * dpos[1] + sub[3 * r + 2] * dpos[2];
                    unsafeAtomicAdd(&gs.grid[(1 + r) * kp.grid_pitch + idx], Wip * (s.vel[r] * m + sd));
                }
            }
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++)
                mark_brick(bl, brick_of(kp, base[0] + 2 * a, base[1] + 2 * b, base[2] + 2 * c));
    return 0;
}

} // namespace icempm
