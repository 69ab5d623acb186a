// kernels_g2p2g.hip — G2P of substep n fused with P2G of substep n+1.
//
// P2G of the next substep depends only on the point state G2P has just
// produced, so one pass per point does both. Per chunk of N points (256, or
// 128 where the autotuner finds it faster), one lane per point:
//   G  read the 14 doubles G2P needs, gather node velocities from grid A
//      (LDS tile), update the point (return mapping included) and store it;
//   P  build the P2G record of the new state from registers, then the
//      workgroup scatters the chunk into grid B (LDS tile, flushed with
//      atomics at the end of the segment).
// Per point that is 112 B read + 208 B written to HBM instead of 200 + 112
// + 208 B for the two separate passes. When the next reader of the points is
// again a G2P (store_vc = 0: another fused pass or the closing G2P of the
// substep loop), velocity and C are not stored either (they are recomputed
// from the grid there): 112 B read + 112 B written. Points outside their
// segment's tile at the start of the substep go to a spill list
// (k_g2p2g_spill).
#include "g2p_point.h"

// 1: a barrier after the brick marking at every segment end (the kernel is
// correct without it; the switch exists to compare the two builds).
#ifndef ICEMPM_SEGMENT_END_BARRIER
#define ICEMPM_SEGMENT_END_BARRIER 0
#endif

namespace icempm {

namespace {

__global__ void k_zero_count(int *c) { *c = 0; }

// N points per chunk and threads per workgroup. With N = 256, LDS (two
// tiles and the records, about 78 KB) limits the kernel to two workgroups per
// CU; with N = 128 (54 KB) three fit. The bound keeps the compiler within the
// 256 registers of two waves per SIMD instead of spilling past them into
// AGPRs.
// kOOP: the point state is read from src and the new state written to pts
// (the calibration adjoint keeps every substep's state); otherwise both are pts.
// kDamage: the return mapping with the damage model (kp.dmg_on).
template <bool kEns, bool kOOP, bool kDamage, int N>
__global__ __launch_bounds__(N) __attribute__((amdgpu_waves_per_eu(2))) void k_g2p2g(KParams kp, PtsView pts, SegInfo seg,
                                                       const double *__restrict__ grid_a, GridState gs_b, int cur_b,
                                                       int record_pq, const StepClock *clk, int store_vc, int *err,
                                                       int *spill, int *spill_count, PtsView src) {
    if (clk) record_pq = clk->record;
    const PtsView in = kOOP ? src : pts;
    __shared__ P2GSharedT<N> sh;
    __shared__ double sv[3][kTileAlloc];
    const int t = threadIdx.x;
    const BrickList bl = brick_list(gs_b, cur_b);
    const int nseg = *seg.nseg;
    const double c4h = 4. * kp.h_inv;
    ChunkCursorT<N> c = cursor_first<N>(seg, nseg, xcd_remap(blockIdx.x, gridDim.x), gridDim.x);
    // what G2P reads of the next chunk's points is loaded during the scatter
    PointState pf;
    pf.util = ICEMPM_STATUS_DISABLED;
    if (c.valid() && c.c0 + t < c.end) load_g2p_state(in, c.c0 + t, pf);
    // the segment before was the next brick down in z and left its tiles
    // moved up by 4 nodes (p2g_flush_lower_slide)
    bool slid = false;
    while (c.valid()) {
        const KParams ks = scene_view_slot<kEns>(kp, c.slot);
        int org[3], bc[3];
        tile_origin_packed(c.coord, org, bc);
        if (c.first()) {
            // the previous segment's gathers and flush are finished (barrier at its end)
            if (slid) {
                vel_tile_slide<N>(sv, t, ks, org, grid_a);
                if (t == 0) sh.brickmask = 0u;
            } else {
                VelTileRegs<N> vt;
                vel_tile_fetch(vt, t, ks, org, grid_a);
                vel_tile_store(sv, vt, t);
                p2g_zero_tile(sh, t);
            }
            __syncthreads();
        }
        const int i = c.c0 + t;
        PointState s;
        s.util = pf.util;
#pragma unroll
        for (int a = 0; a < 3; a++) s.pos[a] = pf.pos[a];
#pragma unroll
        for (int k = 0; k < 9; k++) s.Fe[k] = pf.Fe[k];
        s.Jp = pf.Jp;
        const ChunkCursorT<N> nx = cursor_next(c, seg);
        bool updated = false, spilled = false;
        s.have_j = false;
        if (kOOP && (s.util & ICEMPM_STATUS_DISABLED) && i < c.end)
            pts.at(ICEMPM_IDX_UTILITY, i) = bits_to_double(s.util);
        if (!(s.util & ICEMPM_STATUS_DISABLED)) {
            int base[3];
            double fx[3];
            const int code = locate(ks, s.pos, base, fx);
            if (code) {
                raise_error(err, code);
            } else {
                const int rx = base[0] - org[0], ry = base[1] - org[1], rz = base[2] - org[2];
                if ((unsigned)rx <= kTile - 3 && (unsigned)ry <= kTile - 3 && (unsigned)rz <= kTile - 3) {
                    double w[3][3], d[3][3];
                    stencil_weights(fx, w, d);
                    gather_tile(sv, tile_index(rx, ry, rz), w, d, c4h, s.vel, s.Bp);
                    g2p_finish<true, true, kDamage>(ks, pts, i, s, record_pq, store_vc != 0);
                    updated = true;
                } else {
                    spilled = true;
                }
            }
        }
        wave_append(spilled, i, spill, spill_count);
        // P2G record of the new state, straight from registers. For a point
        // whose F the return mapping left as it was, J and J^{-2/3} are those
        // of G2P's crush test (s.have_j); dev(F F^T) is formed again, since
        // keeping its six doubles across the return mapping as well measured
        // 0.2 ms slower per launch on 16M points (register spills).
        double Je = 1.0, rc = 1.0;
        if (updated && s.have_j) {
            Je = s.Je;
            rc = s.rcJ;
        } else if (updated) {
            Je = det3(s.Fe);
            rc = rcbrt(Je * Je);
        }
        p2g_record(sh, t, ks, updated, s.util, s.pos, s.vel, s.Fe, Je, rc, s.Bp, s.Jp, org, err);
        // the record barrier also publishes the scatter's counts of points
        // outside the tile: one barrier fewer per chunk, and none for the
        // out-of-tile list when the counts are all zero
        p2g_count_over(sh, t);
        __syncthreads();
        pf.util = ICEMPM_STATUS_DISABLED;
        if (nx.valid() && nx.c0 + t < nx.end) load_g2p_state(in, nx.c0 + t, pf);
        p2g_scatter<true>(sh, t, min(N, c.end - c.c0), ks, gs_b, bl);
        if (c.last()) {
            // a successor one brick up in z takes the tile over, moved by 4 nodes
            slid = nx.valid() && nx.brick == c.brick + 1 && bc[2] + 1 < ks.nbz;
            if (slid) p2g_flush_lower_slide(sh, t, ks, org, gs_b);
            else p2g_flush_nodes(sh, t, ks, org, gs_b);
            __syncthreads();
            p2g_mark_tile_bricks(sh, t, ks, bc, bl);
            // No barrier here: the next segment's set-up above writes the
            // velocity tile (last read by the gathers, before the record
            // barrier) and, unless the slide already did, clears the P2G tile
            // (last read by the flush, before the barrier above), so only
            // wavefront 0's brick marking is still running when the other
            // wavefronts start it, and the set-up's barrier joins them.
#if ICEMPM_SEGMENT_END_BARRIER
            __syncthreads();
#endif
        }
        c = nx;
    }
}

// Spilled points: G2P gathering from HBM, then the reference-form scatter
// of the new state straight into grid B.
__global__ __launch_bounds__(kP2GThreads) void k_g2p2g_spill(KParams kp, PtsView pts,
                                                             const double *__restrict__ grid_a, GridState gs_b,
                                                             int cur_b, int record_pq, const StepClock *clk,
                                                             int store_vc, int *err, const int *spill,
                                                             const int *spill_count, PtsView in) {
    if (clk) record_pq = clk->record;
    const int count = *spill_count;
    if (blockIdx.x == 0 && threadIdx.x == 0) count_stat(kp, kStatG2PSpilled, (unsigned long long)count);
    const double c4h = 4. * kp.h_inv;
    const BrickList bl = brick_list(gs_b, cur_b);
    for (int k = blockIdx.x * kP2GThreads + threadIdx.x; k < count; k += gridDim.x * kP2GThreads) {
        const int i = spill[k];
        PointState s;
        load_g2p_state(in, i, s);
        const KParams ks = kp.ens_n > 0 ? scene_of_point(kp, s.pos) : kp;
        int base[3];
        double fx[3];
        (void)locate(ks, s.pos, base, fx); // listed points passed the bounds check
        double w[3][3], d[3][3];
        stencil_weights(fx, w, d);
        gather_global(ks, grid_a, base, w, d, c4h, s.vel, s.Bp);
        g2p_finish_any(ks, pts, i, s, record_pq, store_vc != 0);
        const int code = scatter_point_global(ks, gs_b, bl, s);
        if (code) raise_error(err, code);
    }
}

} // namespace

hipError_t launch_g2p2g(const KParams &kp, PtsView pts, int n, SegInfo seg, const double *grid_a, GridState gs_b,
                        int cur_b, int record_pq, int store_vc, int *err, int *lists, int n_blocks,
                        const StepClock *clk, bool count_zeroed, hipStream_t stream, const PtsView *src,
                        int chunk) {
    // 128-point chunks are built for the forward step only (the adjoint's
    // out-of-place passes keep 256)
    if ((chunk != 128 && chunk != 256) || (src && chunk != 256)) return hipErrorInvalidValue;
    if (n <= 0) return hipSuccess;
    int *spill_count = lists, *spill = lists + 1;
    if (!count_zeroed) k_zero_count<<<1, 1, 0, stream>>>(spill_count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    n_blocks = segment_grid(n_blocks);
    const PtsView in = src ? *src : pts;
#define ICEMPM_G2P2G_N(E, O, D, NT)                                                                              \
    k_g2p2g<E, O, D, NT><<<n_blocks, NT, 0, stream>>>(kp, pts, seg, grid_a, gs_b, cur_b, record_pq, clk, store_vc,  \
                                                      err, spill, spill_count, in)
#define ICEMPM_G2P2G(E, O, D)                                                                                    \
    do {                                                                                                         \
        if (chunk == 128 && !(O)) ICEMPM_G2P2G_N(E, false, D, 128);                                              \
        else ICEMPM_G2P2G_N(E, O, D, 256);                                                                       \
    } while (0)
    // the adjoint's out-of-place passes never run with the damage model
    const bool ens = kp.ens_n > 0, dmg = kp.dmg_on != 0;
    if (src && dmg) return hipErrorInvalidValue;
    if (src) {
        if (ens) ICEMPM_G2P2G(true, true, false);
        else ICEMPM_G2P2G(false, true, false);
    } else if (dmg) {
        if (ens) ICEMPM_G2P2G(true, false, true);
        else ICEMPM_G2P2G(false, false, true);
    } else {
        if (ens) ICEMPM_G2P2G(true, false, false);
        else ICEMPM_G2P2G(false, false, false);
    }
#undef ICEMPM_G2P2G
#undef ICEMPM_G2P2G_N
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_g2p2g_spill<<<256, kP2GThreads, 0, stream>>>(kp, pts, grid_a, gs_b, cur_b, record_pq, clk, store_vc, err,
                                                    spill, spill_count, in);
    return hipGetLastError();
}

} // namespace icempm
