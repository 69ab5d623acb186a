// kernels_g2p.hip — grid-to-particle transfer and constitutive update
// (partition_kernel_g2p, mpm_kernels.cuh:237-351) with the return mapping of
// crushed points (mpm_kernels.cuh:486-544), one lane per point, one
// wavefront per brick segment with the segment's node velocities staged in
// LDS. Point transfer between partitions (:353-398) does not apply.
#include "g2p_point.h"

namespace icempm {

namespace {

__global__ void k_zero_count(int *c) { *c = 0; }

// Each wavefront owns a contiguous range of brick segments and its own LDS
// copy of a segment's 8x8x8 node velocities; it walks the segment in 64-point
// chunks, one lane per point, with no workgroup barrier. A point that
// drifted outside its segment's tile since the sort goes to the spill list.
constexpr int kG2PWaves = kP2GThreads / 64;

// kOOP: the point state is read from `in` and the new state written to pts;
// otherwise in is pts. kDamage: the return mapping with the damage model.
template <bool kEns, bool kOOP, bool kDamage>
__global__ __launch_bounds__(kP2GThreads) void k_g2p(KParams kp, PtsView pts, SegInfo seg,
                                                     const double *__restrict__ grid, int record_pq,
                                                     const StepClock *clk, int *err, int *spill, int *spill_count,
                                                     PtsView in) {
    if (clk) record_pq = clk->record;
    __shared__ double sv_all[kG2PWaves][3][kTileAlloc];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double(*sv)[kTileAlloc] = sv_all[wave];
    const long long nseg = *seg.nseg;
    const long long nw = (long long)gridDim.x * kG2PWaves;
    const long long w = (long long)xcd_remap(blockIdx.x, gridDim.x) * kG2PWaves + wave;
    const double c4h = 4. * kp.h_inv;
    for (long long sg = nseg * w / nw, sg_hi = nseg * (w + 1) / nw; sg < sg_hi; sg++) {
        const int begin = seg.seg_start[sg], end = seg.seg_start[sg + 1];
        const KParams ks = scene_view_slot<kEns>(kp, seg.seg_slot[sg]);
        int org[3], bc[3];
        tile_origin_packed(seg.seg_coord[sg], org, bc);
        // the previous segment's gathers precede these writes
        wave_lds_sync();
        vel_tile_load_wave(sv, lane, ks, org, grid);
        wave_lds_sync();
        for (int c0 = begin; c0 < end; c0 += 64) {
            const int i = c0 + lane;
            PointState s;
            s.util = ICEMPM_STATUS_DISABLED;
            if (i < end) load_g2p_state(in, i, s);
            if (kOOP && (s.util & ICEMPM_STATUS_DISABLED) && i < end)
                pts.at(ICEMPM_IDX_UTILITY, i) = bits_to_double(s.util);
            bool spilled = false;
            if (!(s.util & ICEMPM_STATUS_DISABLED)) {
                int base[3];
                double fx[3];
                const int code = locate(ks, s.pos, base, fx);
                if (code) {
                    raise_error(err, code);
                } else {
                    const int rx = base[0] - org[0], ry = base[1] - org[1], rz = base[2] - org[2];
                    if ((unsigned)rx <= kTile - 3 && (unsigned)ry <= kTile - 3 && (unsigned)rz <= kTile - 3) {
                        double w3[3][3], d[3][3];
                        stencil_weights(fx, w3, d);
                        gather_tile(sv, tile_index(rx, ry, rz), w3, d, c4h, s.vel, s.Bp);
                        g2p_finish<true, true, kDamage>(ks, pts, i, s, record_pq);
                    } else {
                        spilled = true;
                    }
                }
            }
            wave_append(spilled, i, spill, spill_count);
        }
    }
}

// G2P of the spilled points, gathering node velocities from HBM.
__global__ __launch_bounds__(kP2GThreads) void k_g2p_spill(KParams kp, PtsView pts, const double *__restrict__ grid,
                                                           int record_pq, const StepClock *clk, const int *spill,
                                                           const int *spill_count, PtsView in) {
    if (clk) record_pq = clk->record;
    const int count = *spill_count;
    if (blockIdx.x == 0 && threadIdx.x == 0) count_stat(kp, kStatG2PSpilled, (unsigned long long)count);
    const double c4h = 4. * kp.h_inv;
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
        gather_global(ks, grid, base, w, d, c4h, s.vel, s.Bp);
        g2p_finish_any(ks, pts, i, s, record_pq);
    }
}

} // namespace

hipError_t launch_g2p(const KParams &kp, PtsView pts, int n, SegInfo seg, const double *grid, int record_pq, int *err,
                      int *lists, int n_blocks, const StepClock *clk, bool count_zeroed, hipStream_t stream,
                      const PtsView *src) {
    if (n <= 0) return hipSuccess;
    int *spill_count = lists, *spill = lists + 1;
    if (!count_zeroed) k_zero_count<<<1, 1, 0, stream>>>(spill_count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    n_blocks = segment_grid(n_blocks);
    const PtsView in = src ? *src : pts;
#define ICEMPM_G2P(E, O, D)                                                                                      \
    k_g2p<E, O, D><<<n_blocks, kP2GThreads, 0, stream>>>(kp, pts, seg, grid, record_pq, clk, err, spill,           \
                                                         spill_count, in)
    // the adjoint's out-of-place passes never run with the damage model
    const bool ens = kp.ens_n > 0, dmg = kp.dmg_on != 0;
    if (src && dmg) return hipErrorInvalidValue;
    if (src) {
        if (ens) ICEMPM_G2P(true, true, false);
        else ICEMPM_G2P(false, true, false);
    } else if (dmg) {
        if (ens) ICEMPM_G2P(true, false, true);
        else ICEMPM_G2P(false, false, true);
    } else {
        if (ens) ICEMPM_G2P(true, false, false);
        else ICEMPM_G2P(false, false, false);
    }
#undef ICEMPM_G2P
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_g2p_spill<<<256, kP2GThreads, 0, stream>>>(kp, pts, grid, record_pq, clk, spill, spill_count, in);
    return hipGetLastError();
}

} // namespace icempm
