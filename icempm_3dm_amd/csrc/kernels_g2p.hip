// kernels_g2p.hip — grid-to-particle transfer and constitutive update
// (partition_kernel_g2p, mpm_kernels.cuh:237-351), one lane per point, one
// workgroup per brick segment with the segment's node velocities staged in
// LDS. The return mapping of the crushed points of a 256-point chunk
// (mpm_kernels.cuh:486-544) runs as a second phase of the workgroup, lanes
// re-packed over a workgroup-local list, so the SVD registers do not add to
// the gather's. Point transfer between partitions (:353-398) does not apply.
#include "g2p_point.h"

namespace icempm {

namespace {

__global__ void k_zero_count(int *c) { *c = 0; }

// One workgroup per brick segment; a point that drifted outside its
// segment's tile since the sort goes to the spill list.
__global__ __launch_bounds__(kP2GThreads) void k_g2p(KParams kp, PtsView pts, SegInfo seg,
                                                     const double *__restrict__ grid, int record_pq, int *err,
                                                     int *spill, int *spill_count) {
    __shared__ double sv[3][kTileAlloc];
    __shared__ CrushList cl;
    const int t = threadIdx.x;
    const int nseg = *seg.nseg;
    const double c4h = 4. * kp.h_inv;
    ChunkCursor c = cursor_first(seg, nseg, xcd_remap(blockIdx.x, gridDim.x), gridDim.x);
    while (c.valid()) {
        int org[3], bc[3];
        tile_origin(kp, c.brick, org, bc);
        if (c.first()) {
            // the previous segment's gathers are done (barrier in crush_list_build)
            VelTileRegs vt;
            vel_tile_fetch(vt, t, kp, org, grid);
            vel_tile_store(sv, vt, t);
            __syncthreads();
        }
        const int i = c.c0 + t;
        const bool valid = i < c.end;
        PointState s;
        s.util = ICEMPM_STATUS_DISABLED;
        if (valid) load_g2p_inputs(pts, i, s);
        const ChunkCursor nx = cursor_next(c, seg);
        bool spilled = false, crushed = false;
        if (!(s.util & ICEMPM_STATUS_DISABLED)) {
            int base[3];
            double fx[3];
            const int code = locate(kp, s.pos, base, fx);
            if (code) {
                raise_error(err, code);
            } else {
                const int rx = base[0] - org[0], ry = base[1] - org[1], rz = base[2] - org[2];
                if ((unsigned)rx <= kTile - 3 && (unsigned)ry <= kTile - 3 && (unsigned)rz <= kTile - 3) {
                    double w[3][3], d[3][3];
                    stencil_weights(fx, w, d);
                    gather_tile(sv, tile_index(rx, ry, rz), w, d, c4h, s.vel, s.Bp);
                    crushed = g2p_finish(kp, pts, i, s, record_pq);
                } else {
                    spilled = true;
                }
            }
        }
        wave_append(spilled, i, spill, spill_count);
        // return mapping of the chunk's crushed points, from the stored trial F
        crush_list_build(cl, t, crushed, i);
        if (t < cl.n) return_map_point(kp, pts, cl.idx[t]);
        c = nx;
    }
}

// G2P of the spilled points, gathering node velocities from HBM.
__global__ __launch_bounds__(kP2GThreads) void k_g2p_spill(KParams kp, PtsView pts, const double *__restrict__ grid,
                                                           int record_pq, const int *spill, const int *spill_count) {
    const int count = *spill_count;
    const double c4h = 4. * kp.h_inv;
    for (int k = blockIdx.x * kP2GThreads + threadIdx.x; k < count; k += gridDim.x * kP2GThreads) {
        const int i = spill[k];
        PointState s;
        load_g2p_inputs(pts, i, s);
        int base[3];
        double fx[3];
        (void)locate(kp, s.pos, base, fx); // listed points passed the bounds check
        double w[3][3], d[3][3];
        stencil_weights(fx, w, d);
        gather_global(kp, grid, base, w, d, c4h, s.vel, s.Bp);
        if (g2p_finish(kp, pts, i, s, record_pq)) return_map_point(kp, pts, i);
    }
}

} // namespace

hipError_t launch_g2p(const KParams &kp, PtsView pts, int n, SegInfo seg, const double *grid, int record_pq, int *err,
                      int *lists, int n_blocks, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    int *spill_count = lists, *spill = lists + 1;
    k_zero_count<<<1, 1, 0, stream>>>(spill_count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    n_blocks = (n_blocks + kXcd - 1) / kXcd * kXcd;
    k_g2p<<<n_blocks, kP2GThreads, 0, stream>>>(kp, pts, seg, grid, record_pq, err, spill, spill_count);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_g2p_spill<<<256, kP2GThreads, 0, stream>>>(kp, pts, grid, record_pq, spill, spill_count);
    return hipGetLastError();
}

} // namespace icempm
