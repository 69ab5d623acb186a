// kernels_p2g.hip — particle-to-grid transfer (partition_kernel_p2g,
// mpm_kernels.cuh:20-100) for brick-sorted points: one workgroup per brick
// segment, LDS-tiled scatter (p2g_tile.h). The next sub-chunk's point data
// is loaded while the current one is scattered.
#include "p2g_tile.h"

namespace icempm {

namespace {

// What P2G reads of a point: utility, pos, vel, Fe, C (+ Jp_inv if liquid).
struct P2GInputs {
    long long util;
    double pos[3], vel[3], Fe[9], Bp[9];
    double Jp;
};

__device__ __forceinline__ void load_p2g_inputs(PtsView pts, int i, bool valid, P2GInputs &r) {
    r.util = ICEMPM_STATUS_DISABLED;
    r.Jp = 1.0;
    if (!valid) return;
    r.util = double_to_bits(pts.at(ICEMPM_IDX_UTILITY, i));
#pragma unroll
    for (int a = 0; a < 3; a++) {
        r.pos[a] = pts.at(ICEMPM_POSX + a, i);
        r.vel[a] = pts.at(ICEMPM_VELX + a, i);
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
        r.Fe[k] = pts.at(ICEMPM_FE00 + k, i);
        r.Bp[k] = pts.at(ICEMPM_BP00 + k, i);
    }
    if (r.util & ICEMPM_STATUS_LIQUID) r.Jp = pts.at(ICEMPM_IDX_JP_INV, i);
}

template <bool kEns>
__global__ __launch_bounds__(kP2GThreads) void k_p2g(KParams kp, PtsView pts, SegInfo seg, GridState gs, int cur,
                                                     int *err) {
    __shared__ P2GShared sh;
    const int t = threadIdx.x;
    const BrickList bl = brick_list(gs, cur);
    const int nseg = *seg.nseg;
    ChunkCursor c = cursor_first(seg, nseg, xcd_remap(blockIdx.x, gridDim.x), gridDim.x);
    P2GInputs in;
    if (c.valid()) load_p2g_inputs(pts, c.c0 + t, c.c0 + t < c.end, in);
    while (c.valid()) {
        const KParams ks = scene_view_slot<kEns>(kp, c.slot);
        int org[3], bc[3];
        tile_origin_packed(c.coord, org, bc);
        if (c.first()) p2g_zero_tile(sh, t);
        const double Je = (in.util & ICEMPM_STATUS_DISABLED) ? 1.0 : det3(in.Fe);
        p2g_record(sh, t, ks, c.c0 + t < c.end, in.util, in.pos, in.vel, in.Fe, Je, rcbrt(Je * Je), in.Bp, in.Jp,
                   org, err);
        const ChunkCursor nx = cursor_next(c, seg);
        if (nx.valid()) load_p2g_inputs(pts, nx.c0 + t, nx.c0 + t < nx.end, in);
        __syncthreads();
        p2g_scatter(sh, t, min(kP2GChunk, c.end - c.c0), ks, gs, bl);
        if (c.last()) p2g_flush(sh, t, ks, org, bc, gs, bl);
        c = nx;
    }
}

// wg_start[b], b in [0, nblocks]: one thread per workgroup boundary.
__global__ void k_work_split(const int32_t *seg_start, const int32_t *nseg_p, int32_t *wg_start, int nblocks) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nblocks) return;
    const int nseg = *nseg_p;
    const long long total = (long long)seg_start[nseg] + kSegCost * nseg;
    wg_start[b] = (int32_t)(b == nblocks ? nseg : seg_at_cost(seg_start, nseg, total * b / nblocks));
}

} // namespace

hipError_t launch_p2g(const KParams &kp, PtsView pts, int n, SegInfo seg, GridState gs, int cur, int *err,
                      int n_blocks, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    n_blocks = segment_grid(n_blocks);
    if (kp.ens_n > 0) k_p2g<true><<<n_blocks, kP2GThreads, 0, stream>>>(kp, pts, seg, gs, cur, err);
    else k_p2g<false><<<n_blocks, kP2GThreads, 0, stream>>>(kp, pts, seg, gs, cur, err);
    return hipGetLastError();
}

hipError_t launch_work_split(const int32_t *seg_start, const int32_t *nseg, int32_t *wg_start, int nblocks,
                             hipStream_t stream) {
    k_work_split<<<(nblocks + 1 + 255) / 256, 256, 0, stream>>>(seg_start, nseg, wg_start, nblocks);
    return hipGetLastError();
}

} // namespace icempm
