// kernels_grid.hip — grid reset and grid update over the bricks that P2G
// touched, and a copy of the grid into the reference's x-major node order.
//
// The reference clears the whole pitched grid with a memset and updates
// every allocated node (gpu_partition_3d.cu:222-255). Here P2G appends every
// brick it writes to a list; reset and update walk only that list, one
// 64-lane wave per 4x4x4-node brick.
#include "kernels.h"

namespace icempm {

namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / 64;

// Zeroes the bricks of list `cur` and their flags, and the other list's count
// (the list the next P2G appends to); one wave per brick, grid-stride.
__device__ __forceinline__ void reset_bricks(const KParams &kp, const GridState &gs, int cur) {
    const int count = gs.brick_count[cur];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int32_t *list = gs.brick_list[cur];
    const int stride = gridDim.x * kWavesPerBlock;
    // list entries k0, k0 + stride, ... of the wave, up to 64 at a time, one
    // per lane (taken by readlane): the loop over bricks then only stores
    for (int k0 = blockIdx.x * kWavesPerBlock + wave; k0 < count; k0 += 64 * stride) {
        const int kl = k0 + lane * stride;
        const int mine = kl < count ? list[kl] : 0;
        asm volatile("" ::"v"(mine));
        const int nbatch = min(64, (count - k0 + stride - 1) / stride);
        for (int j = 0; j < nbatch; j++) {
            const int b = __builtin_amdgcn_readlane(mine, j);
            const long long idx = (long long)b * kBrickNodes + lane;
#pragma unroll
            for (int l = 0; l < 4; l++) gs.grid[l * kp.grid_pitch + idx] = 0.0;
            if (lane == 0) gs.brick_flag[b] = 0;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) gs.brick_count[cur ^ 1] = 0;
}

__global__ __launch_bounds__(kThreads) void k_reset_grid(KParams kp, GridState gs, int cur) {
    reset_bricks(kp, gs, cur);
}

__global__ void k_clear_all(KParams kp, GridState gs, long long nbricks) {
    long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nbricks * kBrickNodes; i += stride) {
#pragma unroll
        for (int l = 0; l < 4; l++) gs.grid[l * kp.grid_pitch + i] = 0.0;
        if (i < nbricks) gs.brick_flag[i] = 0;
        if (i < 2) gs.brick_count[i] = 0;
    }
}

__global__ void k_brick_info(KParams kp, int32_t *coord, int32_t *slot, long long nbricks) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x; b < nbricks; b += stride) {
        const int bz = (int)(b % kp.nbz), by = (int)((b / kp.nbz) % kp.nby), bx = (int)(b / ((long long)kp.nbz * kp.nby));
        coord[b] = bx | (by << 8) | (bz << 16);
        if (slot) slot[b] = ensemble_slot(kp, bx, by, bz);
    }
}

__global__ void k_clock_set(StepClock *clk, double t, long long step, int every) {
    clk->t = t;
    clk->step = step;
    clk->every = every;
    clk->record = 0;
}

// c0 + t * v rounded as two separate operations, as the host evaluates the
// indenter law (model_3d.cpp:38); no fused multiply-add.
__device__ __forceinline__ double law_center(double c0, double t, double v) {
#pragma clang fp contract(off)
    return c0 + t * v;
}

// partition_kernel_update_nodes (mpm_kernels.cuh:138-232) for the nodes of
// the bricks in list `cur`, one wave per brick at a time, grid-stride over
// the list; the indenter is the reference cylinder along z or a sphere.
template <bool kEns>
__device__ __forceinline__ void update_bricks(const KParams &kp0, const GridState &gs, int cur, double cx0, double cy0,
                                              double cz0, double t, double *bins0) {
    const int count = gs.brick_count[cur];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int32_t *list = gs.brick_list[cur];
    const int stride = gridDim.x * kWavesPerBlock;
    // The wave's bricks are list entries k0, k0 + stride, ...; up to 64 of
    // them are read at once, one per lane, and taken by readlane, and a
    // node's mass and momentum are requested together: one memory round trip
    // per brick instead of three (list entry, mass, momentum).
    for (int k0 = blockIdx.x * kWavesPerBlock + wave; k0 < count; k0 += 64 * stride) {
        const int kl = k0 + lane * stride;
        const int mine = kl < count ? list[kl] : 0;
        // the bricks' coordinates and scene slots, one per lane like the entries
        const int mine_coord = gs.brick_coord[mine];
        const int mine_slot = kEns ? gs.brick_slot[mine] : 0;
        // waited for here, once, not at the head of every brick below (where
        // the wait would also cover the previous brick's stores)
        asm volatile("" ::"v"(mine), "v"(mine_coord), "v"(mine_slot));
        const int nbatch = min(64, (count - k0 + stride - 1) / stride);
        for (int j = 0; j < nbatch; j++) {
            const int b = __builtin_amdgcn_readlane(mine, j);
            const int coord = __builtin_amdgcn_readlane(mine_coord, j);
            const long long idx = (long long)b * kBrickNodes + lane;
            const double mass = gs.grid[idx];
            double v[3];
#pragma unroll
            for (int i = 0; i < 3; i++) v[i] = gs.grid[(i + 1) * kp0.grid_pitch + idx];
            // keeps the momentum loads ahead of the branch below, where the
            // compiler would otherwise move them
            asm volatile("" ::"v"(v[0]), "v"(v[1]), "v"(v[2]));
            if (mass == 0) continue;
            const KParams kp = scene_view_slot<kEns>(kp0, __builtin_amdgcn_readlane(mine_slot, j));
            const double dt = kp.dt;
            const double cellsize = kp.h;
            const bool sphere = kp.ind_shape == ICEMPM_INDENTER_SPHERE;
            double cx = cx0, cy = cy0, cz = cz0;
            if constexpr (kEns) {
                // rounded as the one-scene path and the host law round it
                cx = law_center(kp.ind_c0[0], t, kp.ind_vel[0]);
                cy = law_center(kp.ind_c0[1], t, kp.ind_vel[1]);
                cz = law_center(kp.ind_c0[2], t, kp.ind_vel[2]);
            }
            double *bins = bins0 + kp.bins_off;
            const int bx = coord & 255, by = (coord >> 8) & 255, bz = coord >> 16;
            const int gs_[3] = {bx * 4 + (lane >> 4), by * 4 + ((lane >> 2) & 3), bz * 4 + (lane & 3)};
            const double gnpos[3] = {gs_[0] * cellsize, gs_[1] * cellsize, gs_[2] * cellsize};
            // node index within the scene's grid (walls, sticky band, force bins)
            const int gi[3] = {gs_[0] - kp.ox, gs_[1] - kp.oy, gs_[2] - kp.oz};

#pragma unroll
            for (int i = 0; i < 3; i++) v[i] /= mass;
            v[1] -= kp.dt_gravity;
            double vsq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            if (vsq > kp.vmax_sq) {
                double nrm = sqrt(vsq);
#pragma unroll
                for (int i = 0; i < 3; i++) v[i] = v[i] / nrm * kp.vmax;
            }

            if (kp.setup_type != 2) {
                double n[3] = {gnpos[0] - cx, gnpos[1] - cy, sphere ? gnpos[2] - cz : 0.0};
                double nsq = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
                if (nsq < kp.ind_rsq) {
                    double vrel[3] = {v[0] - kp.ind_vel[0], v[1] - kp.ind_vel[1], v[2] - kp.ind_vel[2]};
                    if (nsq > 0) {
                        double nrm = sqrt(nsq);
#pragma unroll
                        for (int i = 0; i < 3; i++) n[i] /= nrm;
                    }
                    double vn = vrel[0] * n[0] + vrel[1] * n[1] + vrel[2] * n[2];
                    if (vn < 0) {
                        double prev[3] = {v[0], v[1], v[2]};
#pragma unroll
                        for (int i = 0; i < 3; i++) v[i] = kp.ind_vel[i] + (vrel[i] - n[i] * vn);
                        double force[3];
#pragma unroll
                        for (int i = 0; i < 3; i++) force[i] = (prev[i] - v[i]) * mass / dt;
                        // no sensor bins to write when IndenterSubdivisions is 0
                        if (kp.ind_subdiv > 0) {
                        float angle = atan2f((float)n[0], (float)n[1]);
                        angle += ICEMPM_PI;
                        angle *= kp.ind_subdiv / (2 * ICEMPM_PI);
                        int index_angle = min(max((int)angle, 0), kp.ind_subdiv - 1);
                        int index_z = min(max(gi[2], 0), kp.gz - 1);
                        int index = index_z + index_angle * kp.gz;
#pragma unroll
                        for (int i = 0; i < 3; i++) unsafeAtomicAdd(&bins[i + 3 * index], force[i]);
                        }
                    }
                }
            }

            if (kp.setup_type != 2) {
                if (gi[1] <= 2) v[0] = v[1] = v[2] = 0;
            } else {
                if (gi[1] <= 2 && v[1] < 0) v[1] = 0;
            }
            if (gi[1] >= kp.gy - 3 && v[1] > 0) v[1] = 0;
            if (gi[0] <= 2 && v[0] < 0) v[0] = 0;
            else if (gi[0] >= kp.gx - 3 && v[0] > 0) v[0] = 0;
            if (gi[2] <= 2 && v[2] < 0) v[2] = 0;
            else if (gi[2] >= kp.gz - 3 && v[2] > 0) v[2] = 0;

            if (kp.setup_type == 2 && kp.drag != 0) {
#pragma unroll
                for (int i = 0; i < 3; i++) v[i] *= (1 - kp.drag);
            }
#pragma unroll
            for (int i = 0; i < 3; i++) gs.grid[(i + 1) * kp.grid_pitch + idx] = v[i];
        }
    }
}

// One scene: the indenter center is (cx, cy, cz). Ensembles: each scene's
// center is its c0 + t * velocity, walls and force bins are the scene's.
// The work of UpdateExtras x follows in the same launch: the reset touches
// another grid, and the clock slot written is not the one read.
template <bool kEns>
__global__ __launch_bounds__(kThreads) void k_update_nodes(KParams kp0, GridState gs, int cur, double cx0,
                                                           double cy0, double cz0, double t, double *bins0,
                                                           UpdateExtras x) {
    if (x.clk_in) {
        t = x.clk_in->t + kp0.dt;
        cx0 = law_center(kp0.ind_c0[0], t, kp0.ind_vel[0]);
        cy0 = law_center(kp0.ind_c0[1], t, kp0.ind_vel[1]);
        cz0 = law_center(kp0.ind_c0[2], t, kp0.ind_vel[2]);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            const long long step = x.clk_in->step;
            const int every = x.clk_in->every;
            x.clk_out->t = t;
            x.clk_out->step = step + 1;
            x.clk_out->every = every;
            x.clk_out->record = (step % every) == 0 ? 1 : 0;
        }
    }
    if (x.zero && blockIdx.x == 0 && threadIdx.x == 0) *x.zero = 0;
    update_bricks<kEns>(kp0, gs, cur, cx0, cy0, cz0, t, bins0);
    if (x.reset.grid) reset_bricks(kp0, x.reset, x.reset_cur);
}

__global__ void k_grid_to_xmajor(KParams kp, const double *grid, double *out) {
    const long long nn = (long long)kp.gx * kp.gy * kp.gz;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += stride) {
        int z = (int)(i % kp.gz);
        int y = (int)((i / kp.gz) % kp.gy);
        int x = (int)(i / ((long long)kp.gz * kp.gy));
        long long src = node_index(kp, x, y, z);
#pragma unroll
        for (int l = 0; l < 4; l++) out[l * nn + i] = grid[l * kp.grid_pitch + src];
    }
}

} // namespace

hipError_t launch_reset_grid(const KParams &kp, GridState gs, int cur, hipStream_t stream) {
    k_reset_grid<<<1024, kThreads, 0, stream>>>(kp, gs, cur);
    return hipGetLastError();
}

hipError_t launch_clear_grid_all(const KParams &kp, GridState gs, long long nbricks, hipStream_t stream) {
    k_clear_all<<<2048, 256, 0, stream>>>(kp, gs, nbricks);
    return hipGetLastError();
}

hipError_t launch_update_nodes(const KParams &kp, GridState gs, int cur, double cx, double cy, double cz, double t,
                               double *bins, const UpdateExtras &x, hipStream_t stream) {
    if (kp.ens_n > 0) k_update_nodes<true><<<1024, kThreads, 0, stream>>>(kp, gs, cur, cx, cy, cz, t, bins, x);
    else k_update_nodes<false><<<1024, kThreads, 0, stream>>>(kp, gs, cur, cx, cy, cz, t, bins, x);
    return hipGetLastError();
}

hipError_t launch_brick_info(const KParams &kp, int32_t *brick_coord, int32_t *brick_slot, long long nbricks,
                             hipStream_t stream) {
    if (nbricks <= 0) return hipSuccess;
    k_brick_info<<<1024, 256, 0, stream>>>(kp, brick_coord, brick_slot, nbricks);
    return hipGetLastError();
}

hipError_t launch_clock_set(StepClock *clk, double t, long long step, int every, hipStream_t stream) {
    k_clock_set<<<1, 1, 0, stream>>>(clk, t, step, every);
    return hipGetLastError();
}


hipError_t launch_grid_to_xmajor(const KParams &kp, const double *grid, double *out, hipStream_t stream) {
    k_grid_to_xmajor<<<2048, 256, 0, stream>>>(kp, grid, out);
    return hipGetLastError();
}

} // namespace icempm
