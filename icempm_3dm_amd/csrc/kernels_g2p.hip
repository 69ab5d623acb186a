// kernels_g2p.hip — grid-to-particle transfer and constitutive update
// (partition_kernel_g2p, mpm_kernels.cuh:237-351), one lane per particle.
// Point transfer between partitions (:353-398) does not apply: one partition.
//
// The SVD-based return mapping (ComputeSVD + Wolper_Drucker_Prager,
// mpm_kernels.cuh:486-544) only runs for crushed particles. k_g2p appends
// them to a worklist (one atomic per wave) after writing the trial state,
// and k_return_map finishes them, so the common elastic path carries no SVD
// registers. Both kernels compute the same values the fused reference
// kernel does, in the same order.
#include "constitutive.h"
#include "kernels.h"

namespace icempm {

namespace {

constexpr int kThreads = 256;

__global__ void k_zero_count(int *count) { *count = 0; }

__global__ __launch_bounds__(kThreads) void k_g2p(KParams kp, PtsView pts, int n, const double *__restrict__ grid,
                                                  int record_pq, int *err, int *work, int *work_count) {
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int i = lb * kThreads + threadIdx.x;
    bool crushed = false;
    if (i < n) {
        PState p;
        p.util = double_to_bits(pts.at(ICEMPM_IDX_UTILITY, i));
        p.p_tr = 0.0;
        p.q_tr = 0.0;
        double pos[3];
#pragma unroll
        for (int a = 0; a < 3; a++) pos[a] = pts.at(ICEMPM_POSX + a, i);
        int base[3];
        double fx[3];
        const int code = (p.util & ICEMPM_STATUS_DISABLED) ? -1 : locate(kp, pos, base, fx);
        if (code > 0) raise_error(err, code);
        if (code == 0) {
            const bool liquid = p.util & ICEMPM_STATUS_LIQUID;
            double w[3][3], d[3][3];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                bspline3(fx[a], w[a]);
#pragma unroll
                for (int o = 0; o < 3; o++) d[a][o] = o - fx[a];
            }
            // separable brick addressing: node index = X[a] + Y[b] + Z[c]
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
            const double c4h = 4. * kp.h_inv;

            double vel[3] = {0, 0, 0};
            double Bp[9];
#pragma unroll
            for (int k = 0; k < 9; k++) Bp[k] = 0.0;
#pragma unroll 1
            for (int a = 0; a < 3; a++) {
                const double wa = w[0][a];
                const double da = d[0][a];
#pragma unroll
                for (int b = 0; b < 3; b++)
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        const double weight = wa * w[1][b] * w[2][c];
                        const long long idx = X[a] + Y[b] + Z[c];
                        const double nv[3] = {gvx[idx], gvy[idx], gvz[idx]};
                        const double dp[3] = {da, d[1][b], d[2][c]};
                        const double s = c4h * weight;
#pragma unroll
                        for (int r = 0; r < 3; r++) {
                            vel[r] += weight * nv[r];
                            const double sr = s * nv[r];
#pragma unroll
                            for (int q = 0; q < 3; q++) Bp[3 * r + q] += sr * dp[q];
                        }
                    }
            }

            // advection and deformation gradient update (mpm_kernels.cuh:308-309)
            const double dt = kp.dt;
#pragma unroll
            for (int a = 0; a < 3; a++) pos[a] += vel[a] * dt;
#pragma unroll
            for (int k = 0; k < 9; k++) p.Fe[k] = pts.at(ICEMPM_FE00 + k, i);
            p.Jp_inv = pts.at(ICEMPM_IDX_JP_INV, i);
            double A[9];
#pragma unroll
            for (int k = 0; k < 9; k++) A[k] = dt * Bp[k] + ((k % 4 == 0) ? 1.0 : 0.0);
            {
                double Fn[9];
#pragma unroll
                for (int r = 0; r < 3; r++)
#pragma unroll
                    for (int c = 0; c < 3; c++)
                        Fn[3 * r + c] = A[3 * r] * p.Fe[c] + A[3 * r + 1] * p.Fe[3 + c] + A[3 * r + 2] * p.Fe[6 + c];
#pragma unroll
                for (int k = 0; k < 9; k++) p.Fe[k] = Fn[k];
            }

            if (liquid) {
                p.Jp_inv *= det3(A);
            } else {
                compute_pq(kp, p);
                if (!(p.util & ICEMPM_STATUS_CRUSHED)) check_failure_surface(kp, p);
                crushed = (p.util & ICEMPM_STATUS_CRUSHED) != 0;
            }

#pragma unroll
            for (int a = 0; a < 3; a++) {
                pts.at(ICEMPM_POSX + a, i) = pos[a];
                pts.at(ICEMPM_VELX + a, i) = vel[a];
            }
#pragma unroll
            for (int k = 0; k < 9; k++) {
                pts.at(ICEMPM_FE00 + k, i) = p.Fe[k];
                pts.at(ICEMPM_BP00 + k, i) = Bp[k];
            }
            pts.at(ICEMPM_IDX_JP_INV, i) = p.Jp_inv;
            pts.at(ICEMPM_IDX_UTILITY, i) = bits_to_double(p.util);
            if (record_pq) {
                pts.at(ICEMPM_IDX_P, i) = p.p_tr;
                pts.at(ICEMPM_IDX_Q, i) = p.q_tr;
            }
        }
    }
    // crushed particles -> worklist, one atomic per wave
    const unsigned long long cm = __ballot(crushed);
    if (cm) {
        const int lane = __lane_id();
        const int leader = __ffsll((long long)cm) - 1;
        int base_slot = 0;
        if (lane == leader) base_slot = atomicAdd(work_count, __popcll(cm));
        base_slot = __shfl(base_slot, leader, 64);
        if (crushed) work[base_slot + __popcll(cm & ((1ull << lane) - 1ull))] = i;
    }
}

// ComputePQ + ComputeSVD + Wolper_Drucker_Prager for the listed particles,
// starting from the trial Fe that k_g2p stored.
__global__ __launch_bounds__(kThreads) void k_return_map(KParams kp, PtsView pts, const int *work,
                                                         const int *work_count) {
    const int count = *work_count;
    for (int k = blockIdx.x * kThreads + threadIdx.x; k < count; k += gridDim.x * kThreads) {
        const int i = work[k];
        PState p;
#pragma unroll
        for (int q = 0; q < 9; q++) p.Fe[q] = pts.at(ICEMPM_FE00 + q, i);
        p.Jp_inv = pts.at(ICEMPM_IDX_JP_INV, i);
        p.util = 0;
        compute_pq(kp, p);
        return_map(kp, p);
#pragma unroll
        for (int q = 0; q < 9; q++) pts.at(ICEMPM_FE00 + q, i) = p.Fe[q];
        pts.at(ICEMPM_IDX_JP_INV, i) = p.Jp_inv;
    }
}

} // namespace

hipError_t launch_g2p(const KParams &kp, PtsView pts, int n, const double *grid, int record_pq, int *err, int *work,
                      int *work_count, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_zero_count<<<1, 1, 0, stream>>>(work_count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int nb = (n + kThreads - 1) / kThreads;
    nb = (nb + kXcd - 1) / kXcd * kXcd;
    k_g2p<<<nb, kThreads, 0, stream>>>(kp, pts, n, grid, record_pq, err, work, work_count);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_return_map<<<512, kThreads, 0, stream>>>(kp, pts, work, work_count);
    return hipGetLastError();
}

} // namespace icempm
