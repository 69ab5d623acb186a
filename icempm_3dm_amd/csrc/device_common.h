// device_common.h — kernel-side parameter block, grid/brick addressing and
// small FP64 helpers shared by the HIP kernels of the MPM step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "icempm_3dm.h"

namespace icempm {

// Grid nodes are stored in bricks of 4x4x4 nodes. Brick (bx,by,bz) has id
// (bx*nby + by)*nbz + bz; node (x,y,z) sits at brick_id*64 + ((x&3)*4 + (y&3))*4 + (z&3)
// of each of the 4 grid lines (m, vx, vy, vz; line l at grid + l*grid_pitch).
constexpr int kBrick = 4;
constexpr int kBrickNodes = 64;

// Kernel-argument view of SimParams3D plus the indenter law. Passed by value
// so that every context carries its own constants (the reference keeps one
// __constant__ SimParams3D per device, mpm_kernels.cuh:18).
struct KParams {
    int gx, gy, gz;    // GridXTotal, GridY, GridZ of the scene (nodes)
    int ox, oy, oz;    // the scene's node (0,0,0) in the stored grid (0 for one scene)
    int nbx, nby, nbz; // bricks per axis of the stored grid
    long long grid_pitch;

    double h, h_inv, dt;
    double mass_solid, mass_liquid;
    double dt_vol_Dpinv, dt_gravity, vmax, vmax_sq;
    double mu, kappa;
    double ice_comp, ice_tens, ice_shear, ice_tens2, nacc_beta;
    double tan_phi, dp_threshold_p, grain_var;
    // Tension-branch constants of the return mapping (mpm_kernels.cuh:501-503):
    // Je_new = sqrt(2 DP_threshold_p / kappa + 1) and its cube root.
    double rm_je_new, rm_cbrt_je_new;
    double drag;
    int setup_type;
    // damage model: on/off, 1 / softening strain, 1 - residual cohesion
    int dmg_on;
    double dmg_rate, dmg_soft;
    // The return mapping's constants in device memory (kRmc* below). Only
    // crushed points read them, through this pointer at the point of use:
    // as kernel arguments they would be held in scalar registers for the
    // whole fused kernel, which then spills.
    const double *rmc;
    // Step counters of the context (kStat* below), added to by atomics on the
    // rare paths they count; nullptr: not counted.
    unsigned long long *stats;

    int ind_shape, ind_subdiv;
    double ind_rsq;
    double ind_vel[3];
    double ind_c0[3];   // indenter center at t = 0 in stored-grid coordinates (ensembles)
    long long bins_off; // the scene's first indenter accumulator

    // Ensembles: independent scenes, each in its own slot of the stored grid
    // (slots are whole bricks, with at least one empty brick between scene
    // boxes, so no stencil crosses from one scene to another). ens_n = 0 for
    // one scene. A brick's scene block is ens_slots[slot of the brick].
    int ens_n;
    int ens_slot_b[3]; // slot size in bricks
    int ens_nslot[3];  // slots per axis
    const KParams *ens_slots;
};

// Layout of KParams::stats: points that a P2G scatter (plain, fused or the
// adjoint's) found outside its segment's tile and sent straight to HBM, and
// points that a G2P found outside the tile and took the spill path; both
// summed since allocation.
constexpr int kStatP2GOutOfTile = 0, kStatG2PSpilled = 1;
constexpr int kStatCount = 2;

__device__ __forceinline__ void count_stat(const KParams &kp, int which, unsigned long long n) {
    if (kp.stats && n) atomicAdd(kp.stats + which, n);
}

// Layout of KParams::rmc.
constexpr int kRmcThresholdP = 0, kRmcTanPhi = 1, kRmcJeNew = 2, kRmcCbrtJeNew = 3, kRmcDmgRate = 4, kRmcDmgSoft = 5;
constexpr int kRmcSize = 6;

// Ensemble slot of the brick with coordinates (bx, by, bz).
__device__ __forceinline__ int ensemble_slot(const KParams &kp, int bx, int by, int bz) {
    return ((bx / kp.ens_slot_b[0]) * kp.ens_nslot[1] + by / kp.ens_slot_b[1]) * kp.ens_nslot[2] +
           bz / kp.ens_slot_b[2];
}

// Parameter block of ensemble slot `slot`.
__device__ __forceinline__ KParams scene_of_slot(const KParams &kp, int slot) {
    // The table is written by the host only, so it is read through the
    // constant address space: with a workgroup-uniform brick the compiler
    // then uses scalar loads and scalar registers instead of ~40 VGPRs.
    static_assert(sizeof(KParams) % 8 == 0, "copied in 8-byte words");
    using ConstWord = __attribute__((address_space(4))) const unsigned long long;
    ConstWord *src = (ConstWord *)(kp.ens_slots + slot);
    unsigned long long w[sizeof(KParams) / 8];
#pragma unroll
    for (int k = 0; k < (int)(sizeof(KParams) / 8); k++) w[k] = src[k];
    KParams out;
    __builtin_memcpy(&out, w, sizeof(KParams));
    return out;
}

// Parameter block of the scene that owns grid brick b (ensembles).
__device__ __forceinline__ KParams scene_of_brick(const KParams &kp, long long b) {
    const int bz = (int)(b % kp.nbz);
    const int by = (int)((b / kp.nbz) % kp.nby);
    const int bx = (int)(b / ((long long)kp.nbz * kp.nby));
    return scene_of_slot(kp, ensemble_slot(kp, bx, by, bz));
}

// The scene block for work on brick b: kp itself for one scene.
template <bool kEns>
__device__ __forceinline__ KParams scene_view(const KParams &kp, long long b) {
    if constexpr (kEns) return scene_of_brick(kp, b);
    else return kp;
}

// The same for a brick segment whose ensemble slot is known.
template <bool kEns>
__device__ __forceinline__ KParams scene_view_slot(const KParams &kp, int slot) {
    if constexpr (kEns) return scene_of_slot(kp, slot);
    else return kp;
}

// The scene block of a point at pos (ensembles; the point's base node picks
// the brick, clamped into the stored grid).
__device__ __forceinline__ KParams scene_of_point(const KParams &kp, const double pos[3]) {
    const int nb[3] = {kp.nbx, kp.nby, kp.nbz};
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double x = pos[d] * kp.h_inv - 0.5;
        if (!(x >= 0.0)) x = 0.0;
        if (x > (double)(nb[d] * 4 - 1)) x = (double)(nb[d] * 4 - 1);
        c[d] = (int)x >> 2;
    }
    return scene_of_brick(kp, ((long long)c[0] * kp.nby + c[1]) * kp.nbz + c[2]);
}

// Particle storage: 28 SoA lines of `pitch` doubles (HostSideSOA line order).
// A view is a kernel argument (uniform across the workgroup); pitch is at
// most kMaxPointPitch.
constexpr long long kMaxPointPitch = 1ll << 29;
struct PtsView {
    double *lines; // line k at lines + k*pitch
    long long pitch;
    // Element i of line `line`. The line's start is formed in scalar
    // registers where it is used, and the lane's element is a 32-bit byte
    // offset from it (the scalar-base + vector-offset form of the global
    // memory instructions, no vector address arithmetic). The empty asm
    // statements keep the compiler from computing the starts of all lines
    // once and holding them across the kernel (28 scalar register pairs,
    // which in the segment kernels spill into vector registers and come back
    // with two lane reads per access) and from folding the lane offset into
    // a 64-bit vector address.
    __device__ __forceinline__ double &at(int line, int i) const {
        using GDouble = __attribute__((address_space(1))) double;
        using GChar = __attribute__((address_space(1))) char;
        long long p = pitch;
        asm volatile("" : "+s"(p));
        GDouble *row = (GDouble *)lines + (long long)line * p;
        asm volatile("" : "+s"(row));
        return *(double *)((GChar *)row + (unsigned)i * 8u);
    }
};

__device__ __forceinline__ long long node_index(const KParams &kp, int x, int y, int z) {
    long long b = ((long long)(x >> 2) * kp.nby + (y >> 2)) * kp.nbz + (z >> 2);
    return b * kBrickNodes + (((x & 3) << 4) | ((y & 3) << 2) | (z & 3));
}

__device__ __forceinline__ long long brick_of(const KParams &kp, int x, int y, int z) {
    return ((long long)(x >> 2) * kp.nby + (y >> 2)) * kp.nbz + (z >> 2);
}

__device__ __forceinline__ double bits_to_double(long long v) { return __longlong_as_double(v); }
__device__ __forceinline__ long long double_to_bits(double d) { return __double_as_longlong(d); }

// Device error indicator: keeps the largest code raised.
__device__ __forceinline__ void raise_error(int *err, int code) { atomicMax(err, code); }

// Base node, fractional offset and error code for a particle position
// (mpm_kernels.cuh:65-75). Returns 0 if the 3x3x3 stencil lies in the
// scene's grid (nodes [o, o + g) per axis).
__device__ __forceinline__ int locate(const KParams &kp, const double pos[3], int base[3], double fx[3]) {
    double xi[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        xi[d] = pos[d] * kp.h_inv;
        if (!(fabs(xi[d]) <= 1e9)) return ICEMPM_DEVERR_NONFINITE; // also NaN
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        base[d] = (int)(xi[d] - 0.5);
        fx[d] = xi[d] - (double)base[d];
    }
    int code = 0;
    const int g[3] = {kp.gx, kp.gy, kp.gz};
    const int o[3] = {kp.ox, kp.oy, kp.oz};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (base[d] < o[d]) code = max(code, 70 + d);
        if (base[d] > o[d] + g[d] - 3) code = max(code, 73 + d);
    }
    return code;
}

// Quadratic B-spline weights of one axis (mpm_kernels.cuh:78-81).
__device__ __forceinline__ void bspline3(double f, double w[3]) {
    double v0 = 1.5 - f, v1 = f - 1.0, v2 = f - 0.5;
    w[0] = 0.5 * v0 * v0;
    w[1] = 0.75 - v1 * v1;
    w[2] = 0.5 * v2 * v2;
}

// 3x3 determinant in Eigen's order (cofactor expansion along row 0).
__device__ __forceinline__ double det3(const double m[9]) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// Orders the LDS accesses of a wavefront around this point: the fence keeps
// the compiler from moving memory operations across it, and the LDS executes
// one wavefront's operations in issue order, so writes before it are seen by
// every lane's reads after it (no instruction is needed at wavefront scope).
__device__ __forceinline__ void wave_lds_sync() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

// Maps a linear block index so that each of the 8 XCDs (which receive
// blocks round-robin) works on one contiguous range of logical blocks.
// nblocks must be a multiple of 8 (launchers round grids up).
constexpr int kXcd = 8;
__device__ __forceinline__ int xcd_remap(int block, int nblocks) {
    return (block & (kXcd - 1)) * (nblocks / kXcd) + (block / kXcd);
}

} // namespace icempm
