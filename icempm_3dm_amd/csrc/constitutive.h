// constitutive.h — device-side stress, crush test and Drucker-Prager return
// mapping of the reference (mpm_kernels.cuh:434-581), plus an FP64 3x3 SVD.
// Matrices are row-major double[9]. Expression order follows the reference
// so that the device agrees with a host restatement to rounding level.
#pragma once

#include "device_common.h"

namespace icempm {

constexpr double kCoeff1 = 1.224744871391589;    // sqrt(3/2), mpm_kernels.cuh:11
constexpr double kCoeff1Inv = 0.8164965809277260; // sqrt(2/3), mpm_kernels.cuh:12

// b = F F^T (row-major, symmetric: the six distinct dot products, the
// mirrored entries copied, which is bitwise what the nine-product form gives).
__device__ __forceinline__ void left_cauchy_green(const double F[9], double b[9]) {
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = r; c < 3; c++)
            b[3 * r + c] = F[3 * r] * F[3 * c] + F[3 * r + 1] * F[3 * c + 1] + F[3 * r + 2] * F[3 * c + 2];
    b[3] = b[1];
    b[6] = b[2];
    b[7] = b[5];
}

// KirchhoffStress_Wolper (mpm_kernels.cuh:547-557):
// tau = mu J^{-2/3} dev(F F^T) + (kappa/2)(J^2 - 1) I, with J = det F and
// rc = rcbrt(J^2) given (compute_pq yields both for the same F).
__device__ __forceinline__ void kirchhoff_stress_j(const KParams &kp, const double F[9], double Je, double rc,
                                                   double P[9]) {
    double b[9];
    left_cauchy_green(F, b);
    double tr3 = (b[0] + b[4] + b[8]) / 3;
    b[0] -= tr3;
    b[4] -= tr3;
    b[8] -= tr3;
    double s = kp.mu * rc;
    double pterm = kp.kappa * 0.5 * (Je * Je - 1.);
#pragma unroll
    for (int k = 0; k < 9; k++) P[k] = s * b[k];
    P[0] += pterm;
    P[4] += pterm;
    P[8] += pterm;
}

__device__ __forceinline__ void kirchhoff_stress(const KParams &kp, const double F[9], double P[9]) {
    const double Je = det3(F);
    kirchhoff_stress_j(kp, F, Je, rcbrt(Je * Je), P);
}

// Water (mpm_kernels.cuh:559-566): kappa (1 - J^-3) I
__device__ __forceinline__ void water_stress(const KParams &kp, double J, double P[9]) {
    double v = kp.kappa * (1. - pow(J, -3.0));
#pragma unroll
    for (int k = 0; k < 9; k++) P[k] = 0.0;
    P[0] = P[4] = P[8] = v;
}

// ---------------------------------------------------------------- SVD
// Reciprocal and reciprocal square root to about 1 ulp: the hardware
// estimates (relative error below 2^-22) refined by one third-order step,
// y (1 + e + e^2) with e = 1 - x y for 1/x, and y (1 + r/2 + 3r^2/8) with
// r = 1 - x y^2 for 1/sqrt(x). Arguments are finite and positive.
__device__ __forceinline__ double rcp_f64(double x) {
    const double y = __builtin_amdgcn_rcp(x);
    const double e = fma(-x, y, 1.0);
    return fma(y, fma(e, e, e), y);
}

__device__ __forceinline__ double rsq_f64(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double r = fma(-(x * y), y, 1.0);
    return fma(y * r, fma(r, 0.375, 0.5), y);
}

// sqrt(x) for x >= 0 (0 maps to 0; negative or NaN x gives NaN).
__device__ __forceinline__ double sqrt_f64(double x) { return x == 0.0 ? 0.0 : x * rsq_f64(x); }

// One one-sided (Hestenes) Jacobi rotation of the column pair (p, q) of A:
// with alpha = |a_p|^2, beta = |a_q|^2 and gamma = a_p . a_q recomputed from
// the columns themselves, the plane rotation J with
// t = tan(phi) = sgn(d) 2 gamma / (|d| + sqrt(d^2 + 4 gamma^2)), d = beta - alpha
// (the root of t^2 + 2 theta t - 1 = 0 with |phi| <= pi/4), c = 1/sqrt(1 + t^2),
// s = t c, makes the two columns of A J orthogonal. V accumulates J in its
// columns p and q. Returns whether the pair was already orthogonal to
// working precision: |cos| = |gamma| / sqrt(alpha beta) <= 6.3e-16 (about 6
// ulps; recomputed from rounded columns, |cos| does not fall much below a
// few ulps, so a tighter test would only run the sweep limit).
__device__ __forceinline__ bool hestenes_rotate(double A[9], double V[9], int p, int q) {
    double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        al = fma(A[3 * r + p], A[3 * r + p], al);
        be = fma(A[3 * r + q], A[3 * r + q], be);
        ga = fma(A[3 * r + p], A[3 * r + q], ga);
    }
    const double d = be - al;
    const double e2 = ga + ga;
    const double r2 = fma(d, d, e2 * e2);
    const double t0 = (d < 0.0 ? -e2 : e2) * rcp_f64(fabs(d) + r2 * rsq_f64(r2));
    const double t = ga == 0.0 ? 0.0 : t0;
    const double c = rsq_f64(fma(t, t, 1.0));
    const double sn = t * c;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const double ap = A[3 * r + p], aq = A[3 * r + q];
        A[3 * r + p] = c * ap - sn * aq;
        A[3 * r + q] = sn * ap + c * aq;
        const double vp = V[3 * r + p], vq = V[3 * r + q];
        V[3 * r + p] = c * vp - sn * vq;
        V[3 * r + q] = sn * vp + c * vq;
    }
    return ga * ga <= 4e-31 * (al * be);
}

__device__ __forceinline__ void swap_negate_cols(double V[9], int a, int b) {
    // exchanging two columns flips det V; negating one restores +1
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const double t = V[3 * r + a];
        V[3 * r + a] = V[3 * r + b];
        V[3 * r + b] = -t;
    }
}

__device__ __forceinline__ double col_norm2(const double A[9], int c) {
    return A[c] * A[c] + A[3 + c] * A[3 + c] + A[6 + c] * A[6 + c];
}

// F = U diag(S) V^T with U, V rotations (columns U[3r+k], V[3r+k]),
// S[0] >= S[1] >= |S[2]| and S[2] < 0 iff det F < 0 — the convention of the
// reference's svd3x3 (mpm_kernels.cuh:585-601), in FP64 instead of its FP32
// McAdams routine (helper_math.cuh:23-1119).
// One-sided Jacobi on F itself: cyclic sweeps of column rotations of
// B = F V until every column pair is orthogonal to working precision, so the
// singular values and vectors carry errors of rounding times cond(F), not
// cond(F)^2 as an eigen-decomposition of F^T F would. Columns are then sorted
// by norm, descending (det V kept +1), and U is taken by Gram-Schmidt on the
// columns of B, its third column the cross product of the first two, so that
// S[2] = U_2 . B_2 carries the sign of det F. S2 returns S^2.
// svd3_inplace takes B = F and leaves F V in it.
__device__ __forceinline__ void svd3_inplace(double B[9], double U[9], double S[3], double S2[3], double V[9]) {
#pragma unroll
    for (int k = 0; k < 9; k++) V[k] = (k % 4 == 0) ? 1.0 : 0.0;
#pragma unroll 1
    for (int sweep = 0; sweep < 12; sweep++) {
        const bool c01 = hestenes_rotate(B, V, 0, 1);
        const bool c02 = hestenes_rotate(B, V, 0, 2);
        const bool c12 = hestenes_rotate(B, V, 1, 2);
        if (c01 && c02 && c12) break; // also ends on NaN after 12 sweeps
    }
    double n0 = col_norm2(B, 0), n1 = col_norm2(B, 1), n2 = col_norm2(B, 2);
    if (n0 < n1) {
        const double t = n0; n0 = n1; n1 = t;
        swap_negate_cols(V, 0, 1);
        swap_negate_cols(B, 0, 1);
    }
    if (n0 < n2) {
        const double t = n0; n0 = n2; n2 = t;
        swap_negate_cols(V, 0, 2);
        swap_negate_cols(B, 0, 2);
    }
    if (n1 < n2) {
        const double t = n1; n1 = n2; n2 = t;
        swap_negate_cols(V, 1, 2);
        swap_negate_cols(B, 1, 2);
    }
    // first column
    const double bb0 = n0;
    double u0[3] = {1.0, 0.0, 0.0};
    S[0] = 0.0;
    if (bb0 > 0.0) {
        const double r0 = rsq_f64(bb0);
        S[0] = bb0 * r0;
#pragma unroll
        for (int k = 0; k < 3; k++) u0[k] = B[3 * k] * r0;
    }
    // second column, orthogonalised against the first
    const double pr = u0[0] * B[1] + u0[1] * B[4] + u0[2] * B[7];
    double w[3];
#pragma unroll
    for (int k = 0; k < 3; k++) w[k] = fma(-pr, u0[k], B[3 * k + 1]);
    double ww = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    if (!(ww > 1e-30 * bb0)) {
        // rank <= 1: any unit vector orthogonal to u0 (cross with the axis
        // on which u0 is smallest)
        const double ax = fabs(u0[0]), ay = fabs(u0[1]), az = fabs(u0[2]);
        if (ax <= ay && ax <= az) { w[0] = 0.0; w[1] = u0[2]; w[2] = -u0[1]; }
        else if (ay <= az) { w[0] = -u0[2]; w[1] = 0.0; w[2] = u0[0]; }
        else { w[0] = u0[1]; w[1] = -u0[0]; w[2] = 0.0; }
        ww = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    }
    const double r1 = rsq_f64(ww);
    double u1[3];
#pragma unroll
    for (int k = 0; k < 3; k++) u1[k] = w[k] * r1;
    S[1] = u1[0] * B[1] + u1[1] * B[4] + u1[2] * B[7];
    const double u2[3] = {u0[1] * u1[2] - u0[2] * u1[1], u0[2] * u1[0] - u0[0] * u1[2],
                          u0[0] * u1[1] - u0[1] * u1[0]};
    S[2] = u2[0] * B[2] + u2[1] * B[5] + u2[2] * B[8];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        U[3 * k] = u0[k];
        U[3 * k + 1] = u1[k];
        U[3 * k + 2] = u2[k];
    }
    S2[0] = bb0;
    S2[1] = S[1] * S[1];
    S2[2] = S[2] * S[2];
}

__device__ __forceinline__ void svd3(const double F[9], double U[9], double S[3], double S2[3], double V[9]) {
    double B[9];
#pragma unroll
    for (int k = 0; k < 9; k++) B[k] = F[k];
    svd3_inplace(B, U, S, S2, V);
}

// ---------------------------------------------------------------- plasticity
// Per-particle state of the return mapping (Point3D, point_3d.h:14-27).
// Damage D of a point: the IEEE single in bits 32-63 of its utility word,
// read into [0, 1] (NaN as 0).
__device__ __forceinline__ double damage_of(long long util) {
    const float d = __uint_as_float((unsigned)((unsigned long long)util >> 32));
    return d >= 1.0f ? 1.0 : (d > 0.0f ? (double)d : 0.0);
}

// util with D (rounded to single) in bits 32-63.
__device__ __forceinline__ long long with_damage(long long util, double D) {
    const unsigned bits = __float_as_uint((float)D);
    return (long long)(((unsigned long long)util & 0xffffffffull) | ((unsigned long long)bits << 32));
}

struct PState {
    double Fe[9];
    double Jp_inv;
    double p_tr, q_tr, Je_tr;
    double rc_Jsq; // J^{-2/3} (rcbrt(Je_tr^2)), from compute_pq
    long long util;
};

// ComputePQ (mpm_kernels.cuh:477-484)
__device__ __forceinline__ void compute_pq(const KParams &kp, PState &p) {
    const double *F = p.Fe;
    p.Je_tr = det3(F);
    double Jsq = p.Je_tr * p.Je_tr;
    p.p_tr = -(0.5 * kp.kappa) * (Jsq - 1.);
    double b[9];
    left_cauchy_green(F, b);
    double tr3 = (b[0] + b[4] + b[8]) / 3;
    b[0] -= tr3;
    b[4] -= tr3;
    b[8] -= tr3;
    double sq = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) sq += b[k] * b[k];
    p.rc_Jsq = rcbrt(Jsq);
    p.q_tr = kCoeff1 * kp.mu * p.rc_Jsq * sqrt(sq);
}

// GetParametersForGrain + CheckIfPointIsInsideFailureSurface (mpm_kernels.cuh:434-474)
__device__ __forceinline__ void check_failure_surface(const KParams &kp, PState &p) {
    short grain = (short)p.util;
    double var2 = 1.0 + kp.grain_var * 0.033 * (-15 + (grain + 3) % 30);
    double var3 = 1.0 + kp.grain_var * 0.1 * (-10 + (grain + 4) % 11);
    double pmax = kp.ice_comp;
    double pmin = -kp.ice_tens;
    double qmax = kp.ice_shear * var3;
    double pmin2 = -kp.ice_tens2 * var2;
    double beta = kp.nacc_beta;
    double M_sq = (4 * qmax * qmax * (1 + 2 * beta)) / ((pmax - pmin) * (pmax - pmin));
    if (p.p_tr < 0) {
        if (p.p_tr < pmin2) {
            p.util |= ICEMPM_STATUS_CRUSHED;
            return;
        }
        double q0 = 2 * sqrt(-pmax * pmin) * qmax / (pmax - pmin);
        double k = -q0 / pmin2;
        double q_limit = k * (p.p_tr - pmin2);
        if (p.q_tr > q_limit) p.util |= ICEMPM_STATUS_CRUSHED;
    } else {
        double y = (1. + 2. * beta) * p.q_tr * p.q_tr + M_sq * (p.p_tr + beta * pmax) * (p.p_tr - pmax);
        if (y > 0) p.util |= ICEMPM_STATUS_CRUSHED;
    }
}

__device__ __forceinline__ void usvt(const double U[9], const double s[3], const double V[9], double out[9]) {
    double US[9];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) US[3 * r + c] = U[3 * r + c] * s[c];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            out[3 * r + c] = US[3 * r] * V[3 * c] + US[3 * r + 1] * V[3 * c + 1] + US[3 * r + 2] * V[3 * c + 2];
}

// ComputeSVD + Wolper_Drucker_Prager (mpm_kernels.cuh:486-544) on a
// crushed point after compute_pq. J^{2/3} is taken as J^2 * (J^{-2/3})^2 and
// the tension-branch constants come precomputed; the rest keeps the
// reference's expression order.
// kDamage (the kernels instantiate it where kp.dmg_on is set): the cohesion
// of the tension cutoff and of the Drucker-Prager line is scaled by
// f = 1 - dmg_soft * D, and D grows by the equivalent plastic shear strain
// over the softening strain, stored as soon as the branch has fixed the new
// shear stress. Without it the code is the reference's return mapping alone.
// Returns whether F changed (false: inside the yield surface).
template <bool kDamage = false>
__device__ __forceinline__ bool return_map(const KParams &kp, PState &p) {
    constexpr bool damage = kDamage;
    const double *rmc = kp.rmc;
    const double pt0 = rmc[kRmcThresholdP];
    double pt = pt0;
    if constexpr (damage) pt *= 1.0 - rmc[kRmcDmgSoft] * damage_of(p.util);
    // The branch and the new shear stress depend on p, q and Jp_inv alone, so
    // they are settled before the SVD, which then runs only for a point whose
    // F changes and may consume F in place.
    const bool tension = p.p_tr < -pt || p.Jp_inv < 1;
    double q_n_1 = 0.0, s_tension = 0.0;
    if (tension) {
        if (!damage || pt == pt0) {
            s_tension = rmc[kRmcCbrtJeNew];
            p.Jp_inv *= rmc[kRmcJeNew] / p.Je_tr;
        } else {
            const double je_new = sqrt(-2. * (-pt) / kp.kappa + 1.);
            s_tension = cbrt(je_new);
            p.Jp_inv *= je_new / p.Je_tr;
        }
        // the deviatoric stress drops from q_tr to 0
        if constexpr (damage)
            p.util = with_damage(p.util,
                                 fmin(1.0, damage_of(p.util) + p.q_tr / (3.0 * kp.mu) * rmc[kRmcDmgRate]));
    } else {
        const double pmax = kp.ice_comp;
        if (p.p_tr > pmax) {
            q_n_1 = 0;
        } else {
            const double q_from_dp = (p.p_tr + pt) * rmc[kRmcTanPhi];
            const double pmin = -kp.ice_tens;
            const double q_from_failure_surface =
                2 * sqrt((pmax - p.p_tr) * (p.p_tr - pmin)) * kp.ice_shear / (pmax - pmin);
            q_n_1 = fmin(q_from_failure_surface, q_from_dp);
        }
        if (!(p.q_tr >= q_n_1)) return false; // inside the yield surface: F stays as it is
        // the deviatoric stress drops from q_tr to q_n_1
        if constexpr (damage)
            p.util = with_damage(
                p.util, fmin(1.0, damage_of(p.util) + (p.q_tr - q_n_1) / (3.0 * kp.mu) * rmc[kRmcDmgRate]));
    }
    double U[9], V[9], sig[3], sig2[3];
    svd3_inplace(p.Fe, U, sig, sig2, V); // p.Fe now holds F V, overwritten below
    double snew[3];
    if (tension) {
        snew[0] = snew[1] = snew[2] = s_tension;
    } else {
        const double Je_tr_sq = p.Je_tr * p.Je_tr;
        const double s = kp.mu * p.rc_Jsq;
        const double mean = (sig2[0] + sig2[1] + sig2[2]) / 3;
        double nv[3] = {s * (sig2[0] - mean), s * (sig2[1] - mean), s * (sig2[2] - mean)};
        // Eigen normalized(): unchanged when the norm is zero
        const double n2 = nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2];
        if (n2 > 0) {
            const double rn = rsq_f64(n2);
#pragma unroll
            for (int k = 0; k < 3; k++) nv[k] *= rn;
        }
        const double coef = q_n_1 * kCoeff1Inv * (Je_tr_sq * p.rc_Jsq * p.rc_Jsq) / kp.mu;
#pragma unroll
        for (int k = 0; k < 3; k++) snew[k] = sqrt_f64(coef * nv[k] + mean);
    }
    usvt(U, snew, V, p.Fe);
    return true;
}

} // namespace icempm
