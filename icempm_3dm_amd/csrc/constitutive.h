// constitutive.h — device-side stress, crush test and Drucker-Prager return
// mapping of the reference (mpm_kernels.cuh:434-581), plus an FP64 3x3 SVD.
// Matrices are row-major double[9]. Expression order follows the reference
// so that the device agrees with a host restatement to rounding level.
#pragma once

#include "device_common.h"

namespace icempm {

constexpr double kCoeff1 = 1.224744871391589;    // sqrt(3/2), mpm_kernels.cuh:11
constexpr double kCoeff1Inv = 0.8164965809277260; // sqrt(2/3), mpm_kernels.cuh:12

// KirchhoffStress_Wolper (mpm_kernels.cuh:547-557):
// tau = mu J^{-2/3} dev(F F^T) + (kappa/2)(J^2 - 1) I
__device__ __forceinline__ void kirchhoff_stress(const KParams &kp, const double F[9], double P[9]) {
    double Je = det3(F);
    double b[9];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            b[3 * r + c] = F[3 * r] * F[3 * c] + F[3 * r + 1] * F[3 * c + 1] + F[3 * r + 2] * F[3 * c + 2];
    double tr3 = (b[0] + b[4] + b[8]) / 3;
    b[0] -= tr3;
    b[4] -= tr3;
    b[8] -= tr3;
    double s = kp.mu * rcbrt(Je * Je);
    double pterm = kp.kappa * 0.5 * (Je * Je - 1.);
#pragma unroll
    for (int k = 0; k < 9; k++) P[k] = s * b[k];
    P[0] += pterm;
    P[4] += pterm;
    P[8] += pterm;
}

// Water (mpm_kernels.cuh:559-566): kappa (1 - J^-3) I
__device__ __forceinline__ void water_stress(const KParams &kp, double J, double P[9]) {
    double v = kp.kappa * (1. - pow(J, -3.0));
#pragma unroll
    for (int k = 0; k < 9; k++) P[k] = 0.0;
    P[0] = P[4] = P[8] = v;
}

// ---------------------------------------------------------------- SVD
// F = U diag(S) V^T with U, V rotations, S[0] >= S[1] >= |S[2]|, S[2] < 0
// iff det F < 0. V from cyclic Jacobi on F^T F (FP64, converged to machine
// precision), U and S from a Givens QR of F V. This replaces the reference's
// FP32 McAdams routine (helper_math.cuh:23-1119) with an FP64 one.
__device__ __forceinline__ void jacobi_rotate(double &app, double &aqq, double &apq, double &arp, double &arq,
                                              double V[9], int p, int q) {
    if (apq == 0.0) return;
    double tau = (aqq - app) / (2.0 * apq);
    double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
    double c = 1.0 / sqrt(1.0 + t * t);
    double s = t * c;
    app -= t * apq;
    aqq += t * apq;
    apq = 0.0;
    double rp = arp, rq = arq;
    arp = c * rp - s * rq;
    arq = s * rp + c * rq;
#pragma unroll
    for (int r = 0; r < 3; r++) {
        double vp = V[3 * r + p], vq = V[3 * r + q];
        V[3 * r + p] = c * vp - s * vq;
        V[3 * r + q] = s * vp + c * vq;
    }
}

__device__ __forceinline__ void swap_cols(double M[9], int a, int b) {
#pragma unroll
    for (int r = 0; r < 3; r++) {
        double t = M[3 * r + a];
        M[3 * r + a] = M[3 * r + b];
        M[3 * r + b] = t;
    }
}

__device__ __forceinline__ void givens_rows(double B[9], double U[9], int p, int q, int col) {
    // rotate rows p,q of B so that B(q,col) = 0; U accumulates the transpose.
    double a = B[3 * p + col], b = B[3 * q + col];
    double r = sqrt(a * a + b * b);
    double c = 1.0, s = 0.0;
    if (r > 0.0) {
        c = a / r;
        s = b / r;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double bp = B[3 * p + k], bq = B[3 * q + k];
        B[3 * p + k] = c * bp + s * bq;
        B[3 * q + k] = -s * bp + c * bq;
        double up = U[3 * k + p], uq = U[3 * k + q];
        U[3 * k + p] = c * up + s * uq;
        U[3 * k + q] = -s * up + c * uq;
    }
}

__device__ __forceinline__ void svd3(const double F[9], double U[9], double S[3], double V[9]) {
    // A = F^T F
    double a00 = F[0] * F[0] + F[3] * F[3] + F[6] * F[6];
    double a11 = F[1] * F[1] + F[4] * F[4] + F[7] * F[7];
    double a22 = F[2] * F[2] + F[5] * F[5] + F[8] * F[8];
    double a01 = F[0] * F[1] + F[3] * F[4] + F[6] * F[7];
    double a02 = F[0] * F[2] + F[3] * F[5] + F[6] * F[8];
    double a12 = F[1] * F[2] + F[4] * F[5] + F[7] * F[8];
#pragma unroll
    for (int k = 0; k < 9; k++) V[k] = (k % 4 == 0) ? 1.0 : 0.0;
#pragma unroll 1
    for (int sweep = 0; sweep < 12; sweep++) {
        double off = a01 * a01 + a02 * a02 + a12 * a12;
        double dia = a00 * a00 + a11 * a11 + a22 * a22;
        if (!(off > 1e-34 * dia)) break; // ~ (1e-17)^2 relative: converged
        jacobi_rotate(a00, a11, a01, a02, a12, V, 0, 1);
        jacobi_rotate(a00, a22, a02, a01, a12, V, 0, 2);
        jacobi_rotate(a11, a22, a12, a01, a02, V, 1, 2);
    }
    // sort eigenvalues descending; each swap negates one column to keep det V = +1
    if (a00 < a11) {
        double t = a00; a00 = a11; a11 = t;
        swap_cols(V, 0, 1);
#pragma unroll
        for (int r = 0; r < 3; r++) V[3 * r + 1] = -V[3 * r + 1];
    }
    if (a00 < a22) {
        double t = a00; a00 = a22; a22 = t;
        swap_cols(V, 0, 2);
#pragma unroll
        for (int r = 0; r < 3; r++) V[3 * r + 2] = -V[3 * r + 2];
    }
    if (a11 < a22) {
        double t = a11; a11 = a22; a22 = t;
        swap_cols(V, 1, 2);
#pragma unroll
        for (int r = 0; r < 3; r++) V[3 * r + 2] = -V[3 * r + 2];
    }
    // B = F V, then QR: B = U R with U a product of rotations
    double B[9];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) B[3 * r + c] = F[3 * r] * V[c] + F[3 * r + 1] * V[3 + c] + F[3 * r + 2] * V[6 + c];
#pragma unroll
    for (int k = 0; k < 9; k++) U[k] = (k % 4 == 0) ? 1.0 : 0.0;
    givens_rows(B, U, 0, 1, 0);
    givens_rows(B, U, 0, 2, 0);
    givens_rows(B, U, 1, 2, 1);
    S[0] = B[0];
    S[1] = B[4];
    S[2] = B[8];
}

// ---------------------------------------------------------------- plasticity
// Per-particle state of the return mapping (Point3D, point_3d.h:14-27).
struct PState {
    double Fe[9];
    double Jp_inv;
    double p_tr, q_tr, Je_tr;
    long long util;
};

// ComputePQ (mpm_kernels.cuh:477-484)
__device__ __forceinline__ void compute_pq(const KParams &kp, PState &p) {
    const double *F = p.Fe;
    p.Je_tr = det3(F);
    double Jsq = p.Je_tr * p.Je_tr;
    p.p_tr = -(0.5 * kp.kappa) * (Jsq - 1.);
    double b[9];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            b[3 * r + c] = F[3 * r] * F[3 * c] + F[3 * r + 1] * F[3 * c + 1] + F[3 * r + 2] * F[3 * c + 2];
    double tr3 = (b[0] + b[4] + b[8]) / 3;
    b[0] -= tr3;
    b[4] -= tr3;
    b[8] -= tr3;
    double sq = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) sq += b[k] * b[k];
    p.q_tr = kCoeff1 * kp.mu * rcbrt(Jsq) * sqrt(sq);
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

// ComputeSVD + Wolper_Drucker_Prager (mpm_kernels.cuh:486-544), run from
// a separate kernel: only crushed particles take it.
__device__ __forceinline__ void return_map(const KParams &kp, PState &p) {
    double U[9], V[9], sig[3];
    svd3(p.Fe, U, sig, V);
    double sig2[3] = {sig[0] * sig[0], sig[1] * sig[1], sig[2] * sig[2]};
    const double Je_tr_sq = p.Je_tr * p.Je_tr;
    const double mu = kp.mu;
    double s = mu * rcbrt(Je_tr_sq);
    double mean = (sig2[0] + sig2[1] + sig2[2]) / 3;
    double shat[3] = {s * (sig2[0] - mean), s * (sig2[1] - mean), s * (sig2[2] - mean)};

    const double kappa = kp.kappa;
    const double tan_phi = kp.tan_phi;
    const double DP_threshold_p = kp.dp_threshold_p;
    const double pmax = kp.ice_comp;
    const double qmax = kp.ice_shear;

    if (p.p_tr < -DP_threshold_p || p.Jp_inv < 1) {
        double p_new = -DP_threshold_p;
        double Je_new = sqrt(-2. * p_new / kappa + 1.);
        double cbrt_Je_new = cbrt(Je_new);
        double snew[3] = {cbrt_Je_new, cbrt_Je_new, cbrt_Je_new};
        usvt(U, snew, V, p.Fe);
        p.Jp_inv *= Je_new / p.Je_tr;
    } else {
        double q_n_1;
        if (p.p_tr > pmax) {
            q_n_1 = 0;
        } else {
            double q_from_dp = (p.p_tr + DP_threshold_p) * tan_phi;
            const double pmin = -kp.ice_tens;
            double q_from_failure_surface = 2 * sqrt((pmax - p.p_tr) * (p.p_tr - pmin)) * qmax / (pmax - pmin);
            q_n_1 = fmin(q_from_failure_surface, q_from_dp);
        }
        if (p.q_tr >= q_n_1) {
            double s_hat_n_1_norm = q_n_1 * kCoeff1Inv;
            double n2 = shat[0] * shat[0] + shat[1] * shat[1] + shat[2] * shat[2];
            double nv[3] = {shat[0], shat[1], shat[2]};
            if (n2 > 0) {
                double nn = sqrt(n2);
#pragma unroll
                for (int k = 0; k < 3; k++) nv[k] = nv[k] / nn;
            }
            double coef = s_hat_n_1_norm * cbrt(Je_tr_sq) / mu;
            double snew[3];
#pragma unroll
            for (int k = 0; k < 3; k++) snew[k] = sqrt(coef * nv[k] + mean);
            usvt(U, snew, V, p.Fe);
        }
    }
}

} // namespace icempm
