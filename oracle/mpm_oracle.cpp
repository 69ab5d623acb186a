// mpm_oracle.cpp — CPU restatement of the reference MPM step (see mpm_oracle.h).
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the host-core
// baseline. Each function names the reference lines it restates. Arithmetic
// follows the reference's expression order so that the only differences to
// the device are rounding-level (FMA contraction, libm cbrt/sqrt ulps,
// atomic summation order).

#include "mpm_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// mpm_kernels.cuh:11-12
constexpr double kCoeff1 = 1.224744871391589;    // sqrt(3/2)
constexpr double kCoeff1Inv = 0.8164965809277260; // sqrt(2/3)

inline long long util_bits(double d) {
    long long v;
    std::memcpy(&v, &d, sizeof(v));
    return v;
}
inline double bits_util(long long v) {
    double d;
    std::memcpy(&d, &v, sizeof(d));
    return d;
}

struct Mat3 {
    double a[9];
    double &operator()(int i, int j) { return a[3 * i + j]; }
    double operator()(int i, int j) const { return a[3 * i + j]; }
};

inline Mat3 zero3() {
    Mat3 m;
    for (double &x : m.a) x = 0;
    return m;
}
inline Mat3 eye3() {
    Mat3 m = zero3();
    m(0, 0) = m(1, 1) = m(2, 2) = 1.0;
    return m;
}
// Eigen fixed-size product: coefficient (r,c) = sum_k A(r,k)*B(k,c), k ascending.
inline Mat3 mul(const Mat3 &A, const Mat3 &B) {
    Mat3 C;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) C(r, c) = A(r, 0) * B(0, c) + A(r, 1) * B(1, c) + A(r, 2) * B(2, c);
    return C;
}
inline Mat3 mul_abt(const Mat3 &A, const Mat3 &B) { // A * B^T
    Mat3 C;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) C(r, c) = A(r, 0) * B(c, 0) + A(r, 1) * B(c, 1) + A(r, 2) * B(c, 2);
    return C;
}
// Eigen 3.4 Determinant.h, size 3: cofactor expansion along row 0.
inline double det3(const Mat3 &m) {
    return m(0, 0) * (m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1)) -
           m(0, 1) * (m(1, 0) * m(2, 2) - m(1, 2) * m(2, 0)) +
           m(0, 2) * (m(1, 0) * m(2, 1) - m(1, 1) * m(2, 0));
}
// dev (mpm_kernels.cuh:569-572)
inline Mat3 dev(const Mat3 &A) {
    Mat3 D = A;
    double t = (A(0, 0) + A(1, 1) + A(2, 2)) / 3;
    D(0, 0) -= t;
    D(1, 1) -= t;
    D(2, 2) -= t;
    return D;
}
inline double rcbrt(double x) { return 1.0 / std::cbrt(x); }

// KirchhoffStress_Wolper (mpm_kernels.cuh:547-557)
Mat3 kirchhoff(double mu, double kappa, const Mat3 &F) {
    double Je = det3(F);
    Mat3 b = mul_abt(F, F);
    Mat3 db = dev(b);
    double s = mu * rcbrt(Je * Je);
    double pterm = kappa * 0.5 * (Je * Je - 1.);
    Mat3 P;
    for (int i = 0; i < 9; i++) P.a[i] = s * db.a[i];
    P(0, 0) += pterm;
    P(1, 1) += pterm;
    P(2, 2) += pterm;
    return P;
}

// Water (mpm_kernels.cuh:559-566)
Mat3 water(double kappa, double J) {
    constexpr double gamma = 3;
    Mat3 P = zero3();
    double v = kappa * (1. - std::pow(J, -gamma));
    P(0, 0) = P(1, 1) = P(2, 2) = v;
    return P;
}

// ---------------------------------------------------------------- SVD
// One-sided (Hestenes) Jacobi in FP64: rotate column pairs of A = F until
// they are orthogonal; V accumulates the rotations. This is not the
// reference's routine (an FP32 McAdams SVD, helper_math.cuh:23-1119); the
// step only uses U diag(g(S)) V^T, which any accurate SVD reproduces.
void svd3(const Mat3 &F, Mat3 &U, double S[3], Mat3 &V) {
    double A[3][3]; // columns: A[c][r]
    double Vc[3][3];
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) {
            A[c][r] = F(r, c);
            Vc[c][r] = (r == c) ? 1.0 : 0.0;
        }
    const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int sweep = 0; sweep < 60; sweep++) {
        bool rotated = false;
        for (auto &pq : pairs) {
            int p = pq[0], q = pq[1];
            double alpha = 0, beta = 0, gamma = 0;
            for (int r = 0; r < 3; r++) {
                alpha += A[p][r] * A[p][r];
                beta += A[q][r] * A[q][r];
                gamma += A[p][r] * A[q][r];
            }
            if (gamma == 0.0 || std::fabs(gamma) <= DBL_EPSILON * std::sqrt(alpha * beta)) continue;
            rotated = true;
            double zeta = (beta - alpha) / (2.0 * gamma);
            double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
            double c = 1.0 / std::sqrt(1.0 + t * t);
            double s = c * t;
            for (int r = 0; r < 3; r++) {
                double ap = A[p][r], aq = A[q][r];
                A[p][r] = c * ap - s * aq;
                A[q][r] = s * ap + c * aq;
                double vp = Vc[p][r], vq = Vc[q][r];
                Vc[p][r] = c * vp - s * vq;
                Vc[q][r] = s * vp + c * vq;
            }
        }
        if (!rotated) break;
    }
    double sig[3];
    for (int c = 0; c < 3; c++)
        sig[c] = std::sqrt(A[c][0] * A[c][0] + A[c][1] * A[c][1] + A[c][2] * A[c][2]);
    int order[3] = {0, 1, 2};
    std::sort(order, order + 3, [&](int i, int j) { return sig[i] > sig[j]; });
    double Us[3][3], Vs[3][3], ss[3];
    for (int k = 0; k < 3; k++) {
        int c = order[k];
        ss[k] = sig[c];
        for (int r = 0; r < 3; r++) {
            Vs[k][r] = Vc[c][r];
            Us[k][r] = ss[k] > 0 ? A[c][r] / ss[k] : 0.0;
        }
    }
    auto cross = [](const double *a, const double *b, double *o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    auto dot = [](const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    // V must be a rotation: fix the sign of its third column (the column of
    // A = F V follows).
    double vx[3];
    cross(Vs[0], Vs[1], vx);
    double A2[3];
    for (int r = 0; r < 3; r++) A2[r] = ss[2] * Us[2][r];
    if (dot(vx, Vs[2]) < 0)
        for (int r = 0; r < 3; r++) {
            Vs[2][r] = -Vs[2][r];
            A2[r] = -A2[r];
        }
    // U by Gram-Schmidt on the columns of A, completed where a singular value
    // is negligible; U2 = U0 x U1 makes U a rotation and the signed third
    // singular value is A2 . U2 (negative iff det F < 0).
    const double tiny = 8 * DBL_EPSILON * ss[0];
    double u0[3] = {1, 0, 0}, u1[3], u2[3];
    if (ss[0] > 0) std::memcpy(u0, Us[0], sizeof(u0));
    bool have_u1 = false;
    if (ss[1] > tiny && ss[1] > 0) {
        double a1[3], d = 0;
        for (int r = 0; r < 3; r++) a1[r] = ss[1] * Us[1][r];
        d = dot(u0, a1);
        for (int r = 0; r < 3; r++) u1[r] = a1[r] - d * u0[r];
        double nn = std::sqrt(dot(u1, u1));
        if (nn > tiny) {
            for (double &x : u1) x /= nn;
            have_u1 = true;
        }
    }
    if (!have_u1) {
        double e[3] = {0, 0, 0};
        e[std::fabs(u0[0]) < 0.9 ? 0 : 1] = 1.0;
        cross(u0, e, u1);
        double nn = std::sqrt(dot(u1, u1));
        for (double &x : u1) x /= nn;
    }
    cross(u0, u1, u2);
    ss[2] = dot(A2, u2);
    std::memcpy(Us[0], u0, sizeof(u0));
    std::memcpy(Us[1], u1, sizeof(u1));
    std::memcpy(Us[2], u2, sizeof(u2));
    for (int k = 0; k < 3; k++) {
        S[k] = ss[k];
        for (int r = 0; r < 3; r++) {
            U(r, k) = Us[k][r];
            V(r, k) = Vs[k][r];
        }
    }
}

// Point3D (point_3d.h:14-27), the per-particle state inside G2P.
struct Point {
    double pos[3], vel[3];
    Mat3 Bp, Fe;
    double Jp_inv;
    short grain;
    double p_tr, q_tr, Je_tr;
    Mat3 U, V;
    double vSigma[3], vSigmaSquared[3], v_s_hat_tr[3];
    long long utility_data;
};

// GetParametersForGrain (mpm_kernels.cuh:434-450)
void grain_params(const icempm_SimParams3D &P, short grain, double &pmin, double &pmax, double &qmax,
                  double &beta, double &mSq, double &pmin2) {
    double var2 = 1.0 + P.GrainVariability * 0.033 * (-15 + (grain + 3) % 30);
    double var3 = 1.0 + P.GrainVariability * 0.1 * (-10 + (grain + 4) % 11);
    pmax = P.IceCompressiveStrength;
    pmin = -P.IceTensileStrength;
    qmax = P.IceShearStrength * var3;
    pmin2 = -P.IceTensileStrength2 * var2;
    beta = P.NACC_beta;
    mSq = (4 * qmax * qmax * (1 + 2 * beta)) / ((pmax - pmin) * (pmax - pmin));
}

// CheckIfPointIsInsideFailureSurface (mpm_kernels.cuh:453-474)
void check_failure_surface(const icempm_SimParams3D &P, Point &p) {
    double beta, M_sq, pmin, pmax, qmax, pmin2;
    grain_params(P, p.grain, pmin, pmax, qmax, beta, M_sq, pmin2);
    if (p.p_tr < 0) {
        if (p.p_tr < pmin2) {
            p.utility_data |= ICEMPM_STATUS_CRUSHED;
            return;
        }
        double q0 = 2 * std::sqrt(-pmax * pmin) * qmax / (pmax - pmin);
        double k = -q0 / pmin2;
        double q_limit = k * (p.p_tr - pmin2);
        if (p.q_tr > q_limit) {
            p.utility_data |= ICEMPM_STATUS_CRUSHED;
            return;
        }
    } else {
        double y = (1. + 2. * beta) * p.q_tr * p.q_tr + M_sq * (p.p_tr + beta * pmax) * (p.p_tr - pmax);
        if (y > 0) p.utility_data |= ICEMPM_STATUS_CRUSHED;
    }
}

// ComputePQ (mpm_kernels.cuh:477-484)
void compute_pq(Point &p, double kappa, double mu) {
    const Mat3 &F = p.Fe;
    p.Je_tr = det3(F);
    double Jsq = p.Je_tr * p.Je_tr;
    p.p_tr = -(0.5 * kappa) * (Jsq - 1.);
    Mat3 d = dev(mul_abt(F, F));
    double sq = 0;
    for (double x : d.a) sq += x * x;
    p.q_tr = kCoeff1 * mu * rcbrt(Jsq) * std::sqrt(sq);
}

// ComputeSVD (mpm_kernels.cuh:486-492)
void compute_svd(Point &p, double kappa, double mu) {
    (void)kappa;
    svd3(p.Fe, p.U, p.vSigma, p.V);
    for (int i = 0; i < 3; i++) p.vSigmaSquared[i] = p.vSigma[i] * p.vSigma[i];
    const double Je_tr_sq = p.Je_tr * p.Je_tr;
    double s = mu * rcbrt(Je_tr_sq);
    double mean = (p.vSigmaSquared[0] + p.vSigmaSquared[1] + p.vSigmaSquared[2]) / 3;
    for (int i = 0; i < 3; i++) p.v_s_hat_tr[i] = s * (p.vSigmaSquared[i] - mean);
}

// U * diag(s) * V^T, evaluated as (U*diag(s)) * V^T.
Mat3 usvt(const Mat3 &U, const double s[3], const Mat3 &V) {
    Mat3 US;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) US(r, c) = U(r, c) * s[c];
    return mul_abt(US, V);
}

// Wolper_Drucker_Prager (mpm_kernels.cuh:494-544)
void wolper_drucker_prager(const icempm_SimParams3D &P, Point &p) {
    const double mu = P.mu;
    const double kappa = P.kappa;
    const double tan_phi = P.DP_tan_phi;
    const double DP_threshold_p = P.DP_threshold_p;
    const double pmax = P.IceCompressiveStrength;
    const double qmax = P.IceShearStrength;

    if (p.p_tr < -DP_threshold_p || p.Jp_inv < 1) {
        double p_new = -DP_threshold_p;
        double Je_new = std::sqrt(-2. * p_new / kappa + 1.);
        double cbrt_Je_new = std::cbrt(Je_new);
        double s[3] = {cbrt_Je_new, cbrt_Je_new, cbrt_Je_new};
        p.Fe = usvt(p.U, s, p.V);
        p.Jp_inv *= Je_new / p.Je_tr;
    } else {
        double q_n_1;
        if (p.p_tr > pmax) {
            q_n_1 = 0;
        } else {
            double q_from_dp = (p.p_tr + DP_threshold_p) * tan_phi;
            const double pmin = -P.IceTensileStrength;
            double q_from_failure_surface =
                2 * std::sqrt((pmax - p.p_tr) * (p.p_tr - pmin)) * qmax / (pmax - pmin);
            // device min(double,double) is fmin: a NaN operand yields the other one
            q_n_1 = std::fmin(q_from_failure_surface, q_from_dp);
        }
        if (p.q_tr >= q_n_1) {
            double s_hat_n_1_norm = q_n_1 * kCoeff1Inv;
            double Je_tr_sq = p.Je_tr * p.Je_tr;
            // Eigen normalized(): v / sqrt(|v|^2) if |v|^2 > 0, else v unchanged.
            double n2 = p.v_s_hat_tr[0] * p.v_s_hat_tr[0] + p.v_s_hat_tr[1] * p.v_s_hat_tr[1] +
                        p.v_s_hat_tr[2] * p.v_s_hat_tr[2];
            double nv[3] = {p.v_s_hat_tr[0], p.v_s_hat_tr[1], p.v_s_hat_tr[2]};
            if (n2 > 0) {
                double nn = std::sqrt(n2);
                for (double &x : nv) x = x / nn;
            }
            double coef = s_hat_n_1_norm * std::cbrt(Je_tr_sq) / mu;
            double mean = (p.vSigmaSquared[0] + p.vSigmaSquared[1] + p.vSigmaSquared[2]) / 3;
            double snew[3];
            for (int i = 0; i < 3; i++) snew[i] = std::sqrt(coef * nv[i] + mean);
            p.Fe = usvt(p.U, snew, p.V);
        }
    }
}

struct Weights {
    int base[3];
    double fx[3];
    double w[3][3]; // w[i][axis]
};

// mpm_kernels.cuh:65-81 (and :278-286). Returns 0 or an error code; on a
// non-zero code the particle must be skipped.
int weights_and_check(const icempm_SimParams3D &P, const double pos[3], Weights &W) {
    const double h_inv = P.cellsize_inv;
    double xi[3];
    for (int d = 0; d < 3; d++) {
        xi[d] = pos[d] * h_inv;
        if (!std::isfinite(xi[d]) || std::fabs(xi[d]) > 1e9) return ICEMPM_DEVERR_NONFINITE;
    }
    for (int d = 0; d < 3; d++) {
        W.base[d] = (int)(xi[d] - 0.5); // truncation, as the reference's cast<int>()
        W.fx[d] = xi[d] - (double)W.base[d];
    }
    int code = 0;
    const int gdim[3] = {P.GridXTotal, P.GridY, P.GridZ};
    for (int d = 0; d < 3; d++) {
        if (W.base[d] < 0) code = std::max(code, 70 + d);
        if (W.base[d] > gdim[d] - 3) code = std::max(code, 73 + d);
    }
    if (code) return code;
    for (int d = 0; d < 3; d++) {
        double v0 = 1.5 - W.fx[d], v1 = W.fx[d] - 1.0, v2 = W.fx[d] - 0.5;
        W.w[0][d] = 0.5 * v0 * v0;
        W.w[1][d] = 0.75 - v1 * v1;
        W.w[2][d] = 0.5 * v2 * v2;
    }
    return 0;
}

inline size_t node_index(const icempm_SimParams3D &P, int x, int y, int z) {
    return ((size_t)x * P.GridY + y) * P.GridZ + z;
}

// Body of partition_kernel_p2g for one particle (mpm_kernels.cuh:24-99).
int p2g_particle(const icempm_SimParams3D &P, const double *soa, int pitch, int i, double *grid,
                 size_t nn) {
    long long util = util_bits(soa[i]);
    if (util & ICEMPM_STATUS_DISABLED) return 0;
    const double h = P.cellsize;
    const int isLiquid = (int)((util & ICEMPM_STATUS_LIQUID) >> 18);
    const double masses[2] = {P.ParticleMass, P.LiquidParticleMass};
    const double m = masses[isLiquid];

    double pos[3], vel[3];
    Mat3 Fe, Bp;
    for (int a = 0; a < 3; a++) {
        pos[a] = soa[i + (size_t)pitch * (ICEMPM_POSX + a)];
        vel[a] = soa[i + (size_t)pitch * (ICEMPM_VELX + a)];
        for (int b = 0; b < 3; b++) {
            Fe(a, b) = soa[i + (size_t)pitch * (ICEMPM_FE00 + 3 * a + b)];
            Bp(a, b) = soa[i + (size_t)pitch * (ICEMPM_BP00 + 3 * a + b)];
        }
    }
    Mat3 PFt = isLiquid ? water(P.kappa, soa[i + (size_t)pitch * ICEMPM_IDX_JP_INV])
                        : kirchhoff(P.mu, P.kappa, Fe);
    Mat3 sub;
    for (int k = 0; k < 9; k++) sub.a[k] = m * Bp.a[k] - P.dt_vol_Dpinv * PFt.a[k];

    Weights W;
    int code = weights_and_check(P, pos, W);
    if (code) return code;

    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            for (int c = 0; c < 3; c++) {
                double Wip = W.w[a][0] * W.w[b][1] * W.w[c][2];
                double dpos[3] = {(a - W.fx[0]) * h, (b - W.fx[1]) * h, (c - W.fx[2]) * h};
                double incM = Wip * m;
                size_t idx = node_index(P, W.base[0] + a, W.base[1] + b, W.base[2] + c);
                grid[idx] += incM;
                for (int r = 0; r < 3; r++) {
                    double sd = sub(r, 0) * dpos[0] + sub(r, 1) * dpos[1] + sub(r, 2) * dpos[2];
                    grid[(1 + r) * nn + idx] += Wip * (vel[r] * m + sd);
                }
            }
    return 0;
}

} // namespace

extern "C" {

void oracle_weights(const icempm_SimParams3D *prms, const double pos[3], int base[3], double fx[3],
                    double w[9]) {
    Weights W;
    icempm_SimParams3D P = *prms;
    // no bounds: callers probe the weights themselves
    P.GridXTotal = P.GridY = P.GridZ = 1 << 29;
    weights_and_check(P, pos, W);
    for (int d = 0; d < 3; d++) {
        base[d] = W.base[d];
        fx[d] = W.fx[d];
        for (int i = 0; i < 3; i++) w[3 * i + d] = W.w[i][d];
    }
}

void oracle_kirchhoff(double mu, double kappa, const double F[9], double tau[9]) {
    Mat3 f;
    std::memcpy(f.a, F, sizeof(f.a));
    Mat3 t = kirchhoff(mu, kappa, f);
    std::memcpy(tau, t.a, sizeof(t.a));
}

void oracle_svd3(const double F[9], double U[9], double S[3], double V[9]) {
    Mat3 f, u, v;
    std::memcpy(f.a, F, sizeof(f.a));
    svd3(f, u, S, v);
    std::memcpy(U, u.a, sizeof(u.a));
    std::memcpy(V, v.a, sizeof(v.a));
}

double oracle_det3(const double F[9]) {
    Mat3 f;
    std::memcpy(f.a, F, sizeof(f.a));
    return det3(f);
}

// ProxyPoint3D::getCellIndex (proxypoint3d.cpp:158-163)
void oracle_cell_keys_reference(const icempm_SimParams3D *P, const double *soa, int pitch, int n,
                                uint32_t *keys) {
    const double hinv = P->cellsize_inv;
    for (int i = 0; i < n; i++) {
        if (util_bits(soa[i]) & ICEMPM_STATUS_DISABLED) {
            keys[i] = 0xFFFFFFFFu;
            continue;
        }
        int ix = (int)(soa[i + (size_t)pitch * ICEMPM_POSX] * hinv + 0.5);
        int iy = (int)(soa[i + (size_t)pitch * (ICEMPM_POSX + 1)] * hinv + 0.5);
        int iz = (int)(soa[i + (size_t)pitch * (ICEMPM_POSX + 2)] * hinv + 0.5);
        keys[i] = (uint32_t)(ix * P->GridY * P->GridZ + iy * P->GridZ + iz);
    }
}

// Brick key: the same nearest-node cell, bricks of 4x4x4 cells in x-major
// brick order, cells x-major inside the brick. Cells are clamped into the
// grid so that every finite position has a key.
void oracle_cell_keys_brick(const icempm_SimParams3D *P, const double *soa, int pitch, int n,
                            uint32_t *keys) {
    const double hinv = P->cellsize_inv;
    const int g[3] = {P->GridXTotal, P->GridY, P->GridZ};
    const int nb[3] = {(g[0] + 3) / 4, (g[1] + 3) / 4, (g[2] + 3) / 4};
    for (int i = 0; i < n; i++) {
        if (util_bits(soa[i]) & ICEMPM_STATUS_DISABLED) {
            keys[i] = 0xFFFFFFFFu;
            continue;
        }
        int c[3];
        for (int d = 0; d < 3; d++) {
            double x = soa[i + (size_t)pitch * (ICEMPM_POSX + d)] * hinv + 0.5;
            if (!(x >= 0.0)) x = 0.0; // also NaN
            if (x > (double)(g[d] - 1)) x = (double)(g[d] - 1);
            c[d] = (int)x;
        }
        uint32_t brick = (uint32_t)(((c[0] >> 2) * nb[1] + (c[1] >> 2)) * nb[2] + (c[2] >> 2));
        uint32_t local = (uint32_t)(((c[0] & 3) * 4 + (c[1] & 3)) * 4 + (c[2] & 3));
        keys[i] = brick * 64u + local;
    }
}

void oracle_stable_sort(const uint32_t *keys, int n, int32_t *perm) {
    std::iota(perm, perm + n, 0);
    std::stable_sort(perm, perm + n, [&](int32_t a, int32_t b) { return keys[a] < keys[b]; });
}

int oracle_remove_disabled_and_sort(const icempm_SimParams3D *P, double *soa, int pitch, int n) {
    std::vector<int32_t> keep;
    keep.reserve(n);
    for (int i = 0; i < n; i++)
        if (!(util_bits(soa[i]) & ICEMPM_STATUS_DISABLED)) keep.push_back(i);
    int m = (int)keep.size();
    std::vector<double> tmp((size_t)m * ICEMPM_N_PTS_ARRAYS);
    for (int l = 0; l < ICEMPM_N_PTS_ARRAYS; l++)
        for (int j = 0; j < m; j++) tmp[(size_t)l * m + j] = soa[(size_t)l * pitch + keep[j]];
    std::vector<uint32_t> keys(m);
    oracle_cell_keys_reference(P, tmp.data(), m, m, keys.data());
    std::vector<int32_t> perm(m);
    oracle_stable_sort(keys.data(), m, perm.data());
    for (int l = 0; l < ICEMPM_N_PTS_ARRAYS; l++)
        for (int j = 0; j < m; j++) soa[(size_t)l * pitch + j] = tmp[(size_t)l * m + perm[j]];
    return m;
}

int oracle_p2g(const icempm_SimParams3D *prms, const double *soa, int pitch, int n, double *grid,
               int nthreads) {
    const icempm_SimParams3D &P = *prms;
    const size_t nn = (size_t)P.GridXTotal * P.GridY * P.GridZ;
    int code = 0;
    if (nthreads <= 1) {
        for (int i = 0; i < n; i++) code = std::max(code, p2g_particle(P, soa, pitch, i, grid, nn));
        return code;
    }
    // Parallel form: particles bucketed by x-slab of 4 base cells; slabs of
    // one parity never share a node (stencil width 3), so each color runs
    // without atomics. Deterministic for a given input.
    const int S = 4;
    const int nslab = P.GridXTotal / S + 2;
    std::vector<int> slab(n);
    for (int i = 0; i < n; i++) {
        double x = soa[i + (size_t)pitch * ICEMPM_POSX] * P.cellsize_inv - 0.5;
        int b = (std::isfinite(x) && x > -1e9 && x < 1e9) ? (int)x : 0;
        b = std::min(std::max(b, 0), P.GridXTotal);
        slab[i] = b / S;
    }
    std::vector<int> start(nslab + 1, 0);
    for (int i = 0; i < n; i++) start[slab[i] + 1]++;
    for (int s = 0; s < nslab; s++) start[s + 1] += start[s];
    std::vector<int> order(n);
    {
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int i = 0; i < n; i++) order[fill[slab[i]]++] = i;
    }
    for (int color = 0; color < 2; color++) {
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1) reduction(max : code)
        for (int s = color; s < nslab; s += 2)
            for (int k = start[s]; k < start[s + 1]; k++)
                code = std::max(code, p2g_particle(P, soa, pitch, order[k], grid, nn));
    }
    return code;
}

void oracle_update_nodes(const icempm_SimParams3D *prms, const icempm_Indenter *ind,
                         const double center[3], double *grid, double *bins, int nthreads) {
    const icempm_SimParams3D &P = *prms;
    const int gridY = P.GridY, gridZ = P.GridZ, gridXTotal = P.GridXTotal;
    const size_t nn = (size_t)gridXTotal * gridY * gridZ;
    const double indRsq = P.IndRSq;
    const double dt = P.InitialTimeStep;
    const double cellsize = P.cellsize;
    const double vmax = P.vmax, vmax_squared = P.vmax_squared;
    const double vco[3] = {ind->velocity[0], ind->velocity[1], ind->velocity[2]};
    const bool sphere = ind->shape == ICEMPM_INDENTER_SPHERE;

#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static)
    for (long long li = 0; li < (long long)nn; li++) {
        size_t idx = (size_t)li;
        double mass = grid[idx];
        if (mass == 0) continue;
        double v[3];
        for (int i = 0; i < 3; i++) v[i] = grid[(i + 1) * nn + idx];
        int idx_z = (int)(idx % gridZ);
        int idx_y = (int)((idx / gridZ) % gridY);
        int idx_x = (int)(idx / ((size_t)gridZ * gridY));
        int gi[3] = {idx_x, idx_y, idx_z};
        double gnpos[3] = {gi[0] * cellsize, gi[1] * cellsize, gi[2] * cellsize};

        for (int i = 0; i < 3; i++) v[i] /= mass;
        v[1] -= P.dt_Gravity;
        double vsq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if (vsq > vmax_squared) {
            double nrm = std::sqrt(vsq);
            for (int i = 0; i < 3; i++) v[i] = v[i] / nrm * vmax;
        }

        if (P.SetupType != 2) {
            double n[3] = {gnpos[0] - center[0], gnpos[1] - center[1], sphere ? gnpos[2] - center[2] : 0.0};
            double nsq = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
            if (nsq < indRsq) {
                double vrel[3] = {v[0] - vco[0], v[1] - vco[1], v[2] - vco[2]};
                if (nsq > 0) {
                    double nrm = std::sqrt(nsq);
                    for (int i = 0; i < 3; i++) n[i] /= nrm;
                }
                double vn = vrel[0] * n[0] + vrel[1] * n[1] + vrel[2] * n[2];
                if (vn < 0) {
                    double prev[3] = {v[0], v[1], v[2]};
                    for (int i = 0; i < 3; i++) v[i] = vco[i] + (vrel[i] - n[i] * vn);
                    double force[3];
                    for (int i = 0; i < 3; i++) force[i] = (prev[i] - v[i]) * mass / dt;
                    if (P.IndenterSubdivisions <= 0) goto bins_done; // no sensor bins
                    {
                    float angle = std::atan2((float)n[0], (float)n[1]);
                    angle += ICEMPM_PI;
                    angle *= P.IndenterSubdivisions / (2 * ICEMPM_PI);
                    int index_angle = std::min(std::max((int)angle, 0), P.IndenterSubdivisions - 1);
                    int index_z = std::min(std::max(idx_z, 0), gridZ - 1);
                    int index = index_z + index_angle * gridZ;
                    for (int i = 0; i < 3; i++) {
#pragma omp atomic
                        bins[i + 3 * index] += force[i];
                    }
                    }
                bins_done:;
                }
            }
        }

        if (P.SetupType != 2) {
            if (gi[1] <= 2) v[0] = v[1] = v[2] = 0;
        } else {
            if (gi[1] <= 2 && v[1] < 0) v[1] = 0;
        }
        if (gi[1] >= gridY - 3 && v[1] > 0) v[1] = 0;
        if (gi[0] <= 2 && v[0] < 0) v[0] = 0;
        else if (gi[0] >= gridXTotal - 3 && v[0] > 0) v[0] = 0;
        if (gi[2] <= 2 && v[2] < 0) v[2] = 0;
        else if (gi[2] >= gridZ - 3 && v[2] > 0) v[2] = 0;

        if (P.SetupType == 2 && P.DragGlobal != 0)
            for (int i = 0; i < 3; i++) v[i] *= (1 - P.DragGlobal);

        for (int i = 0; i < 3; i++) grid[(i + 1) * nn + idx] = v[i];
    }
}

int oracle_g2p(const icempm_SimParams3D *prms, double *soa, int pitch, int n, const double *grid,
               int record_pq, int nthreads) {
    const icempm_SimParams3D &P = *prms;
    const size_t nn = (size_t)P.GridXTotal * P.GridY * P.GridZ;
    const double h_inv = P.cellsize_inv;
    const double dt = P.InitialTimeStep;
    const double mu = P.mu, kappa = P.kappa;
    int code = 0;

#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static) reduction(max : code)
    for (int i = 0; i < n; i++) {
        Point p;
        double *ptr = &soa[i];
        p.utility_data = util_bits(*ptr);
        // liquid points never compute p, q; the reference records whatever its
        // uninitialised Point3D holds (mpm_kernels.cuh:250, :349-350), here 0
        p.p_tr = p.q_tr = 0.0;
        if (p.utility_data & ICEMPM_STATUS_DISABLED) continue;
        bool isLiquid = p.utility_data & ICEMPM_STATUS_LIQUID;
        for (int a = 0; a < 3; a++) {
            p.pos[a] = soa[i + (size_t)pitch * (ICEMPM_POSX + a)];
            for (int b = 0; b < 3; b++) p.Fe(a, b) = soa[i + (size_t)pitch * (ICEMPM_FE00 + 3 * a + b)];
        }
        p.Jp_inv = soa[i + (size_t)pitch * ICEMPM_IDX_JP_INV];
        p.grain = (short)p.utility_data;

        Weights W;
        int c = weights_and_check(P, p.pos, W);
        if (c) {
            code = std::max(code, c);
            continue;
        }
        for (int a = 0; a < 3; a++) p.vel[a] = 0;
        p.Bp = zero3();
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int cc = 0; cc < 3; cc++) {
                    double dpos[3] = {a - W.fx[0], b - W.fx[1], cc - W.fx[2]};
                    double weight = W.w[a][0] * W.w[b][1] * W.w[cc][2];
                    size_t idx = node_index(P, W.base[0] + a, W.base[1] + b, W.base[2] + cc);
                    double nv[3];
                    for (int m = 0; m < 3; m++) nv[m] = grid[(1 + m) * nn + idx];
                    for (int m = 0; m < 3; m++) p.vel[m] += weight * nv[m];
                    double s = (4. * h_inv) * weight;
                    for (int r = 0; r < 3; r++)
                        for (int q = 0; q < 3; q++) p.Bp(r, q) += s * (nv[r] * dpos[q]);
                }

        for (int a = 0; a < 3; a++) p.pos[a] += p.vel[a] * dt;
        Mat3 A = eye3();
        for (int k = 0; k < 9; k++) A.a[k] += dt * p.Bp.a[k];
        p.Fe = mul(A, p.Fe);

        if (isLiquid) {
            p.Jp_inv *= det3(A);
        } else {
            compute_pq(p, kappa, mu);
            if (!(p.utility_data & ICEMPM_STATUS_CRUSHED)) check_failure_surface(P, p);
            if (p.utility_data & ICEMPM_STATUS_CRUSHED) {
                compute_svd(p, kappa, mu);
                wolper_drucker_prager(P, p);
            }
        }

        for (int a = 0; a < 3; a++) {
            soa[i + (size_t)pitch * (ICEMPM_POSX + a)] = p.pos[a];
            soa[i + (size_t)pitch * (ICEMPM_VELX + a)] = p.vel[a];
            for (int b = 0; b < 3; b++) {
                soa[i + (size_t)pitch * (ICEMPM_FE00 + 3 * a + b)] = p.Fe(a, b);
                soa[i + (size_t)pitch * (ICEMPM_BP00 + 3 * a + b)] = p.Bp(a, b);
            }
        }
        soa[i + (size_t)pitch * ICEMPM_IDX_JP_INV] = p.Jp_inv;
        *ptr = bits_util(p.utility_data);
        if (record_pq) {
            soa[i + (size_t)pitch * ICEMPM_IDX_P] = p.p_tr;
            soa[i + (size_t)pitch * ICEMPM_IDX_Q] = p.q_tr;
        }
    }
    return code;
}

int oracle_step(const icempm_SimParams3D *prms, const icempm_Indenter *ind, double *soa, int pitch,
                int n, double *grid, double *bins, int n_substeps, double *sim_time,
                int simulation_step, int nthreads) {
    const icempm_SimParams3D &P = *prms;
    const size_t nn = (size_t)P.GridXTotal * P.GridY * P.GridZ;
    int code = 0;
    double t = *sim_time;
    for (int s = 0; s < n_substeps; s++) {
        t += P.InitialTimeStep;
        double center[3];
        for (int d = 0; d < 3; d++) center[d] = ind->center_initial[d] + t * ind->velocity[d];
        std::memset(grid, 0, sizeof(double) * nn * 4);
        code = std::max(code, oracle_p2g(prms, soa, pitch, n, grid, nthreads));
        oracle_update_nodes(prms, ind, center, grid, bins, nthreads);
        bool rec = ((simulation_step + s) % P.UpdateEveryNthStep) == 0;
        code = std::max(code, oracle_g2p(prms, soa, pitch, n, grid, rec ? 1 : 0, nthreads));
    }
    *sim_time = t;
    return code;
}

} // extern "C"
