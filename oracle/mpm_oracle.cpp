// mpm_oracle.cpp — CPU restatement of the reference MPM step (see mpm_oracle.h).
//
// TEST INFRASTRUCTURE ONLY: the reference against which the HIP path is
// tested, and the host-core baseline. Each function names the reference lines it restates. Arithmetic
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
// ---------------------------------------------------------------- FP32 SVD
// The reference computes its 3x3 SVD in single precision with the method of
// McAdams, Selle, Tamstorf, Teran and Sifakis, "Computing the Singular Value
// Decomposition of 3x3 matrices with minimal branching and elementary
// floating point operations" (UW-Madison CS TR1690, 2011): Jacobi
// eigen-analysis of A^T A with approximate Givens rotations (a fixed number of
// sweeps), columns of A V sorted by decreasing norm with V kept a rotation,
// then a QR factorisation by Givens rotations giving U and the singular
// values (the last one negative for det A < 0). This is that method, written
// from the paper in float arithmetic, with the rotations accumulated as
// matrices; it is not bit-faithful to the reference (parity unpinned) and
// serves to measure how far an FP32 SVD moves the step from the FP64 one.
bool g_svd_fp32 = false;

// Approximate Givens rotation (cos, sin) that nearly diagonalises the
// symmetric 2x2 block [[app, apq], [apq, aqq]] (the paper's cheap angle, or
// pi/8 when that angle would be too coarse).
void approx_givens(float app, float apq, float aqq, float &c, float &s) {
    const float gamma = 5.82842712474619f; // 3 + 2 sqrt(2)
    float ch = 2.0f * (app - aqq), sh = apq;
    if (gamma * sh * sh < ch * ch) {
        const float w = 1.0f / std::sqrt(ch * ch + sh * sh);
        ch *= w;
        sh *= w;
    } else {
        ch = 0.92387953251128674f; // cos(pi/8)
        sh = 0.38268343236508977f; // sin(pi/8)
    }
    c = ch * ch - sh * sh;
    s = 2.0f * sh * ch;
}

// QR Givens rotation zeroing a2 against a1 (the paper's robust variant).
void qr_givens(float a1, float a2, float &c, float &s) {
    const float eps = 1e-6f;
    const float rho = std::sqrt(a1 * a1 + a2 * a2);
    float sh = rho > eps ? a2 : 0.0f;
    float ch = std::fabs(a1) + std::fmax(rho, eps);
    if (a1 < 0.0f) std::swap(sh, ch);
    const float w = 1.0f / std::sqrt(ch * ch + sh * sh);
    ch *= w;
    sh *= w;
    c = ch * ch - sh * sh;
    s = 2.0f * ch * sh;
}

void svd3_fp32(const Mat3 &F, Mat3 &U, double S[3], Mat3 &V) {
    float A[3][3], M[3][3], Vf[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            A[r][c] = (float)F(r, c);
            Vf[r][c] = r == c ? 1.0f : 0.0f;
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) M[i][j] = A[0][i] * A[0][j] + A[1][i] * A[1][j] + A[2][i] * A[2][j];
    // Jacobi sweeps: M <- G^T M G in the (p, q) plane, V <- V G
    const int pairs[3][2] = {{0, 1}, {1, 2}, {0, 2}};
    for (int sweep = 0; sweep < 4; sweep++)
        for (auto &pq : pairs) {
            const int p = pq[0], q = pq[1];
            float c, s;
            approx_givens(M[p][p], M[p][q], M[q][q], c, s);
            for (int k = 0; k < 3; k++) { // columns
                const float mp = M[k][p], mq = M[k][q];
                M[k][p] = c * mp + s * mq;
                M[k][q] = -s * mp + c * mq;
            }
            for (int k = 0; k < 3; k++) { // rows
                const float mp = M[p][k], mq = M[q][k];
                M[p][k] = c * mp + s * mq;
                M[q][k] = -s * mp + c * mq;
            }
            for (int k = 0; k < 3; k++) {
                const float vp = Vf[k][p], vq = Vf[k][q];
                Vf[k][p] = c * vp + s * vq;
                Vf[k][q] = -s * vp + c * vq;
            }
        }
    // B = A V, columns sorted by decreasing norm (a swap negates one column
    // so that V stays a rotation)
    float B[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) B[r][c] = A[r][0] * Vf[0][c] + A[r][1] * Vf[1][c] + A[r][2] * Vf[2][c];
    auto norm2 = [&](int c) { return B[0][c] * B[0][c] + B[1][c] * B[1][c] + B[2][c] * B[2][c]; };
    auto swap_cols = [&](int i, int j) {
        for (int r = 0; r < 3; r++) {
            const float b = B[r][i], v = Vf[r][i];
            B[r][i] = -B[r][j];
            Vf[r][i] = -Vf[r][j];
            B[r][j] = b;
            Vf[r][j] = v;
        }
    };
    if (norm2(0) < norm2(1)) swap_cols(0, 1);
    if (norm2(0) < norm2(2)) swap_cols(0, 2);
    if (norm2(1) < norm2(2)) swap_cols(1, 2);
    // QR: Givens rotations on rows (1,0), (2,0), (2,1); U accumulates their transposes
    float Uf[3][3] = {{1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}};
    const int rows[3][3] = {{0, 1, 0}, {0, 2, 0}, {1, 2, 1}}; // (pivot row, zeroed row, column)
    for (auto &rr : rows) {
        const int i = rr[0], j = rr[1], col = rr[2];
        float c, s;
        qr_givens(B[i][col], B[j][col], c, s);
        for (int k = 0; k < 3; k++) {
            const float bi = B[i][k], bj = B[j][k];
            B[i][k] = c * bi + s * bj;
            B[j][k] = -s * bi + c * bj;
        }
        for (int k = 0; k < 3; k++) {
            const float ui = Uf[k][i], uj = Uf[k][j];
            Uf[k][i] = c * ui + s * uj;
            Uf[k][j] = -s * ui + c * uj;
        }
    }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            U(r, c) = Uf[r][c];
            V(r, c) = Vf[r][c];
        }
    for (int k = 0; k < 3; k++) S[k] = B[k][k];
}

void compute_svd(Point &p, double kappa, double mu) {
    (void)kappa;
    if (g_svd_fp32) svd3_fp32(p.Fe, p.U, p.vSigma, p.V);
    else svd3(p.Fe, p.U, p.vSigma, p.V);
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

// The damage model (no reference counterpart; process-wide, set by
// oracle_set_damage): disabled unless set.
icempm_Damage g_damage = {0, 1.0, 1.0};

// Damage of a point: the IEEE single in bits 32-63 of the utility word,
// read into [0, 1] (NaN as 0).
double damage_of(long long util) {
    const uint32_t bits = (uint32_t)((unsigned long long)util >> 32);
    float d;
    std::memcpy(&d, &bits, sizeof(d));
    return d >= 1.0f ? 1.0 : (d > 0.0f ? (double)d : 0.0);
}

long long with_damage(long long util, double D) {
    const float d = (float)D;
    uint32_t bits;
    std::memcpy(&bits, &d, sizeof(bits));
    return (long long)(((unsigned long long)util & 0xffffffffull) | ((unsigned long long)bits << 32));
}

// Wolper_Drucker_Prager (mpm_kernels.cuh:494-544). With the damage model,
// the cohesion terms are scaled by f = 1 - (1 - residual) D and D grows by
// (q_tr - q_new) / (3 mu) / softening_strain afterwards.
void wolper_drucker_prager(const icempm_SimParams3D &P, Point &p) {
    const double mu = P.mu;
    const double kappa = P.kappa;
    const double tan_phi = P.DP_tan_phi;
    double DP_threshold_p = P.DP_threshold_p;
    const double pmax = P.IceCompressiveStrength;
    const double qmax = P.IceShearStrength;
    double D = 0.0;
    if (g_damage.enabled) {
        D = damage_of(p.utility_data);
        DP_threshold_p *= 1.0 - (1.0 - g_damage.residual_cohesion) * D;
    }
    double q_new = p.q_tr;

    if (p.p_tr < -DP_threshold_p || p.Jp_inv < 1) {
        double p_new = -DP_threshold_p;
        double Je_new = std::sqrt(-2. * p_new / kappa + 1.);
        double cbrt_Je_new = std::cbrt(Je_new);
        double s[3] = {cbrt_Je_new, cbrt_Je_new, cbrt_Je_new};
        p.Fe = usvt(p.U, s, p.V);
        p.Jp_inv *= Je_new / p.Je_tr;
        q_new = 0.0;
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
            q_new = q_n_1;
        }
    }
    if (g_damage.enabled) {
        const double dD = (p.q_tr - q_new) / (3.0 * mu) / g_damage.softening_strain;
        p.utility_data = with_damage(p.utility_data, std::fmin(1.0, D + dD));
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

void oracle_set_svd_fp32(int on) { g_svd_fp32 = on != 0; }

void oracle_svd3_fp32(const double F[9], double U[9], double S[3], double V[9]) {
    Mat3 f, u, v;
    std::memcpy(f.a, F, sizeof(f.a));
    svd3_fp32(f, u, S, v);
    std::memcpy(U, u.a, sizeof(u.a));
    std::memcpy(V, v.a, sizeof(v.a));
}

void oracle_set_damage(const icempm_Damage *damage) {
    g_damage = damage ? *damage : icempm_Damage{0, 1.0, 1.0};
}

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

// ====================================================================
// Calibration adjoint (no reference counterpart: the reference has no
// differentiation). Loss L = sum_k <dL/dF_k, F_k> to first order, where F_k
// is the total indenter force averaged over the substeps of sample k
// (sample_every substeps each; the per-substep force is the sum over
// contact nodes of (v_before - v_after) m / dt, mpm_kernels.cuh:189-201).
// Reverse mode over the step: P2G (Kirchhoff stress, APIC), the grid
// update (gravity, velocity clamp, indenter contact, walls), G2P and the
// plastic return mapping of crushed points, with every branch (crush test,
// tension cap or shear projection, contact, walls) held as the forward
// took it. Liquid points end the run with code 91.
// ====================================================================
namespace {

constexpr int kAdjUnsupported = 91;

inline void bspline_dw(double f, double dw[3]) {
    dw[0] = -(1.5 - f);
    dw[1] = -2.0 * (f - 1.0);
    dw[2] = f - 0.5;
}

struct PtState {
    double x[3], v[3];
    Mat3 F, C;
};

inline PtState load_state(const double *soa, int pitch, int i) {
    PtState s;
    for (int a = 0; a < 3; a++) {
        s.x[a] = soa[i + (size_t)pitch * (ICEMPM_POSX + a)];
        s.v[a] = soa[i + (size_t)pitch * (ICEMPM_VELX + a)];
    }
    for (int k = 0; k < 9; k++) {
        s.F.a[k] = soa[i + (size_t)pitch * (ICEMPM_FE00 + k)];
        s.C.a[k] = soa[i + (size_t)pitch * (ICEMPM_BP00 + k)];
    }
    return s;
}

// Forward grid update of one node (the oracle_update_nodes branch chain),
// keeping what the reverse pass needs. Returns the node's contact force.
struct NodeFwd {
    double u0[3], u1[3], u2[3], u3[3], v[3];
    bool clamp, contact;
    double n[3];
    bool wall_zero[3];
};

void node_forward(const icempm_SimParams3D &P, const icempm_Indenter &ind, const double center[3], int gi[3],
                  double mass, const double mv[3], NodeFwd &nf, double force[3]) {
    const double h = P.cellsize, dt = P.InitialTimeStep;
    const bool sphere = ind.shape == ICEMPM_INDENTER_SPHERE;
    for (int i = 0; i < 3; i++) nf.u0[i] = mv[i] / mass;
    for (int i = 0; i < 3; i++) nf.u1[i] = nf.u0[i];
    nf.u1[1] -= P.dt_Gravity;
    double vsq = nf.u1[0] * nf.u1[0] + nf.u1[1] * nf.u1[1] + nf.u1[2] * nf.u1[2];
    nf.clamp = vsq > P.vmax_squared;
    for (int i = 0; i < 3; i++) nf.u2[i] = nf.clamp ? nf.u1[i] / std::sqrt(vsq) * P.vmax : nf.u1[i];
    for (int i = 0; i < 3; i++) nf.u3[i] = nf.u2[i];
    nf.contact = false;
    force[0] = force[1] = force[2] = 0.0;
    if (P.SetupType != 2) {
        double g[3] = {gi[0] * h, gi[1] * h, gi[2] * h};
        double n[3] = {g[0] - center[0], g[1] - center[1], sphere ? g[2] - center[2] : 0.0};
        double nsq = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
        if (nsq < P.IndRSq) {
            double vrel[3] = {nf.u2[0] - ind.velocity[0], nf.u2[1] - ind.velocity[1], nf.u2[2] - ind.velocity[2]};
            if (nsq > 0) {
                double nrm = std::sqrt(nsq);
                for (int i = 0; i < 3; i++) n[i] /= nrm;
            }
            double vn = vrel[0] * n[0] + vrel[1] * n[1] + vrel[2] * n[2];
            if (vn < 0) {
                nf.contact = true;
                for (int i = 0; i < 3; i++) {
                    nf.n[i] = n[i];
                    nf.u3[i] = ind.velocity[i] + (vrel[i] - n[i] * vn);
                    force[i] = (nf.u2[i] - nf.u3[i]) * mass / dt;
                }
            }
        }
    }
    for (int i = 0; i < 3; i++) {
        nf.v[i] = nf.u3[i];
        nf.wall_zero[i] = false;
    }
    const int GX = P.GridXTotal, GY = P.GridY, GZ = P.GridZ;
    if (P.SetupType != 2) {
        if (gi[1] <= 2) nf.wall_zero[0] = nf.wall_zero[1] = nf.wall_zero[2] = true;
    } else if (gi[1] <= 2 && nf.v[1] < 0) {
        nf.wall_zero[1] = true;
    }
    if (gi[1] >= GY - 3 && nf.v[1] > 0) nf.wall_zero[1] = true;
    if (gi[0] <= 2 && nf.v[0] < 0) nf.wall_zero[0] = true;
    else if (gi[0] >= GX - 3 && nf.v[0] > 0) nf.wall_zero[0] = true;
    if (gi[2] <= 2 && nf.v[2] < 0) nf.wall_zero[2] = true;
    else if (gi[2] >= GZ - 3 && nf.v[2] > 0) nf.wall_zero[2] = true;
    for (int i = 0; i < 3; i++)
        if (nf.wall_zero[i]) nf.v[i] = 0.0;
}

// d(tau):tau_bar for the Kirchhoff stress tau = mu J^(-2/3) dev(F F^T) +
// (kappa/2)(J^2 - 1) I: returns F_bar (added), mu_bar and kappa_bar (added).
void kirchhoff_adjoint(double mu, double kappa, const Mat3 &F, const Mat3 &tb, Mat3 &Fbar, double &mub,
                       double &kappab) {
    const double J = det3(F);
    const Mat3 b = mul_abt(F, F);
    const Mat3 db = dev(b);
    const double s = rcbrt(J * J); // J^(-2/3)
    double dbt = 0, trt = tb(0, 0) + tb(1, 1) + tb(2, 2);
    for (int k = 0; k < 9; k++) dbt += db.a[k] * tb.a[k];
    mub += s * dbt;
    kappab += 0.5 * (J * J - 1.0) * trt;
    // cofactor matrix: J F^{-T}
    Mat3 cof;
    cof(0, 0) = F(1, 1) * F(2, 2) - F(1, 2) * F(2, 1);
    cof(0, 1) = F(1, 2) * F(2, 0) - F(1, 0) * F(2, 2);
    cof(0, 2) = F(1, 0) * F(2, 1) - F(1, 1) * F(2, 0);
    cof(1, 0) = F(0, 2) * F(2, 1) - F(0, 1) * F(2, 2);
    cof(1, 1) = F(0, 0) * F(2, 2) - F(0, 2) * F(2, 0);
    cof(1, 2) = F(0, 1) * F(2, 0) - F(0, 0) * F(2, 1);
    cof(2, 0) = F(0, 1) * F(1, 2) - F(0, 2) * F(1, 1);
    cof(2, 1) = F(0, 2) * F(1, 0) - F(0, 0) * F(1, 2);
    cof(2, 2) = F(0, 0) * F(1, 1) - F(0, 1) * F(1, 0);
    // dev(tb) symmetrised, times F
    Mat3 T = dev(tb);
    Mat3 TS;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) TS(r, c) = T(r, c) + T(c, r);
    const Mat3 TSF = mul(TS, F);
    // dJ = cof : dF; d(J^(-2/3)) = -(2/3) J^(-2/3) dJ / J
    const double cJ = mu * s * (-2.0 / 3.0) * dbt / J + kappa * J * trt;
    for (int k = 0; k < 9; k++) Fbar.a[k] += mu * s * TSF.a[k] + cJ * cof.a[k];
}

// Reverse of wolper_drucker_prager for a point crushed in this substep:
// Ftr is the trial F (A F), Jp_inv its value before the update, Fb the
// adjoint of the returned F on entry and of Ftr on exit; th[0..3] gain the
// adjoints of (mu, kappa, DP_threshold_p, DP_tan_phi). The returned F is an
// isotropic function U diag(s(sigma)) V^T of Ftr: in the singular basis the
// diagonal carries ds/dsigma, and each off-diagonal pair (i, j) the factors
// (s_i - s_j)/(sigma_i - sigma_j) on its symmetric part and
// (s_i + s_j)/(sigma_i + sigma_j) on its antisymmetric part.
void return_map_adjoint(const icempm_SimParams3D &P, const Mat3 &Ftr, double Jp_inv, Mat3 &Fb, double th[4]) {
    Point p;
    p.Fe = Ftr;
    p.Jp_inv = Jp_inv;
    compute_pq(p, P.kappa, P.mu);
    compute_svd(p, P.kappa, P.mu);
    const double mu = P.mu, kappa = P.kappa, pt = P.DP_threshold_p, tphi = P.DP_tan_phi;
    const double *sig = p.vSigma;
    // Uhat = U^T Fb V
    Mat3 G;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double acc = 0;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) acc += p.U(a, i) * Fb(a, b) * p.V(b, j);
            G(i, j) = acc;
        }
    double sh[3], ds[3][3] = {{0}}, dth[3][4] = {{0}};
    if (p.p_tr < -pt || p.Jp_inv < 1) {
        const double Jn = std::sqrt(-2. * (-pt) / kappa + 1.);
        const double c = std::cbrt(Jn);
        const double dc_dJn = c / (3 * Jn);
        for (int k = 0; k < 3; k++) {
            sh[k] = c;
            dth[k][2] = dc_dJn / (kappa * Jn);           // dJn/dpt = 1 / (kappa Jn)
            dth[k][1] = -dc_dJn * pt / (kappa * kappa * Jn); // dJn/dkappa
        }
    } else {
        const double pmax = P.IceCompressiveStrength, pmin = -P.IceTensileStrength, qmax = P.IceShearStrength;
        double q, dq_dp = 0, dq_dpt = 0, dq_dt = 0;
        if (p.p_tr > pmax) {
            q = 0;
        } else {
            const double q_dp = (p.p_tr + pt) * tphi;
            const double root = std::sqrt((pmax - p.p_tr) * (p.p_tr - pmin));
            const double q_fs = 2 * root * qmax / (pmax - pmin);
            q = std::fmin(q_fs, q_dp);
            if (q_dp < q_fs) {
                dq_dp = tphi;
                dq_dpt = tphi;
                dq_dt = p.p_tr + pt;
            } else {
                dq_dp = qmax * (pmax + pmin - 2 * p.p_tr) / ((pmax - pmin) * root);
            }
        }
        if (!(p.q_tr >= q)) return; // not projected: F unchanged
        const double J = p.Je_tr, J23 = std::cbrt(J * J);
        const double K = q * kCoeff1Inv / mu; // coef = K J^(2/3)
        const double coef = K * J23;
        double s2[3], d[3];
        for (int k = 0; k < 3; k++) s2[k] = sig[k] * sig[k];
        const double mean = (s2[0] + s2[1] + s2[2]) / 3;
        for (int k = 0; k < 3; k++) d[k] = s2[k] - mean;
        const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double n[3];
        for (int k = 0; k < 3; k++) n[k] = r > 0 ? d[k] / r : 0.0;
        const double dJ[3] = {sig[1] * sig[2], sig[0] * sig[2], sig[0] * sig[1]};
        // dcoef/dsigma_m = (kCoeff1Inv/mu) J^(2/3) dJ_m (dq/dp dp/dJ + (2/3) q / J), dp/dJ = -kappa J
        const double g = kCoeff1Inv / mu * J23 * (dq_dp * (-kappa * J) + (2.0 / 3.0) * q / J);
        for (int k = 0; k < 3; k++) sh[k] = std::sqrt(coef * n[k] + mean);
        for (int k = 0; k < 3; k++) {
            for (int m = 0; m < 3; m++) {
                const double dn = r > 0 ? 2 * sig[m] / r * ((k == m ? 1.0 : 0.0) - n[k] * n[m] - 1.0 / 3.0) : 0.0;
                ds[k][m] = (coef * dn + n[k] * g * dJ[m] + 2 * sig[m] / 3) / (2 * sh[k]);
            }
            const double f = n[k] / (2 * sh[k]);
            dth[k][0] = f * (-coef / mu);
            dth[k][1] = f * kCoeff1Inv / mu * J23 * dq_dp * (-(J * J - 1) / 2);
            dth[k][2] = f * kCoeff1Inv / mu * J23 * dq_dpt;
            dth[k][3] = f * kCoeff1Inv / mu * J23 * dq_dt;
        }
    }
    Mat3 H = zero3();
    for (int m = 0; m < 3; m++)
        for (int k = 0; k < 3; k++) H(m, m) += ds[k][m] * G(k, k);
    for (int k = 0; k < 3; k++)
        for (int t = 0; t < 4; t++) th[t] += dth[k][t] * G(k, k);
    for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 3; j++) {
            double a;
            if (std::fabs(sig[i] - sig[j]) > 1e-8 * std::max(1.0, std::fabs(sig[i])))
                a = (sh[i] - sh[j]) / (sig[i] - sig[j]);
            else
                a = 0.5 * ((ds[i][i] - ds[i][j]) + (ds[j][j] - ds[j][i]));
            const double b = (sh[i] + sh[j]) / (sig[i] + sig[j]);
            const double sym = 0.5 * (G(i, j) + G(j, i)), asym = 0.5 * (G(i, j) - G(j, i));
            H(i, j) = a * sym + b * asym;
            H(j, i) = a * sym - b * asym;
        }
    // Fb = U H V^T
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            double acc = 0;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) acc += p.U(a, i) * H(i, j) * p.V(b, j);
            Fb(a, b) = acc;
        }
}

} // namespace

extern "C" {

int oracle_adj_forward(const icempm_SimParams3D *prms, const icempm_Indenter *ind, const double *soa0, int pitch,
                       int n, int n_steps, int sample_every, double *forces) {
    if (g_damage.enabled) return ICEMPM_ADJ_UNSUPPORTED; // the damage model is not differentiated
    const icempm_SimParams3D &P = *prms;
    const size_t nn = (size_t)P.GridXTotal * P.GridY * P.GridZ;
    std::vector<double> soa(soa0, soa0 + (size_t)pitch * ICEMPM_N_PTS_ARRAYS);
    std::vector<double> grid(nn * 4);
    const int nsamp = n_steps / sample_every;
    for (int k = 0; k < 3 * nsamp; k++) forces[k] = 0.0;
    double t = 0.0;
    for (int s = 0; s < n_steps; s++) {
        t += P.InitialTimeStep;
        double center[3];
        for (int d = 0; d < 3; d++) center[d] = ind->center_initial[d] + t * ind->velocity[d];
        std::fill(grid.begin(), grid.end(), 0.0);
        int code = oracle_p2g(prms, soa.data(), pitch, n, grid.data(), 1);
        if (code) return code;
        double fs[3] = {0, 0, 0};
        for (size_t idx = 0; idx < nn; idx++) {
            const double mass = grid[idx];
            if (mass == 0) continue;
            int gi[3] = {(int)(idx / ((size_t)P.GridZ * P.GridY)), (int)((idx / P.GridZ) % P.GridY),
                         (int)(idx % P.GridZ)};
            double mv[3] = {grid[nn + idx], grid[2 * nn + idx], grid[3 * nn + idx]};
            NodeFwd nf;
            double f[3];
            node_forward(P, *ind, center, gi, mass, mv, nf, f);
            for (int i = 0; i < 3; i++) {
                fs[i] += f[i];
                grid[(1 + i) * nn + idx] = nf.v[i];
            }
        }
        if (s / sample_every < nsamp)
            for (int i = 0; i < 3; i++) forces[3 * (s / sample_every) + i] += fs[i] / sample_every;
        code = oracle_g2p(prms, soa.data(), pitch, n, grid.data(), 0, 1);
        if (code) return code;
    }
    return 0;
}

int oracle_adj_gradient(const icempm_SimParams3D *prms, const icempm_Indenter *ind, const double *soa0, int pitch,
                        int n, int n_steps, int sample_every, const double *dL_dF, double *grad_mu_kappa) {
    return oracle_adj_gradient_ex(prms, ind, soa0, pitch, n, n_steps, sample_every, dL_dF, nullptr, grad_mu_kappa);
}

int oracle_adj_gradient_ex(const icempm_SimParams3D *prms, const icempm_Indenter *ind, const double *soa0, int pitch,
                           int n, int n_steps, int sample_every, const double *dL_dF, const double *dL_dx,
                           double *grad_mu_kappa) {
    if (g_damage.enabled) return ICEMPM_ADJ_UNSUPPORTED; // the damage model is not differentiated
    // grad_mu_kappa: dL/d(mu, kappa, DP_threshold_p, DP_tan_phi)
    const icempm_SimParams3D &P = *prms;
    const size_t nn = (size_t)P.GridXTotal * P.GridY * P.GridZ;
    const size_t lines = (size_t)pitch * ICEMPM_N_PTS_ARRAYS;
    const double h = P.cellsize, h_inv = P.cellsize_inv, dt = P.InitialTimeStep;
    const double mu = P.mu, kappa = P.kappa, D = P.dt_vol_Dpinv;
    const int nsamp = n_steps / sample_every;
    // forward, keeping every substep's start state
    std::vector<double> states((size_t)(n_steps + 1) * lines);
    std::copy(soa0, soa0 + lines, states.begin());
    std::vector<double> grid(nn * 4);
    std::vector<double> centers(3 * (size_t)n_steps);
    double t = 0.0;
    for (int s = 0; s < n_steps; s++) {
        double *cur = &states[(size_t)(s + 1) * lines];
        std::copy(&states[(size_t)s * lines], &states[(size_t)s * lines] + lines, cur);
        t += dt;
        for (int d = 0; d < 3; d++) centers[3 * s + d] = ind->center_initial[d] + t * ind->velocity[d];
        std::fill(grid.begin(), grid.end(), 0.0);
        int code = oracle_p2g(prms, cur, pitch, n, grid.data(), 1);
        if (code) return code;
        std::vector<double> bins(std::max(P.IndenterSubdivisions * P.GridZ * 3, 1));
        oracle_update_nodes(prms, ind, &centers[3 * s], grid.data(), bins.data(), 1);
        code = oracle_g2p(prms, cur, pitch, n, grid.data(), 0, 1);
        if (code) return code;
        for (int i = 0; i < n; i++)
            if (util_bits(cur[i]) & ICEMPM_STATUS_LIQUID) return kAdjUnsupported;
    }
    // reverse
    // the adjoint of the final state: dL/dx_N (3 lines of n) or zero
    std::vector<PtState> bar(n);
    for (int p = 0; p < n; p++) {
        PtState &b = bar[p];
        for (int a = 0; a < 3; a++) {
            b.x[a] = dL_dx ? dL_dx[(size_t)a * n + p] : 0.0;
            b.v[a] = 0.0;
        }
        b.F = zero3();
        b.C = zero3();
    }
    std::vector<double> gm(nn * 4), vgrid(nn * 3), vbar(nn * 3), gbar(nn * 4);
    double mub = 0.0, kappab = 0.0;
    double th[4] = {0.0, 0.0, 0.0, 0.0}; // return-map part: mu, kappa, DP_threshold_p, DP_tan_phi
    for (int s = n_steps - 1; s >= 0; s--) {
        const double *st = &states[(size_t)s * lines];
        const double *nx = &states[(size_t)(s + 1) * lines];
        const double *center = &centers[3 * s];
        double fbar[3] = {0, 0, 0};
        if (s / sample_every < nsamp)
            for (int i = 0; i < 3; i++) fbar[i] = dL_dF ? dL_dF[3 * (s / sample_every) + i] / sample_every : 0.0;
        // recompute the grid of substep s
        std::fill(gm.begin(), gm.end(), 0.0);
        oracle_p2g(prms, st, pitch, n, gm.data(), 1);
        std::vector<NodeFwd> nodes(nn);
        for (size_t idx = 0; idx < nn; idx++) {
            for (int i = 0; i < 3; i++) vgrid[i * nn + idx] = 0.0;
            if (gm[idx] == 0) continue;
            int gi[3] = {(int)(idx / ((size_t)P.GridZ * P.GridY)), (int)((idx / P.GridZ) % P.GridY),
                         (int)(idx % P.GridZ)};
            double mv[3] = {gm[nn + idx], gm[2 * nn + idx], gm[3 * nn + idx]};
            double f[3];
            node_forward(P, *ind, center, gi, gm[idx], mv, nodes[idx], f);
            for (int i = 0; i < 3; i++) vgrid[i * nn + idx] = nodes[idx].v[i];
        }
        // G2P reverse: scatter v_bar to the nodes, partial point adjoints
        std::fill(vbar.begin(), vbar.end(), 0.0);
        std::vector<PtState> part(n);
        for (int p = 0; p < n; p++) {
            const PtState S0 = load_state(st, pitch, p);
            const PtState S1 = load_state(nx, pitch, p);
            PtState &B = bar[p];
            Weights W;
            weights_and_check(P, S0.x, W);
            Mat3 A = eye3();
            for (int k = 0; k < 9; k++) A.a[k] += dt * S1.C.a[k];
            if (util_bits(nx[p]) & ICEMPM_STATUS_CRUSHED) {
                // the G2P of this substep returned F to the yield surface
                const Mat3 Ftr = mul(A, S0.F);
                return_map_adjoint(P, Ftr, st[p + (size_t)pitch * ICEMPM_IDX_JP_INV], B.F, th);
            }
            Mat3 chat = B.C;
            const Mat3 FbFt = mul_abt(B.F, S0.F);
            for (int k = 0; k < 9; k++) chat.a[k] += dt * FbFt.a[k];
            double vhat[3];
            for (int a = 0; a < 3; a++) vhat[a] = B.v[a] + dt * B.x[a];
            PtState &Q = part[p];
            // F_bar through F' = A F
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++)
                    Q.F(r, c) = A(0, r) * B.F(0, c) + A(1, r) * B.F(1, c) + A(2, r) * B.F(2, c);
            double dw[3][3];
            for (int d = 0; d < 3; d++) bspline_dw(W.fx[d], dw[d]);
            double fb[3] = {0, 0, 0};
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    for (int c = 0; c < 3; c++) {
                        const int o[3] = {a, b, c};
                        const double Wa = W.w[a][0] * W.w[b][1] * W.w[c][2];
                        const double dWa[3] = {dw[0][a] * W.w[b][1] * W.w[c][2], W.w[a][0] * dw[1][b] * W.w[c][2],
                                               W.w[a][0] * W.w[b][1] * dw[2][c]};
                        const size_t idx = node_index(P, W.base[0] + a, W.base[1] + b, W.base[2] + c);
                        double va[3], df[3];
                        for (int i = 0; i < 3; i++) {
                            va[i] = vgrid[i * nn + idx];
                            df[i] = o[i] - W.fx[i];
                        }
                        double cd[3]; // chat (a - f)
                        for (int r = 0; r < 3; r++) cd[r] = chat(r, 0) * df[0] + chat(r, 1) * df[1] + chat(r, 2) * df[2];
                        for (int r = 0; r < 3; r++) vbar[r * nn + idx] += Wa * (vhat[r] + 4.0 * h_inv * cd[r]);
                        const double g = va[0] * vhat[0] + va[1] * vhat[1] + va[2] * vhat[2] +
                                         4.0 * h_inv * (va[0] * cd[0] + va[1] * cd[1] + va[2] * cd[2]);
                        for (int d = 0; d < 3; d++) {
                            const double ctv = chat(0, d) * va[0] + chat(1, d) * va[1] + chat(2, d) * va[2];
                            fb[d] += dWa[d] * g - 4.0 * h_inv * Wa * ctv;
                        }
                    }
            for (int d = 0; d < 3; d++) Q.x[d] = B.x[d] + fb[d] * h_inv;
        }
        // grid update reverse
        std::fill(gbar.begin(), gbar.end(), 0.0);
        for (size_t idx = 0; idx < nn; idx++) {
            const double mass = gm[idx];
            if (mass == 0) continue;
            const NodeFwd &nf = nodes[idx];
            double u3b[3], u2b[3], u1b[3];
            for (int i = 0; i < 3; i++) u3b[i] = nf.wall_zero[i] ? 0.0 : vbar[i * nn + idx];
            double mb = 0.0;
            if (nf.contact) {
                // force = (u2 - u3) m / dt
                for (int i = 0; i < 3; i++) {
                    u3b[i] -= fbar[i] * mass / dt;
                    mb += fbar[i] * (nf.u2[i] - nf.u3[i]) / dt;
                }
                const double nu = nf.n[0] * u3b[0] + nf.n[1] * u3b[1] + nf.n[2] * u3b[2];
                for (int i = 0; i < 3; i++) u2b[i] = u3b[i] - nf.n[i] * nu + fbar[i] * mass / dt;
            } else {
                for (int i = 0; i < 3; i++) u2b[i] = u3b[i];
            }
            if (nf.clamp) {
                const double nrm = std::sqrt(nf.u1[0] * nf.u1[0] + nf.u1[1] * nf.u1[1] + nf.u1[2] * nf.u1[2]);
                double uh[3] = {nf.u1[0] / nrm, nf.u1[1] / nrm, nf.u1[2] / nrm};
                const double d = uh[0] * u2b[0] + uh[1] * u2b[1] + uh[2] * u2b[2];
                for (int i = 0; i < 3; i++) u1b[i] = P.vmax / nrm * (u2b[i] - uh[i] * d);
            } else {
                for (int i = 0; i < 3; i++) u1b[i] = u2b[i];
            }
            // u0 = mv / m
            double dot = 0;
            for (int i = 0; i < 3; i++) {
                gbar[(1 + i) * nn + idx] = u1b[i] / mass;
                dot += u1b[i] * nf.u0[i];
            }
            gbar[idx] = mb - dot / mass;
        }
        // P2G reverse
        for (int p = 0; p < n; p++) {
            const PtState S0 = load_state(st, pitch, p);
            const long long util = util_bits(st[p]);
            const double m = (util & ICEMPM_STATUS_LIQUID) ? P.LiquidParticleMass : P.ParticleMass;
            Weights W;
            weights_and_check(P, S0.x, W);
            const Mat3 tau = kirchhoff(mu, kappa, S0.F);
            Mat3 sub;
            for (int k = 0; k < 9; k++) sub.a[k] = m * S0.C.a[k] - D * tau.a[k];
            double dw[3][3];
            for (int d = 0; d < 3; d++) bspline_dw(W.fx[d], dw[d]);
            double vb[3] = {0, 0, 0}, fb[3] = {0, 0, 0};
            Mat3 subb = zero3();
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    for (int c = 0; c < 3; c++) {
                        const int o[3] = {a, b, c};
                        const double Wa = W.w[a][0] * W.w[b][1] * W.w[c][2];
                        const double dWa[3] = {dw[0][a] * W.w[b][1] * W.w[c][2], W.w[a][0] * dw[1][b] * W.w[c][2],
                                               W.w[a][0] * W.w[b][1] * dw[2][c]};
                        const size_t idx = node_index(P, W.base[0] + a, W.base[1] + b, W.base[2] + c);
                        const double mb = gbar[idx];
                        double pb[3], dp[3];
                        for (int i = 0; i < 3; i++) {
                            pb[i] = gbar[(1 + i) * nn + idx];
                            dp[i] = (o[i] - W.fx[i]) * h;
                        }
                        double q = m * mb;
                        for (int r = 0; r < 3; r++) {
                            const double sd = sub(r, 0) * dp[0] + sub(r, 1) * dp[1] + sub(r, 2) * dp[2];
                            q += pb[r] * (S0.v[r] * m + sd);
                            vb[r] += Wa * m * pb[r];
                            for (int cc = 0; cc < 3; cc++) subb(r, cc) += Wa * pb[r] * dp[cc];
                        }
                        for (int d = 0; d < 3; d++) {
                            const double stp = sub(0, d) * pb[0] + sub(1, d) * pb[1] + sub(2, d) * pb[2];
                            fb[d] += dWa[d] * q - h * Wa * stp;
                        }
                    }
            PtState &B = bar[p];
            const PtState &Q = part[p];
            for (int d = 0; d < 3; d++) {
                B.x[d] = Q.x[d] + fb[d] * h_inv;
                B.v[d] = vb[d];
            }
            for (int k = 0; k < 9; k++) B.C.a[k] = m * subb.a[k];
            Mat3 tb;
            for (int k = 0; k < 9; k++) tb.a[k] = -D * subb.a[k];
            B.F = Q.F;
            kirchhoff_adjoint(mu, kappa, S0.F, tb, B.F, mub, kappab);
        }
    }
    grad_mu_kappa[0] = mub + th[0];
    grad_mu_kappa[1] = kappab + th[1];
    grad_mu_kappa[2] = th[2];
    grad_mu_kappa[3] = th[3];
    return 0;
}

} // extern "C"
