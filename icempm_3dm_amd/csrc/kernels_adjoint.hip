// kernels_adjoint.hip — reverse mode of the step for the calibration
// adjoint: the forward run and its replay from checkpoints use the step's
// own kernels (P2G, grid update, G2P); these kernels run the reverse.
//
// The reference has no differentiation; these kernels differentiate its
// step (partition_kernel_p2g, _update_nodes, _g2p, mpm_kernels.cuh:20-351)
// with respect to the Lame parameters mu and kappa, for a loss that depends
// on the indenter force. Grids use the step's bricked layout.
//
// The reverse of substep s, given the adjoints of the point state at its
// end (x', v', F', C'), the point states at its start and end, the grid
// after P2G (m, mv) and the grid after the update (m, v):
//   G2P   x' = x + dt v',  F' = (I + dt C') F,  v' = sum W v_a,
//         C' = (4/h) sum W v_a (a - f)^T       -> v_bar at the nodes, x_bar, F_bar
//         (k_adj_g2p_bwd: brick segments, node velocities gathered from an
//         LDS tile, v_bar scattered through an LDS tile like P2G's: its
//         contribution W (v_hat + (4/h) c_hat (a - f)) has P2G's form)
//   grid  v = walls(contact(clamp(mv/m - g dt)))  (branches as taken forward),
//         force = (v_before_contact - v_after) m / dt enters the loss
//                                                -> m_bar, mv_bar at the nodes
//   P2G   m_a = sum W m,  mv_a = sum W (m v + (m C - D tau(F)) (a - f) h)
//                                                -> x_bar, v_bar, C_bar, F_bar,
//                                                   mu_bar, kappa_bar
#include "g2p_point.h"

namespace icempm {

namespace {

constexpr int kAdjThreads = 256;

// Derivatives of the quadratic B-spline weights with respect to f.
__device__ __forceinline__ void bspline3_d(double f, double dw[3]) {
    dw[0] = -(1.5 - f);
    dw[1] = -2.0 * (f - 1.0);
    dw[2] = f - 0.5;
}

__device__ __forceinline__ void load_mat(PtsView p, int line, int i, double M[9]) {
#pragma unroll
    for (int k = 0; k < 9; k++) M[k] = p.at(line + k, i);
}

// Sums v over the block and adds the result to *out (one atomic per block).
template <int kN>
__device__ __forceinline__ void block_add(double (&v)[kN], double *out) {
    __shared__ double red[kN][kAdjThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kN; k++) {
        double x = v[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
        if (lane == 0) red[k][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x < kN) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kAdjThreads / 64; w++) s += red[threadIdx.x][w];
        if (s != 0.0) unsafeAtomicAdd(&out[threadIdx.x], s);
    }
}

// The grid update of one node (partition_kernel_update_nodes,
// mpm_kernels.cuh:138-232, for one scene and SetupType != 2), keeping the
// intermediate velocities and branches the reverse pass needs.
struct NodeFwd {
    double u0[3], u1[3], u2[3], u3[3], v[3], n[3];
    bool clamp, contact;
    bool zero[3];
};

__device__ __forceinline__ void node_forward(const KParams &kp, const double c[3], int x, int y, int z, double mass,
                                             const double mv[3], NodeFwd &nf, double force[3]) {
    const double h = kp.h, dt = kp.dt;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        nf.u0[i] = mv[i] / mass;
        nf.u1[i] = nf.u0[i];
    }
    nf.u1[1] -= kp.dt_gravity;
    const double vsq = nf.u1[0] * nf.u1[0] + nf.u1[1] * nf.u1[1] + nf.u1[2] * nf.u1[2];
    nf.clamp = vsq > kp.vmax_sq;
    const double nrm = nf.clamp ? sqrt(vsq) : 1.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        nf.u2[i] = nf.clamp ? nf.u1[i] / nrm * kp.vmax : nf.u1[i];
        nf.u3[i] = nf.u2[i];
        force[i] = 0.0;
        nf.n[i] = 0.0;
    }
    nf.contact = false;
    const bool sphere = kp.ind_shape == ICEMPM_INDENTER_SPHERE;
    double nn_[3] = {x * h - c[0], y * h - c[1], sphere ? z * h - c[2] : 0.0};
    const double nsq = nn_[0] * nn_[0] + nn_[1] * nn_[1] + nn_[2] * nn_[2];
    if (nsq < kp.ind_rsq) {
        const double vrel[3] = {nf.u2[0] - kp.ind_vel[0], nf.u2[1] - kp.ind_vel[1], nf.u2[2] - kp.ind_vel[2]};
        if (nsq > 0) {
            const double r = sqrt(nsq);
#pragma unroll
            for (int i = 0; i < 3; i++) nn_[i] /= r;
        }
        const double vn = vrel[0] * nn_[0] + vrel[1] * nn_[1] + vrel[2] * nn_[2];
        if (vn < 0) {
            nf.contact = true;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                nf.n[i] = nn_[i];
                nf.u3[i] = kp.ind_vel[i] + (vrel[i] - nn_[i] * vn);
                force[i] = (nf.u2[i] - nf.u3[i]) * mass / dt;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        nf.v[i] = nf.u3[i];
        nf.zero[i] = false;
    }
    if (y <= 2) nf.zero[0] = nf.zero[1] = nf.zero[2] = true;
    if (y >= kp.gy - 3 && nf.v[1] > 0) nf.zero[1] = true;
    if (x <= 2 && nf.v[0] < 0) nf.zero[0] = true;
    else if (x >= kp.gx - 3 && nf.v[0] > 0) nf.zero[0] = true;
    if (z <= 2 && nf.v[2] < 0) nf.zero[2] = true;
    else if (z >= kp.gz - 3 && nf.v[2] > 0) nf.zero[2] = true;
#pragma unroll
    for (int i = 0; i < 3; i++)
        if (nf.zero[i]) nf.v[i] = 0.0;
}

// Reverse of the grid update over every node of the stored grid: Gbar
// (m_bar, mv_bar; all nodes written) from Vbar (lines 1-3) and the force
// adjoint fbar of this substep. G: the grid after P2G (m, mv).
__global__ __launch_bounds__(kAdjThreads) void k_adj_update_bwd(KParams kp, const double *G, double *Vbar,
                                                                double *Gbar, double cx, double cy, double cz,
                                                                double fb0, double fb1, double fb2) {
    const long long nn = kp.grid_pitch;
    const long long idx = (long long)blockIdx.x * kAdjThreads + threadIdx.x;
    if (idx >= nn) return;
    const double mass = G[idx];
    if (mass == 0.0) {
#pragma unroll
        for (int l = 0; l < 4; l++) {
            Gbar[l * nn + idx] = 0.0;
            Vbar[l * nn + idx] = 0.0;
        }
        return;
    }
    const long long b = idx >> 6;
    const int local = (int)(idx & 63);
    const int bz = (int)(b % kp.nbz), by = (int)((b / kp.nbz) % kp.nby), bx = (int)(b / ((long long)kp.nbz * kp.nby));
    const int x = bx * 4 + (local >> 4), y = by * 4 + ((local >> 2) & 3), z = bz * 4 + (local & 3);
    const double mv[3] = {G[nn + idx], G[2 * nn + idx], G[3 * nn + idx]};
    const double c[3] = {cx, cy, cz};
    NodeFwd nf;
    double f[3];
    node_forward(kp, c, x, y, z, mass, mv, nf, f);
    const double fbar[3] = {fb0, fb1, fb2};
    const double k = mass / kp.dt;
    double u3b[3], u2b[3], u1b[3], mb = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) u3b[i] = nf.zero[i] ? 0.0 : Vbar[(1 + i) * nn + idx];
    // this kernel is v_bar's only reader: leave it zeroed for the next substep
#pragma unroll
    for (int l = 0; l < 4; l++) Vbar[l * nn + idx] = 0.0;
    if (nf.contact) {
        // force = (u2 - u3) m / dt, u3 = u2 - n n.(u2 - v_ind)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            u3b[i] -= fbar[i] * k;
            mb += fbar[i] * (nf.u2[i] - nf.u3[i]) / kp.dt;
        }
        const double nu = nf.n[0] * u3b[0] + nf.n[1] * u3b[1] + nf.n[2] * u3b[2];
#pragma unroll
        for (int i = 0; i < 3; i++) u2b[i] = u3b[i] - nf.n[i] * nu + fbar[i] * k;
    } else {
#pragma unroll
        for (int i = 0; i < 3; i++) u2b[i] = u3b[i];
    }
    if (nf.clamp) {
        const double r = sqrt(nf.u1[0] * nf.u1[0] + nf.u1[1] * nf.u1[1] + nf.u1[2] * nf.u1[2]);
        const double uh[3] = {nf.u1[0] / r, nf.u1[1] / r, nf.u1[2] / r};
        const double d = uh[0] * u2b[0] + uh[1] * u2b[1] + uh[2] * u2b[2];
#pragma unroll
        for (int i = 0; i < 3; i++) u1b[i] = kp.vmax / r * (u2b[i] - uh[i] * d);
    } else {
#pragma unroll
        for (int i = 0; i < 3; i++) u1b[i] = u2b[i];
    }
    double dot = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        Gbar[(1 + i) * nn + idx] = u1b[i] / mass;
        dot += u1b[i] * nf.u0[i];
    }
    Gbar[idx] = mb - dot / mass;
}

// d(tau) : tau_bar for tau = mu J^(-2/3) dev(F F^T) + (kappa/2)(J^2 - 1) I:
// adds to F_bar, mu_bar and kappa_bar.
__device__ __forceinline__ void kirchhoff_adjoint(const KParams &kp, const double F[9], const double tb[9],
                                                  double Fb[9], double &mub, double &kb) {
    const double J = det3(F);
    double b[9];
    left_cauchy_green(F, b);
    const double trb = (b[0] + b[4] + b[8]) / 3;
    b[0] -= trb;
    b[4] -= trb;
    b[8] -= trb;
    const double s = rcbrt(J * J);
    double dbt = 0.0;
#pragma unroll
    for (int k = 0; k < 9; k++) dbt += b[k] * tb[k];
    const double trt = tb[0] + tb[4] + tb[8];
    mub += s * dbt;
    kb += 0.5 * (J * J - 1.0) * trt;
    const double cof[9] = {F[4] * F[8] - F[5] * F[7], F[5] * F[6] - F[3] * F[8], F[3] * F[7] - F[4] * F[6],
                           F[2] * F[7] - F[1] * F[8], F[0] * F[8] - F[2] * F[6], F[1] * F[6] - F[0] * F[7],
                           F[1] * F[5] - F[2] * F[4], F[2] * F[3] - F[0] * F[5], F[0] * F[4] - F[1] * F[3]};
    double T[9];
    const double t3 = trt / 3;
#pragma unroll
    for (int k = 0; k < 9; k++) T[k] = tb[k];
    T[0] -= t3;
    T[4] -= t3;
    T[8] -= t3;
    const double cJ = kp.mu * s * (-2.0 / 3.0) * dbt / J + kp.kappa * J * trt;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            double tsf = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) tsf += (T[3 * r + k] + T[3 * k + r]) * F[3 * k + q];
            Fb[3 * r + q] += kp.mu * s * tsf + cJ * cof[3 * r + q];
        }
}

// Reverse of return_map (Wolper_Drucker_Prager, mpm_kernels.cuh:494-544)
// for a point the G2P of this substep crushed or kept crushed: Ftr is the
// trial F (A F), Jp_inv its value before the update, Fb the adjoint of the
// returned F on entry and of Ftr on exit; th[0..3] gain the adjoints of
// (mu, kappa, DP_threshold_p, DP_tan_phi). The returned F is an isotropic
// function U diag(s(sigma)) V^T of Ftr: in the singular basis the diagonal
// carries ds/dsigma, and each off-diagonal pair (i, j) the factor
// (s_i - s_j)/(sigma_i - sigma_j) on its symmetric part and
// (s_i + s_j)/(sigma_i + sigma_j) on its antisymmetric part.
__device__ __noinline__ void return_map_adjoint(const KParams &kp, const double Ftr[9], double Jp_inv, double Fb[9],
                                                double th[4]) {
    PState p;
#pragma unroll
    for (int k = 0; k < 9; k++) p.Fe[k] = Ftr[k];
    p.Jp_inv = Jp_inv;
    p.util = 0;
    compute_pq(kp, p);
    double U[9], V[9], sig[3], sig2[3];
    svd3(p.Fe, U, sig, sig2, V);
    double G[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) acc += U[3 * a + i] * Fb[3 * a + b] * V[3 * b + j];
            G[3 * i + j] = acc;
        }
    const double mu = kp.mu, kappa = kp.kappa, pt = kp.dp_threshold_p, tphi = kp.tan_phi;
    double sh[3], ds[3][3], dth[3][4];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int m = 0; m < 4; m++) {
            if (m < 3) ds[k][m] = 0.0;
            dth[k][m] = 0.0;
        }
    if (p.p_tr < -pt || p.Jp_inv < 1) {
        const double Jn = kp.rm_je_new, c = kp.rm_cbrt_je_new;
        const double dc_dJn = c / (3 * Jn);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            sh[k] = c;
            dth[k][2] = dc_dJn / (kappa * Jn);
            dth[k][1] = -dc_dJn * pt / (kappa * kappa * Jn);
        }
    } else {
        const double pmax = kp.ice_comp, pmin = -kp.ice_tens, qmax = kp.ice_shear;
        double q, dq_dp = 0.0, dq_dpt = 0.0, dq_dt = 0.0;
        if (p.p_tr > pmax) {
            q = 0.0;
        } else {
            const double q_dp = (p.p_tr + pt) * tphi;
            const double root = sqrt((pmax - p.p_tr) * (p.p_tr - pmin));
            const double q_fs = 2 * root * qmax / (pmax - pmin);
            q = fmin(q_fs, q_dp);
            if (q_dp < q_fs) {
                dq_dp = tphi;
                dq_dpt = tphi;
                dq_dt = p.p_tr + pt;
            } else {
                dq_dp = qmax * (pmax + pmin - 2 * p.p_tr) / ((pmax - pmin) * root);
            }
        }
        if (!(p.q_tr >= q)) return; // not projected: F unchanged
        const double J = p.Je_tr, J23 = J * J * p.rc_Jsq * p.rc_Jsq;
        const double coef = q * kCoeff1Inv / mu * J23;
        const double mean = (sig2[0] + sig2[1] + sig2[2]) / 3;
        double d[3], n[3];
#pragma unroll
        for (int k = 0; k < 3; k++) d[k] = sig2[k] - mean;
        const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
        for (int k = 0; k < 3; k++) n[k] = r > 0 ? d[k] / r : 0.0;
        const double dJ[3] = {sig[1] * sig[2], sig[0] * sig[2], sig[0] * sig[1]};
        const double g = kCoeff1Inv / mu * J23 * (dq_dp * (-kappa * J) + (2.0 / 3.0) * q / J);
#pragma unroll
        for (int k = 0; k < 3; k++) sh[k] = sqrt(coef * n[k] + mean);
#pragma unroll
        for (int k = 0; k < 3; k++) {
#pragma unroll
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
    double H[9];
#pragma unroll
    for (int k = 0; k < 9; k++) H[k] = 0.0;
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
        for (int k = 0; k < 3; k++) H[4 * m] += ds[k][m] * G[4 * k];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int t = 0; t < 4; t++) th[t] += dth[k][t] * G[4 * k];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = i + 1; j < 3; j++) {
            double a;
            if (fabs(sig[i] - sig[j]) > 1e-8 * fmax(1.0, fabs(sig[i])))
                a = (sh[i] - sh[j]) / (sig[i] - sig[j]);
            else
                a = 0.5 * ((ds[i][i] - ds[i][j]) + (ds[j][j] - ds[j][i]));
            const double b = (sh[i] + sh[j]) / (sig[i] + sig[j]);
            const double sym = 0.5 * (G[3 * i + j] + G[3 * j + i]), asym = 0.5 * (G[3 * i + j] - G[3 * j + i]);
            H[3 * i + j] = a * sym + b * asym;
            H[3 * j + i] = a * sym - b * asym;
        }
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) acc += U[3 * a + i] * H[3 * i + j] * V[3 * b + j];
            Fb[3 * a + b] = acc;
        }
}

// Reverse of P2G for substep s: gathers Gbar at the point, combines with the
// G2P part and writes the adjoints of the state at the start of s into bar;
// mu_bar and kappa_bar are summed into theta[0..1].
__global__ __launch_bounds__(kAdjThreads) void k_adj_p2g_bwd(KParams kp, PtsView s0, const double *Gbar,
                                                             PtsView part, PtsView bar, double *theta, int n) {
    const long long nn = kp.grid_pitch;
    const int i = blockIdx.x * kAdjThreads + threadIdx.x;
    double th[2] = {0.0, 0.0};
    if (i < n) {
        const long long util = double_to_bits(s0.at(ICEMPM_IDX_UTILITY, i));
        if (!(util & ICEMPM_STATUS_DISABLED)) {
            double pos[3], vel[3], F[9], C[9];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                pos[a] = s0.at(ICEMPM_POSX + a, i);
                vel[a] = s0.at(ICEMPM_VELX + a, i);
            }
            load_mat(s0, ICEMPM_FE00, i, F);
            load_mat(s0, ICEMPM_BP00, i, C);
            int base[3];
            double fx[3];
            const bool inside = locate(kp, pos, base, fx) == 0; // else the forward refused the run
            double tau[9], sub[9];
            kirchhoff_stress(kp, F, tau);
            const double m = kp.mass_solid, h = kp.h, D = kp.dt_vol_Dpinv;
#pragma unroll
            for (int k = 0; k < 9; k++) sub[k] = m * C[k] - D * tau[k];
            double w[3][3], dw[3][3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                bspline3(fx[d], w[d]);
                bspline3_d(fx[d], dw[d]);
            }
            double vb[3] = {0.0, 0.0, 0.0}, fb[3] = {0.0, 0.0, 0.0}, subb[9];
#pragma unroll
            for (int k = 0; k < 9; k++) subb[k] = 0.0;
            for (int a = 0; a < (inside ? 3 : 0); a++)
                for (int b = 0; b < 3; b++)
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        const double W = w[0][a] * w[1][b] * w[2][c];
                        const double dW[3] = {dw[0][a] * w[1][b] * w[2][c], w[0][a] * dw[1][b] * w[2][c],
                                              w[0][a] * w[1][b] * dw[2][c]};
                        const double dp[3] = {(a - fx[0]) * h, (b - fx[1]) * h, (c - fx[2]) * h};
                        const long long idx = node_index(kp, base[0] + a, base[1] + b, base[2] + c);
                        const double mb = Gbar[idx];
                        const double pb[3] = {Gbar[nn + idx], Gbar[2 * nn + idx], Gbar[3 * nn + idx]};
                        double q = m * mb;
#pragma unroll
                        for (int r = 0; r < 3; r++) {
                            const double sd = sub[3 * r] * dp[0] + sub[3 * r + 1] * dp[1] + sub[3 * r + 2] * dp[2];
                            q += pb[r] * (vel[r] * m + sd);
                            vb[r] += W * m * pb[r];
#pragma unroll
                            for (int cc = 0; cc < 3; cc++) subb[3 * r + cc] += W * pb[r] * dp[cc];
                        }
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            const double stp = sub[d] * pb[0] + sub[3 + d] * pb[1] + sub[6 + d] * pb[2];
                            fb[d] += dW[d] * q - h * W * stp;
                        }
                    }
            double Fb[9], tb[9];
            load_mat(part, ICEMPM_FE00, i, Fb);
#pragma unroll
            for (int k = 0; k < 9; k++) tb[k] = -D * subb[k];
            kirchhoff_adjoint(kp, F, tb, Fb, th[0], th[1]);
#pragma unroll
            for (int a = 0; a < 3; a++) {
                bar.at(ICEMPM_POSX + a, i) = part.at(ICEMPM_POSX + a, i) + fb[a] * kp.h_inv;
                bar.at(ICEMPM_VELX + a, i) = vb[a];
            }
#pragma unroll
            for (int k = 0; k < 9; k++) {
                bar.at(ICEMPM_FE00 + k, i) = Fb[k];
                bar.at(ICEMPM_BP00 + k, i) = m * subb[k];
            }
        }
    }
    block_add<2>(th, theta);
}

// Reverse of G2P over brick segments. Per 256-point chunk, one lane per
// point: node velocities from the segment's LDS tile of the grid after the
// update (from HBM for points whose stencil left the tile since the sort),
// the G2P part of the point adjoints to `part`, and a P2G-form record
// (unit mass, alpha = v_hat, M = (4/h) c_hat) scattered into Vbar's tile and
// flushed with atomics at the segment end.
__device__ __forceinline__ void adj_record(P2GShared &sh, int t, bool valid, const int base[3], const double fx[3],
                                           const double alpha[3], const double M[9], const int org[3]) {
    int packed = -1;
    if (valid) {
        sh.rec(0, t) = 1.0; // the flush writes the nodes whose line 0 is non-zero
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double wd[3];
            bspline3(fx[d], wd);
            sh.rec(1 + 3 * d, t) = wd[0];
            sh.rec(2 + 3 * d, t) = wd[1];
            sh.rec(3 + 3 * d, t) = wd[2];
        }
#pragma unroll
        for (int r = 0; r < 3; r++) {
            // W (alpha + M (o - f)) = W (a + Mh o) with a = alpha - M f, Mh = M
            sh.rec(10 + r, t) = alpha[r] - (M[3 * r] * fx[0] + M[3 * r + 1] * fx[1] + M[3 * r + 2] * fx[2]);
#pragma unroll
            for (int c = 0; c < 3; c++) sh.rec(13 + 3 * r + c, t) = M[3 * r + c];
        }
        const int rx = base[0] - org[0], ry = base[1] - org[1], rz = base[2] - org[2];
        if ((unsigned)rx <= kTile - 3 && (unsigned)ry <= kTile - 3 && (unsigned)rz <= kTile - 3) {
            packed = rx | (ry << 3) | (rz << 6);
        } else {
            packed = -2;
            sh.absb[t] = base[0] | (base[1] << 10) | (base[2] << 20);
        }
    }
    sh.base[t] = packed;
}

// The gather of the reverse G2P: with node velocities V(a,b,c), returns the
// f-gradient fb of  v_hat . v' + c_hat : C'  (v' = sum W v, C' = c4 sum W v (o - f)^T).
template <typename Fetch>
__device__ __forceinline__ void adj_gather(const double w[3][3], const double dw[3][3], const double fx[3],
                                           const double vhat[3], const double chat[9], double c4, Fetch V,
                                           double fb[3]) {
    fb[0] = fb[1] = fb[2] = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double W = w[0][a] * w[1][b] * w[2][c];
                const double dW[3] = {dw[0][a] * w[1][b] * w[2][c], w[0][a] * dw[1][b] * w[2][c],
                                      w[0][a] * w[1][b] * dw[2][c]};
                const double d[3] = {a - fx[0], b - fx[1], c - fx[2]};
                double va[3];
                V(a, b, c, va);
                double cd[3];
#pragma unroll
                for (int r = 0; r < 3; r++) cd[r] = chat[3 * r] * d[0] + chat[3 * r + 1] * d[1] + chat[3 * r + 2] * d[2];
                const double g = va[0] * vhat[0] + va[1] * vhat[1] + va[2] * vhat[2] +
                                 c4 * (va[0] * cd[0] + va[1] * cd[1] + va[2] * cd[2]);
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    const double ctv = chat[q] * va[0] + chat[3 + q] * va[1] + chat[6 + q] * va[2];
                    fb[q] += dW[q] * g - c4 * W * ctv;
                }
            }
}

// kPlastic: some point is crushed in the run (else the return-mapping
// reverse, and its registers, stay out of the kernel). Two workgroups share
// a CU (LDS); the waves-per-EU bound keeps the kernel within the 256
// registers that lets both run.
template <bool kPlastic>
__global__ __launch_bounds__(kP2GThreads) __attribute__((amdgpu_waves_per_eu(2))) void k_adj_g2p_bwd(KParams kp, PtsView s0, PtsView s1, PtsView bar,
                                                             PtsView part, SegInfo seg, const double *__restrict__ V,
                                                             GridState gb, int cur_b, double *theta) {
    __shared__ P2GShared sh;
    __shared__ double sv[3][kTileAlloc];
    const int t = threadIdx.x;
    const BrickList bl = brick_list(gb, cur_b);
    const int nseg = *seg.nseg;
    const double dt = kp.dt, c4 = 4.0 * kp.h_inv;
    ChunkCursor c = cursor_first(seg, nseg, xcd_remap(blockIdx.x, gridDim.x), gridDim.x);
    while (c.valid()) {
        int org[3], bc[3];
        tile_origin_packed(c.coord, org, bc);
        if (c.first()) {
            VelTileRegs vt;
            vel_tile_fetch(vt, t, kp, org, V);
            vel_tile_store(sv, vt, t);
            p2g_zero_tile(sh, t);
            __syncthreads();
        }
        const int i = c.c0 + t;
        bool ok = false;
        int base[3] = {0, 0, 0};
        double fx[3] = {0.0, 0.0, 0.0}, vhat[3], M[9];
        double th[4] = {0.0, 0.0, 0.0, 0.0};
        if (i < c.end && !(double_to_bits(s0.at(ICEMPM_IDX_UTILITY, i)) & ICEMPM_STATUS_DISABLED)) {
            double pos[3];
#pragma unroll
            for (int a = 0; a < 3; a++) pos[a] = s0.at(ICEMPM_POSX + a, i);
            ok = locate(kp, pos, base, fx) == 0; // else the forward refused the run
            if (ok) {
                double F[9], C1[9], Fb[9], Cb[9], chat[9];
                load_mat(s0, ICEMPM_FE00, i, F);
                load_mat(s1, ICEMPM_BP00, i, C1);
                load_mat(bar, ICEMPM_FE00, i, Fb);
                load_mat(bar, ICEMPM_BP00, i, Cb);
                if (kPlastic && (double_to_bits(s1.at(ICEMPM_IDX_UTILITY, i)) & ICEMPM_STATUS_CRUSHED)) {
                    // this substep's G2P returned F to the yield surface: F'_bar -> trial F_bar
                    double Ftr[9];
#pragma unroll
                    for (int r = 0; r < 3; r++)
#pragma unroll
                        for (int q = 0; q < 3; q++)
                            Ftr[3 * r + q] = ((r == 0 ? 1.0 : 0.0) + dt * C1[3 * r]) * F[q] +
                                             ((r == 1 ? 1.0 : 0.0) + dt * C1[3 * r + 1]) * F[3 + q] +
                                             ((r == 2 ? 1.0 : 0.0) + dt * C1[3 * r + 2]) * F[6 + q];
                    return_map_adjoint(kp, Ftr, s0.at(ICEMPM_IDX_JP_INV, i), Fb, th);
                }
                double xb[3];
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    xb[a] = bar.at(ICEMPM_POSX + a, i);
                    vhat[a] = bar.at(ICEMPM_VELX + a, i) + dt * xb[a];
                }
                // F' = A F with A = I + dt C': F_bar = A^T F'_bar, C'_bar += dt F'_bar F^T
#pragma unroll
                for (int r = 0; r < 3; r++)
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        chat[3 * r + q] = Cb[3 * r + q] + dt * (Fb[3 * r] * F[3 * q] + Fb[3 * r + 1] * F[3 * q + 1] +
                                                               Fb[3 * r + 2] * F[3 * q + 2]);
                        double s = 0.0;
#pragma unroll
                        for (int k = 0; k < 3; k++) s += ((k == r ? 1.0 : 0.0) + dt * C1[3 * k + r]) * Fb[3 * k + q];
                        part.at(ICEMPM_FE00 + 3 * r + q, i) = s;
                    }
                double w[3][3], dw[3][3];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    bspline3(fx[d], w[d]);
                    bspline3_d(fx[d], dw[d]);
                }
                double fb[3];
                const int rx = base[0] - org[0], ry = base[1] - org[1], rz = base[2] - org[2];
                if ((unsigned)rx <= kTile - 3 && (unsigned)ry <= kTile - 3 && (unsigned)rz <= kTile - 3) {
                    const int t0 = tile_index(rx, ry, rz);
                    adj_gather(w, dw, fx, vhat, chat, c4,
                               [&](int a, int b, int cc, double va[3]) {
                                   const int tk = t0 + a * kTileSX + b * kTileSY + cc;
                                   va[0] = sv[0][tk];
                                   va[1] = sv[1][tk];
                                   va[2] = sv[2][tk];
                               },
                               fb);
                } else {
                    const long long pitch = kp.grid_pitch;
                    adj_gather(w, dw, fx, vhat, chat, c4,
                               [&](int a, int b, int cc, double va[3]) {
                                   const long long idx = node_index(kp, base[0] + a, base[1] + b, base[2] + cc);
                                   va[0] = V[pitch + idx];
                                   va[1] = V[2 * pitch + idx];
                                   va[2] = V[3 * pitch + idx];
                               },
                               fb);
                }
#pragma unroll
                for (int a = 0; a < 3; a++) part.at(ICEMPM_POSX + a, i) = xb[a] + fb[a] * kp.h_inv;
#pragma unroll
                for (int k = 0; k < 9; k++) M[k] = c4 * chat[k];
            }
        }
        adj_record(sh, t, ok, base, fx, vhat, M, org);
        // return-mapping parameter adjoints: wavefront sums, one atomic each
#pragma unroll
        for (int k = 0; k < (kPlastic ? 4 : 0); k++) {
            double x = th[k];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
            if ((t & 63) == 0 && x != 0.0) unsafeAtomicAdd(&theta[k], x);
        }
        __syncthreads();
        p2g_scatter(sh, t, min(kP2GChunk, c.end - c.c0), kp, gb, bl);
        if (c.last()) p2g_flush(sh, t, kp, org, bc, gb, bl);
        c = cursor_next(c, seg);
    }
}

// Total force of the accumulators, times scale, added to out[0..2].
__global__ __launch_bounds__(kAdjThreads) void k_sum_bins(const double *bins, long long nbins, double scale,
                                                          double *out) {
    double f[3] = {0.0, 0.0, 0.0};
    for (long long k = threadIdx.x; k < nbins / 3; k += kAdjThreads)
#pragma unroll
        for (int i = 0; i < 3; i++) f[i] += bins[3 * k + i];
#pragma unroll
    for (int i = 0; i < 3; i++) f[i] *= scale;
    block_add<3>(f, out);
}

// Marks crushed and liquid points in flags[0] / flags[1].
__global__ __launch_bounds__(kAdjThreads) void k_adj_flags(PtsView st, int n, int *flags) {
    const int i = blockIdx.x * kAdjThreads + threadIdx.x;
    if (i >= n) return;
    const long long util = double_to_bits(st.at(ICEMPM_IDX_UTILITY, i));
    if (util & ICEMPM_STATUS_DISABLED) return;
    if (util & ICEMPM_STATUS_CRUSHED) atomicOr(&flags[0], 1);
    if (util & ICEMPM_STATUS_LIQUID) atomicOr(&flags[1], 1);
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + kAdjThreads - 1) / kAdjThreads); }

// Point lines between the device order and the upload order: with
// to_upload, out line l [slot[i]] = in line l [i]; otherwise out line l [i] =
// in line l [slot[i]] (lines of n doubles; slot is a permutation of 0..n-1).
__global__ void k_lines_by_slot(const double *in, const int32_t *slot, double *out, int n, int lines,
                                int to_upload) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = slot[i];
    for (int l = 0; l < lines; l++) {
        if (to_upload) out[(long long)l * n + k] = in[(long long)l * n + i];
        else out[(long long)l * n + i] = in[(long long)l * n + k];
    }
}

} // namespace

hipError_t launch_adj_g2p_bwd(const KParams &kp, PtsView s0, PtsView s1, PtsView bar, PtsView part, int n,
                              SegInfo seg, const double *V, GridState gb, int cur_b, double *theta, bool plastic,
                              int n_blocks, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    n_blocks = segment_grid(n_blocks);
    if (plastic)
        k_adj_g2p_bwd<true><<<n_blocks, kP2GThreads, 0, stream>>>(kp, s0, s1, bar, part, seg, V, gb, cur_b, theta);
    else
        k_adj_g2p_bwd<false><<<n_blocks, kP2GThreads, 0, stream>>>(kp, s0, s1, bar, part, seg, V, gb, cur_b, theta);
    return hipGetLastError();
}

hipError_t launch_adj_update_bwd(const KParams &kp, const double *G, double *Vbar, double *Gbar,
                                 const double c[3], const double fbar[3], hipStream_t stream) {
    k_adj_update_bwd<<<blocks_for(kp.grid_pitch), kAdjThreads, 0, stream>>>(kp, G, Vbar, Gbar, c[0], c[1], c[2],
                                                                            fbar[0], fbar[1], fbar[2]);
    return hipGetLastError();
}

hipError_t launch_adj_p2g_bwd(const KParams &kp, PtsView s0, const double *Gbar, PtsView part, PtsView bar,
                              double *theta, int n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_adj_p2g_bwd<<<blocks_for(n), kAdjThreads, 0, stream>>>(kp, s0, Gbar, part, bar, theta, n);
    return hipGetLastError();
}

hipError_t launch_sum_bins(const double *bins, long long nbins, double scale, double *out, hipStream_t stream) {
    k_sum_bins<<<1, kAdjThreads, 0, stream>>>(bins, nbins, scale, out);
    return hipGetLastError();
}

hipError_t launch_adj_flags(PtsView st, int n, int *flags, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_adj_flags<<<blocks_for(n), kAdjThreads, 0, stream>>>(st, n, flags);
    return hipGetLastError();
}

hipError_t launch_lines_by_slot(const double *in, const int32_t *slot, double *out, int n, int lines, bool to_upload,
                                hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_lines_by_slot<<<(n + 255) / 256, 256, 0, stream>>>(in, slot, out, n, lines, to_upload ? 1 : 0);
    return hipGetLastError();
}

} // namespace icempm
