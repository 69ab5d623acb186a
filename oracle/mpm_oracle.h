/*
 * mpm_oracle.h — CPU restatement of the reference's single-partition MPM step.
 *
 * TEST INFRASTRUCTURE ONLY. This library checks the HIP path: it is loaded by
 * tests/, by the repository's smoke test and by bench.py's host-core
 * baseline, and nothing in the product links or calls it.
 *
 * Every function restates a reference function (file:line given beside it)
 * in plain C++ on FP64 host arithmetic. Buffers use the reference layouts:
 *   - particles: HostSideSOA lines (28 lines of `pitch` doubles, line k at
 *     soa + k*pitch; line 0 holds int64 bit patterns);
 *   - grid: 4 lines (m, mvx, mvy, mvz; velocity after the update) of
 *     GX*GY*GZ doubles each, node (x,y,z) at x*GY*GZ + y*GZ + z — the
 *     reference's x-major order without its halo slices.
 *
 * Deliberate deviations from the reference (all documented in the repo's
 * design notes):
 *   - the 3x3 SVD is an FP64 one-sided Jacobi SVD instead of the
 *     reference's FP32 McAdams routine (helper_math.cuh:23-1119);
 *   - a particle whose stencil leaves the grid or whose position is not
 *     finite raises the error code and is skipped (the reference raises the
 *     code and still writes, mpm_kernels.cuh:70-75);
 *   - the indenter may be the reference's z-cylinder or a sphere.
 */
#ifndef ICEMPM_MPM_ORACLE_H
#define ICEMPM_MPM_ORACLE_H

#include <stdint.h>
#include "icempm_3dm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ProxyPoint3D::getCellIndex (proxypoint3d.cpp:158-163): nearest-node key. */
void oracle_cell_keys_reference(const icempm_SimParams3D *prms, const double *soa, int pitch, int n,
                                uint32_t *keys);
/* Brick-major key used by the device sort (same cell, grouped in 4^3 bricks). */
void oracle_cell_keys_brick(const icempm_SimParams3D *prms, const double *soa, int pitch, int n,
                            uint32_t *keys);
/* std::stable_sort of indices 0..n-1 by key (the defined order of the device sort). */
void oracle_stable_sort(const uint32_t *keys, int n, int32_t *perm);
/* HostSideSOA::RemoveDisabledAndSort (host_side_soa.cpp:56-67) with a stable sort:
 * compacts enabled points (stable), then sorts by reference key. Returns new size. */
int oracle_remove_disabled_and_sort(const icempm_SimParams3D *prms, double *soa, int pitch, int n);

/* SVD in single precision by the method the reference uses (McAdams et al.
 * 2011), written from the paper; not bit-faithful (parity unpinned). With
 * oracle_set_svd_fp32(1) the return mapping uses it instead of the FP64 SVD
 * (process-wide), to measure the gap between the two. */
void oracle_set_svd_fp32(int on);
void oracle_svd3_fp32(const double F[9], double U[9], double S[3], double V[9]);

/* The damage model of the return mapping (no reference counterpart; see
 * icempm_Damage). Process-wide; NULL disables it. */
void oracle_set_damage(const icempm_Damage *damage);

/* Quadratic B-spline base node, fractional offset and weights (mpm_kernels.cuh:65-81). */
void oracle_weights(const icempm_SimParams3D *prms, const double pos[3], int base[3], double fx[3],
                    double w[9] /* w[3*i + axis] */);
/* KirchhoffStress_Wolper (mpm_kernels.cuh:547-557), row-major in/out. */
void oracle_kirchhoff(double mu, double kappa, const double F[9], double tau[9]);
/* FP64 SVD, F = U diag(S) V^T, U,V in SO(3), S[0] >= S[1] >= |S[2]|, S[2] < 0 iff det F < 0. */
void oracle_svd3(const double F[9], double U[9], double S[3], double V[9]);
/* 3x3 determinant as Eigen 3.4 evaluates it (cofactor expansion along row 0). */
double oracle_det3(const double F[9]);

/* partition_kernel_p2g (mpm_kernels.cuh:20-100). grid must be zeroed by the caller.
 * nthreads <= 1: particles in index order (deterministic). Returns error code (0 ok). */
int oracle_p2g(const icempm_SimParams3D *prms, const double *soa, int pitch, int n, double *grid,
               int nthreads);
/* partition_kernel_update_nodes (mpm_kernels.cuh:138-232) over all GX*GY*GZ nodes.
 * center: indenter position; bins: IndenterSubdivisions*GridZ*3 accumulators. */
void oracle_update_nodes(const icempm_SimParams3D *prms, const icempm_Indenter *ind,
                         const double center[3], double *grid, double *bins, int nthreads);
/* partition_kernel_g2p (mpm_kernels.cuh:237-351) without point transfer. Returns error code. */
int oracle_g2p(const icempm_SimParams3D *prms, double *soa, int pitch, int n, const double *grid,
               int record_pq, int nthreads);
/* Model3D::Step substep loop (model_3d.cpp:35-51) for one partition:
 * per substep t += dt, indenter center from the law, zero grid, P2G, update, G2P.
 * sim_time is advanced in place. Returns the max error code seen. */
int oracle_step(const icempm_SimParams3D *prms, const icempm_Indenter *ind, double *soa, int pitch,
                int n, double *grid, double *bins, int n_substeps, double *sim_time,
                int simulation_step, int nthreads);

/* Calibration adjoint (no reference counterpart). Forward: n_steps substeps
 * of oracle_step's model from soa0 (left untouched), force samples
 * F_k = (1/S) sum of the total indenter force over substeps [kS, (k+1)S),
 * S = sample_every, written to forces (3 per sample). Returns an error code. */
int oracle_adj_forward(const icempm_SimParams3D *prms, const icempm_Indenter *ind, const double *soa0, int pitch,
                       int n, int n_steps, int sample_every, double *forces);
/* Reverse mode of the same run for L with dL/dF_k given (3 per sample),
 * the plastic return mapping included: grad_mu_kappa[0..3] = dL/d(mu, kappa,
 * DP_threshold_p, DP_tan_phi). Returns 91 for liquid points. */
int oracle_adj_gradient(const icempm_SimParams3D *prms, const icempm_Indenter *ind, const double *soa0, int pitch,
                        int n, int n_steps, int sample_every, const double *dL_dF, double *grad_mu_kappa);
/* The same with a loss that also depends on the final positions: dL_dx holds
 * dL/dx_N as 3 lines of n doubles (x, y, z; point order of soa0). Either
 * dL_dF or dL_dx may be NULL (zero). */
int oracle_adj_gradient_ex(const icempm_SimParams3D *prms, const icempm_Indenter *ind, const double *soa0, int pitch,
                           int n, int n_steps, int sample_every, const double *dL_dF, const double *dL_dx,
                           double *grad_mu_kappa);

#ifdef __cplusplus
}
#endif

#endif
