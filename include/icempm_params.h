/*
 * icempm_params.h — plain-data parameter block and particle-record layout of
 * the single-partition 3D MPM step.
 *
 * icempm_SimParams3D is declared member-for-member in the order and with the
 * types of the reference's SimParams3D (simulation/parameters_sim_3d.h:21-110),
 * so that a caller holding the reference's struct can pass its address
 * unchanged (both are standard-layout; sizeof == 520 on x86-64 / gfx950).
 * The static constants of that struct (parameters_sim_3d.h:24-38) become the
 * ICEMPM_* macros below.
 */
#ifndef ICEMPM_PARAMS_H
#define ICEMPM_PARAMS_H

#ifdef __cplusplus
extern "C" {
#endif

/* SimParams3D::pi, dim, nGridArrays (parameters_sim_3d.h:24-26) */
#define ICEMPM_PI 3.14159265358979323846
#define ICEMPM_DIM 3
#define ICEMPM_N_GRID_ARRAYS 4

/* Host-side SoA line indices (parameters_sim_3d.h:29-38). Line k of a
 * buffer with capacity C holds C doubles starting at base + k*C. */
#define ICEMPM_IDX_UTILITY 0 /* int64 bit pattern stored in a double slot */
#define ICEMPM_IDX_P 1
#define ICEMPM_IDX_Q 2
#define ICEMPM_IDX_JP_INV 3
#define ICEMPM_POSX 4
#define ICEMPM_VELX 7
#define ICEMPM_FE00 10 /* row-major 3x3: Fe(i,j) at FE00 + 3*i + j */
#define ICEMPM_BP00 19 /* row-major 3x3 APIC affine matrix C */
#define ICEMPM_N_PTS_ARRAYS 28

/* utility_data bits (mpm_kernels.cuh:13-15, proxypoint3d.cpp:89-156):
 * bits 0-15 grain id (read as a signed short), 16 crushed, 17 disabled,
 * 18 liquid, 24-31 partition id. */
#define ICEMPM_STATUS_CRUSHED 0x10000LL
#define ICEMPM_STATUS_DISABLED 0x20000LL
#define ICEMPM_STATUS_LIQUID 0x40000LL

typedef struct icempm_SimParams3D {
    int SetupType; /* 0 block indentation, 1 cone compression, 2 wave tank */

    int nPtsTotal;
    int GridXTotal, GridY, GridZ;
    int IndenterSubdivisions;
    double DomainDimensionX;

    double InitialTimeStep, SimulationEndTime;
    double SimulationTime;
    int UpdateEveryNthStep;
    int SnapshotPeriod;
    int SimulationStep;

    double Gravity, Density, WaterDensity, PoissonsRatio, YoungsModulus;
    double DragGlobal;
    double lambda, mu, kappa;

    double IceCompressiveStrength, IceTensileStrength, IceShearStrength, IceTensileStrength2;
    double NACC_beta;
    double DP_tan_phi, DP_threshold_p;
    double GrainVariability;

    double IndDiameter, IndRSq, IndVelocity, IndDepth;

    double cellsize;
    double xmin, xmax, ymin, ymax, zmin, zmax;
    double BlockCutout[3];

    double ParticleVolume, ParticleMass, ParticleViewSize, LiquidParticleMass;

    double indenter_x, indenter_x_initial, indenter_y, indenter_y_initial;
    double Volume;

    int nPartitions;
    int GridHaloSize;
    int PointTransferFrequency;
    int VectorCapacity_transfer;
    double ExtraSpaceForIncomingPoints;
    double PointsTransferBufferFraction;
    double RebalanceThresholdFreeSpaceRemaining;
    double RebalanceThresholdDisabledPercentage;

    double dt_vol_Dpinv, dt_Gravity, vmax, vmax_squared;
    int gbOffset;
    double cellsize_inv, Dp_inv;
    int tpb_P2G, tpb_Upd, tpb_G2P;
    double animation_threshold_pos;
} icempm_SimParams3D;

/* Host-side restatements of the reference's SimParams3D member functions.
 * They touch no GPU and may be called before any context exists. */

/* SimParams3D::Reset (parameters_sim_3d.cpp:5-62): defaults + derived. */
void icempm_params_reset(icempm_SimParams3D *p);
/* SimParams3D::ComputeLame (parameters_sim_3d.cpp:164-169) */
void icempm_params_compute_lame(icempm_SimParams3D *p);
/* SimParams3D::ComputeCamClayParams (parameters_sim_3d.cpp:171-175) */
void icempm_params_compute_camclay(icempm_SimParams3D *p);
/* SimParams3D::ComputeHelperVariables (parameters_sim_3d.cpp:65-86) */
void icempm_params_compute_helper_variables(icempm_SimParams3D *p);
/* SimParams3D::SetPhi (parameters_sim_3d.h:100): DP_tan_phi = tan(phi deg) */
void icempm_params_set_phi(icempm_SimParams3D *p, double phi_degrees);
/* sizeof(icempm_SimParams3D), for layout checks from other languages */
int icempm_params_sizeof(void);

#ifdef __cplusplus
}
#endif

#endif /* ICEMPM_PARAMS_H */
