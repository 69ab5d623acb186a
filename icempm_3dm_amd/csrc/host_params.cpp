// host_params.cpp — restatements of the reference's SimParams3D member
// functions for icempm_SimParams3D, and the synthetic block generator.
// Host-only; no GPU is touched.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "icempm_3dm.h"

namespace {

// Truncation toward zero like a plain (int) cast, saturated at the int range
// (NaN gives 0), so that parameters read from a damaged file cannot make the
// conversion undefined. Equal to the plain cast for every in-range value.
int sat_int(double v) {
    if (!(v == v)) return 0;
    if (v >= 2147483647.0) return INT32_MAX;
    if (v <= -2147483648.0) return INT32_MIN;
    return (int)v;
}

} // namespace

extern "C" {

// SimParams3D::ComputeLame (parameters_sim_3d.cpp:164-169)
void icempm_params_compute_lame(icempm_SimParams3D *p) {
    p->lambda = p->YoungsModulus * p->PoissonsRatio / ((1 + p->PoissonsRatio) * (1 - 2 * p->PoissonsRatio));
    p->mu = p->YoungsModulus / (2 * (1 + p->PoissonsRatio));
    p->kappa = p->mu * 2. / 3. + p->lambda;
}

// SimParams3D::ComputeCamClayParams (parameters_sim_3d.cpp:171-175)
void icempm_params_compute_camclay(icempm_SimParams3D *p) {
    icempm_params_compute_lame(p);
    p->NACC_beta = p->IceTensileStrength / p->IceCompressiveStrength;
}

// SimParams3D::ComputeHelperVariables (parameters_sim_3d.cpp:65-86)
void icempm_params_compute_helper_variables(icempm_SimParams3D *p) {
    p->ParticleMass = p->ParticleVolume * p->Density;
    p->LiquidParticleMass = p->ParticleVolume * p->WaterDensity;
    p->UpdateEveryNthStep = sat_int(1.f / (200 * p->InitialTimeStep));
    p->cellsize = p->DomainDimensionX / p->GridXTotal;
    p->cellsize_inv = 1. / p->cellsize;
    p->Dp_inv = 4. / (p->cellsize * p->cellsize);
    p->IndRSq = p->IndDiameter * p->IndDiameter / 4.;
    p->dt_vol_Dpinv = p->InitialTimeStep * p->ParticleVolume * p->Dp_inv;
    p->dt_Gravity = p->InitialTimeStep * p->Gravity;
    p->vmax = 0.5 * p->cellsize / p->InitialTimeStep;
    p->vmax_squared = p->vmax * p->vmax;
    p->VectorCapacity_transfer =
        p->nPartitions > 0 ? sat_int(p->nPtsTotal / p->nPartitions * p->PointsTransferBufferFraction) : 0;
    // the reference's int product; formed in 64 bits so that absurd grid
    // sizes saturate instead of overflowing
    p->gbOffset = sat_int((double)((long long)p->GridZ * p->GridY * p->GridHaloSize));
    p->IndenterSubdivisions = sat_int(ICEMPM_PI * p->IndDiameter / p->cellsize);
}

// SimParams3D::SetPhi (parameters_sim_3d.h:100)
void icempm_params_set_phi(icempm_SimParams3D *p, double phi) { p->DP_tan_phi = std::tan(phi * ICEMPM_PI / 180.); }

// SimParams3D::Reset (parameters_sim_3d.cpp:5-62)
void icempm_params_reset(icempm_SimParams3D *p) {
    std::memset(p, 0, sizeof(*p));
    p->nPtsTotal = 0;
    p->SimulationEndTime = 12;
    p->SimulationStep = 0;
    p->SimulationTime = 0;
    p->InitialTimeStep = 3.e-5;
    p->YoungsModulus = 5.e8;
    p->GridXTotal = 128;
    p->GridY = 55;
    p->GridZ = 55;
    p->ParticleViewSize = 2.5f;
    p->DomainDimensionX = 3.33;
    p->PoissonsRatio = 0.3;
    p->Gravity = 9.81;
    p->Density = 916;
    p->WaterDensity = 1000;
    p->DragGlobal = 0;
    p->IndDiameter = 0.324;
    p->IndVelocity = 0.2;
    p->IndDepth = 0.25;
    p->IceCompressiveStrength = 100e6;
    p->IceTensileStrength = 10e6;
    p->IceShearStrength = 6e6;
    p->IceTensileStrength2 = 5e6;
    icempm_params_set_phi(p, 45);
    p->DP_threshold_p = 1e2;
    p->tpb_P2G = 256;
    p->tpb_Upd = 512;
    p->tpb_G2P = 128;
    p->indenter_x = p->indenter_x_initial = p->indenter_y = p->indenter_y_initial = 0;
    p->SetupType = 0;
    p->GrainVariability = 0.50;
    p->GridHaloSize = 15;
    p->ExtraSpaceForIncomingPoints = 0.2;
    p->PointsTransferBufferFraction = 0.05;
    p->nPartitions = 4;
    p->RebalanceThresholdFreeSpaceRemaining = 0.10;
    p->RebalanceThresholdDisabledPercentage = 0.05;
    p->PointTransferFrequency = 2;
    p->SnapshotPeriod = 100;
    p->animation_threshold_pos = 1e-3;
    icempm_params_compute_lame(p);
    icempm_params_compute_camclay(p);
    icempm_params_compute_helper_variables(p);
}

int icempm_params_sizeof(void) { return (int)sizeof(icempm_SimParams3D); }

// SplitMix64 output function (Steele, Lea & Flood 2014), used as a
// counter-based generator: draw(seed, index, k) = mix(mix(seed) + 4*index + k).
static inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline double unit_double(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

int icempm_scene_block(double *hssoa, unsigned capacity, const int cells[3], int ppc_axis, const double origin[3],
                       double cell_size, double jitter, unsigned long long seed, int n_grains, int *n_out) {
    if (!hssoa || !cells || !origin || ppc_axis <= 0 || cells[0] < 0 || cells[1] < 0 || cells[2] < 0 ||
        n_grains <= 0 || n_grains > 32767 || !(cell_size > 0))
        return ICEMPM_ERR_INVALID_ARG;
    const long long nx = (long long)cells[0] * ppc_axis, ny = (long long)cells[1] * ppc_axis,
                    nz = (long long)cells[2] * ppc_axis;
    const long long n = nx * ny * nz;
    if (n > (long long)capacity || n > 0x7FFFFFFF) return ICEMPM_ERR_CAPACITY;
    const double s = cell_size / ppc_axis;
    const uint64_t key = splitmix64((uint64_t)seed);
    for (int l = 0; l < ICEMPM_N_PTS_ARRAYS; l++) std::memset(hssoa + (size_t)l * capacity, 0, sizeof(double) * n);
    for (long long idx = 0; idx < n; idx++) {
        const long long iz = idx % nz, iy = (idx / nz) % ny, ix = idx / (nz * ny);
        const long long lat[3] = {ix, iy, iz};
        for (int d = 0; d < 3; d++) {
            double u = unit_double(splitmix64(key + 4ull * (uint64_t)idx + (uint64_t)d));
            hssoa[(size_t)(ICEMPM_POSX + d) * capacity + idx] =
                origin[d] + ((double)lat[d] + 0.5) * s + (2.0 * u - 1.0) * jitter * s;
        }
        double ug = unit_double(splitmix64(key + 4ull * (uint64_t)idx + 3ull));
        long long grain = (long long)(ug * n_grains);
        if (grain >= n_grains) grain = n_grains - 1;
        std::memcpy(&hssoa[idx], &grain, sizeof(grain));
        hssoa[(size_t)ICEMPM_IDX_JP_INV * capacity + idx] = 1.0;
        for (int d = 0; d < 3; d++) hssoa[(size_t)(ICEMPM_FE00 + 4 * d) * capacity + idx] = 1.0;
    }
    if (n_out) *n_out = (int)n;
    return ICEMPM_OK;
}

} // extern "C"
