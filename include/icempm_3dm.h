/*
 * icempm_3dm.h — C-ABI of the MI355X single-partition 3D MPM step.
 *
 * The library replaces the one-partition slice of the reference's device
 * runtime: GPU_Partition_3D (simulation/gpu_partition_3d.{h,cu}) driven by
 * GPU_Implementation6 (simulation/gpu_implementation6.{h,cpp}) and the
 * substep loop of Model3D::Step (simulation/model_3d.cpp:35-51). Each entry
 * point below names the reference member it stands in for.
 *
 * Conventions
 *  - Every entry point returns ICEMPM_OK (0) or a positive ICEMPM_ERR_* code;
 *    icempm_last_error() then returns a message (thread-local). Where the
 *    reference throws std::runtime_error, this ABI returns an error code; the
 *    C++ wrapper in icempm_gpu_implementation6.hpp turns it back into a throw.
 *  - Host buffers are HostSideSOA buffers (host_side_soa.h:51-70): 28 lines
 *    of `capacity` doubles, line k at base + k*capacity. The library never
 *    keeps a host pointer beyond a call.
 *  - A context is driven by one host thread at a time; distinct contexts are
 *    independent. All work is issued on the context's own HIP stream, in
 *    order; calls that return host data synchronize that stream.
 *  - Points come back in the order they were uploaded (HostSideSOA order),
 *    whatever order the device keeps them in.
 */
#ifndef ICEMPM_3DM_H
#define ICEMPM_3DM_H

#include <stdint.h>
#include "icempm_params.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define ICEMPM_OK 0
#define ICEMPM_ERR_INVALID_ARG 1   /* bad pointer, size or state */
#define ICEMPM_ERR_HIP 2           /* a HIP runtime call failed */
#define ICEMPM_ERR_NOT_READY 3     /* allocate/update_constants not done yet */
#define ICEMPM_ERR_CAPACITY 4      /* host or device capacity exceeded */
#define ICEMPM_ERR_DEVICE_FLAG 5   /* the device error indicator is set */
#define ICEMPM_ERR_NO_DEVICE 6     /* no HIP device is visible */

/* ---- device error indicator values -----------------------------------
 * 70..75: a particle's 3x3x3 stencil leaves the grid (mpm_kernels.cuh:70-75:
 * 70 x low, 71 y low, 72 z low, 73 x high, 74 y high, 75 z high).
 * 76: a particle position is not finite (no reference counterpart).
 * The indicator keeps the largest code raised since the last
 * update_constants / clear. Offending particles are skipped, not written. */
#define ICEMPM_DEVERR_NONFINITE 76
/* Adjoint runs only: 91 a liquid point (no adjoint of the water pressure) or
 * the damage model enabled. */
#define ICEMPM_ADJ_UNSUPPORTED 91

/* ---- indenter law (extension) -----------------------------------------
 * The reference's indenter is an infinite cylinder along z of radius
 * sqrt(IndRSq), moving with velocity (IndVelocity, 0, 0)
 * (mpm_kernels.cuh:154, :176-179; model_3d.cpp:38). A sphere of the same
 * radius is added. The center at simulation time t is
 * center_initial + t * velocity (z ignored for the cylinder). */
#define ICEMPM_INDENTER_CYLINDER_Z 0
#define ICEMPM_INDENTER_SPHERE 1
typedef struct icempm_Indenter {
    int shape;
    double center_initial[3];
    double velocity[3];
} icempm_Indenter;

/* Extension (no reference counterpart): softening damage in the return
 * mapping of crushed solid points. Each point carries a damage D in [0, 1],
 * stored as an IEEE single in bits 32-63 of its utility word (bits the
 * reference never sets and setPartition keeps, proxypoint3d.cpp:89-110; an
 * all-zero upper word is D = 0, and a value that is not a number in [0, 1]
 * reads as the nearest of 0 and 1, NaN as 0). With damage enabled the
 * cohesion of the Drucker-Prager surface is scaled by
 * f = 1 - (1 - residual_cohesion) * D: the tension cutoff pressure becomes
 * f * DP_threshold_p (also in the tension branch's Je_new) and the
 * Drucker-Prager line (p + f * DP_threshold_p) * tan(phi). After the return
 * mapping, D += (q_tr - q_new) / (3 mu) / softening_strain, capped at 1,
 * where q_tr is the trial equivalent shear stress and q_new the one after the
 * return (q_n+1 for a projection, 0 for the tension branch, q_tr when
 * inside): D grows with the equivalent plastic shear strain and reaches 1 at
 * softening_strain. The yield surface of a substep uses the damage from
 * before it. */
typedef struct icempm_Damage {
    int enabled;
    double softening_strain;  /* > 0 */
    double residual_cohesion; /* in [0, 1]: cohesion fraction left at D = 1 */
} icempm_Damage;

/* Timed phases reported by icempm_read_timings (ms accumulated). */
#define ICEMPM_PHASE_SORT 0
#define ICEMPM_PHASE_RESET 1
#define ICEMPM_PHASE_P2G 2
#define ICEMPM_PHASE_UPDATE 3
#define ICEMPM_PHASE_G2P 4
#define ICEMPM_PHASE_G2P2G 5 /* fused G2P + next P2G (icempm_set_fusion) */
#define ICEMPM_PHASE_GRAPH 6 /* one replay of a captured run of substeps (icempm_set_graph) */
#define ICEMPM_N_PHASES 7

typedef struct icempm_ctx icempm_ctx;

/* Message for the last failed call on this thread ("" if none). */
const char *icempm_last_error(void);
/* Library version string. */
const char *icempm_version(void);
/* Number of visible HIP devices (0 without a GPU). */
int icempm_get_device_count(int *count);

/* GPU_Implementation6::initialize (gpu_implementation6.cpp:319-337) with one
 * partition + GPU_Partition_3D::initialize (gpu_partition_3d.cu:52-73):
 * binds a device and creates the context's stream. */
int icempm_create(int device, icempm_ctx **out);
/* GPU_Partition_3D::~GPU_Partition_3D (gpu_partition_3d.cu:27-50). */
int icempm_destroy(icempm_ctx *ctx);

/* GPU_Implementation6::allocate_arrays (gpu_implementation6.cpp:385-401) and
 * GPU_Partition_3D::allocate (gpu_partition_3d.cu:76-132): device particle
 * storage for n_points_capacity points, the GridXTotal x GridY x GridZ grid
 * (no halo slices: one partition) and IndenterSubdivisions*GridZ*3 bins.
 * ICEMPM_ERR_CAPACITY if n_points_capacity > 2^29 - 64 (536,870,848 points,
 * 240 GB of point storage: the kernels address a point line with a 32-bit
 * byte offset). */
int icempm_allocate(icempm_ctx *ctx, const icempm_SimParams3D *prms, int n_points_capacity);

/* GPU_Partition_3D::update_constants (gpu_partition_3d.cu:135-143): takes a
 * copy of the parameters and clears the device error indicator. Unless
 * icempm_set_indenter was called, the indenter law is reset to the
 * reference cylinder: center (indenter_x_initial, indenter_y_initial, 0),
 * velocity (IndVelocity, 0, 0). */
int icempm_update_constants(icempm_ctx *ctx, const icempm_SimParams3D *prms);

/* Extension: replace the indenter law (shape, initial center, velocity). */
int icempm_set_indenter(icempm_ctx *ctx, const icempm_Indenter *ind);

/* Extension: the damage model (icempm_Damage; disabled by default), for the
 * single scene or every scene of an ensemble. ICEMPM_ERR_INVALID_ARG for a
 * softening strain that is not positive or a residual cohesion outside
 * [0, 1]. The calibration adjoint does not differentiate it:
 * icempm_adjoint_forward returns ICEMPM_ADJ_UNSUPPORTED while it is enabled. */
int icempm_set_damage(icempm_ctx *ctx, const icempm_Damage *damage);

/* GPU_Partition_3D::transfer_points_from_soa_to_device
 * (gpu_partition_3d.cu:201-215): n_points points from host lines
 * [offset, offset+n_points) become the device's points, in that order. */
int icempm_transfer_points_to_device(icempm_ctx *ctx, const double *hssoa, unsigned capacity,
                                     unsigned offset, int n_points);

/* GPU_Partition_3D::transfer_from_device (gpu_partition_3d.cu:169-199) plus
 * the checks of GPU_Implementation6::transfer_from_device
 * (gpu_implementation6.cpp:417-448): writes the points back to host lines
 * [offset, offset+n) in upload order, copies the raw indenter accumulators
 * (IndenterSubdivisions*GridZ*3 doubles) if indenter_bins is not NULL and
 * reads the error indicator. Returns ICEMPM_ERR_DEVICE_FLAG (data still
 * copied) when the indicator is set, ICEMPM_ERR_CAPACITY if
 * offset + n > capacity. */
int icempm_transfer_from_device(icempm_ctx *ctx, double *hssoa, unsigned capacity, unsigned offset,
                                int *n_points_out, double *indenter_bins);

/* GPU_Partition_3D::reset_indenter_force_accumulator (gpu_partition_3d.cu:145-150) */
int icempm_reset_indenter_force_accumulator(icempm_ctx *ctx);
/* Extension: the indenter force accumulators alone (IndenterArraySize()
 * doubles; for an ensemble all scenes' concatenated in scene order), without
 * the point copy of icempm_transfer_from_device. Synchronizes. */
int icempm_read_indenter_bins(icempm_ctx *ctx, double *indenter_bins);
/* GPU_Partition_3D::reset_grid (gpu_partition_3d.cu:222-229) */
int icempm_reset_grid(icempm_ctx *ctx);
/* GPU_Partition_3D::p2g (gpu_partition_3d.cu:231-241) */
int icempm_p2g(icempm_ctx *ctx);
/* GPU_Partition_3D::update_nodes (gpu_partition_3d.cu:243-255); the
 * indenter center is passed explicitly, as the reference passes
 * (indenter_x, indenter_y) as a kernel argument (cz used by the sphere). */
int icempm_update_nodes(icempm_ctx *ctx, double cx, double cy, double cz);
/* GPU_Partition_3D::g2p (gpu_partition_3d.cu:258-292) without point
 * transfer (one partition). record_pq != 0 writes P and Q lines. */
int icempm_g2p(icempm_ctx *ctx, int record_pq);

/* Model3D::Step substep loop (model_3d.cpp:35-51) on the device, without a
 * host synchronization per substep: for s in [0, n_substeps):
 * *simulation_time += dt; center = law(*simulation_time); reset_grid; p2g;
 * update_nodes(center); g2p((simulation_step + s) % UpdateEveryNthStep == 0).
 * If a sort period is set, points are regrouped by grid brick first on every
 * substep whose global index is a multiple of it. Asynchronous. */
int icempm_step(icempm_ctx *ctx, int n_substeps, double *simulation_time, int simulation_step);

/* Extension: with fusion on (the default), icempm_step runs the G2P of each
 * substep and the P2G of the next in one kernel over two grid buffers; off, it
 * issues the four phases per substep like the reference. Results agree to
 * rounding (summation order). */
int icempm_set_fusion(icempm_ctx *ctx, int on);

/* Extension (no reference counterpart): with graph mode on, icempm_step
 * captures each run of substeps between two brick re-sorts (the re-sort
 * included) once into a hipGraph and replays it; the simulation time and the
 * p/q record flag then come from a device-side clock that the replays
 * advance, so one graph serves every run of the same length that starts from
 * the same buffer state (at most 16 are kept; a change of parameters, point
 * count or launch settings captures anew). Results are those of the launched
 * path. Turning it off destroys the graphs. With timing on, each replay is
 * timed as ICEMPM_PHASE_GRAPH (per-phase times are then not recorded). */
int icempm_set_graph(icempm_ctx *ctx, int on);
/* Graphs captured so far and graphs currently kept. */
int icempm_graph_captures(icempm_ctx *ctx, long long *captures_out, int *cached_out);

/* GPU_Implementation6::synchronize (gpu_implementation6.cpp:513-520) */
int icempm_synchronize(icempm_ctx *ctx);
/* Reads the device error indicator (synchronizes). */
int icempm_error_code(icempm_ctx *ctx, int *code);
/* Number of points currently on the device. */
int icempm_point_count(icempm_ctx *ctx, int *n);

/* HostSideSOA::RemoveDisabledAndSort (host_side_soa.cpp:56-67) on the device:
 * disabled points are dropped and the rest are ordered by the reference key
 * ProxyPoint3D::getCellIndex (proxypoint3d.cpp:158-163) with a stable LSD
 * radix sort (ties keep their previous order). The new order becomes the
 * upload order, so the next transfer_from_device returns the sorted set. */
int icempm_remove_disabled_and_sort(icempm_ctx *ctx, int *n_points_out);
/* Regroups the device points by brick key (4^3-cell bricks, cells ordered
 * inside a brick); upload order is preserved for transfer_from_device. */
int icempm_sort_by_brick(icempm_ctx *ctx);
/* Extension: icempm_step regroups by brick every `period` substeps (0: never). */
int icempm_set_sort_period(icempm_ctx *ctx, int period);

/* Extension: counters of the step (synchronizes). out[ICEMPM_STAT_SPILLED]:
 * points the last G2P found outside their brick tile (they take the slow
 * global-memory path); out[ICEMPM_STAT_SORTS]: brick regroupings done since
 * allocation; out[ICEMPM_STAT_SEGMENTS]: brick segments of the last one;
 * out[ICEMPM_STAT_P2G_OUT_OF_TILE]: points whose stencil a P2G scatter (plain,
 * fused, or the adjoint's P2G-form scatter) found outside its segment's tile
 * and added straight to the grid in HBM, summed since allocation;
 * out[ICEMPM_STAT_G2P_SPILLED_TOTAL]: ICEMPM_STAT_SPILLED summed over every
 * G2P since allocation. Writes min(n_out, ICEMPM_N_STATS) values. */
#define ICEMPM_STAT_SPILLED 0
#define ICEMPM_STAT_SORTS 1
#define ICEMPM_STAT_SEGMENTS 2
#define ICEMPM_STAT_P2G_OUT_OF_TILE 3
#define ICEMPM_STAT_G2P_SPILLED_TOTAL 4
#define ICEMPM_N_STATS 5
int icempm_read_stats(icempm_ctx *ctx, long long *out, int n_out);

/* Keys of the device points in device order (kind 0: reference key,
 * kind 1: brick key; disabled points get 0xFFFFFFFF), copied to keys_out. */
int icempm_cell_keys(icempm_ctx *ctx, int kind, uint32_t *keys_out);
/* Device order of the points: slot_out[i] = upload index of device point i. */
int icempm_device_order(icempm_ctx *ctx, int32_t *slot_out);
/* Stable LSD radix sort of n (key, index) pairs on the device: keys and
 * perm_out are host arrays; keys_sorted_out/perm_out receive the result. */
int icempm_radix_sort_pairs(icempm_ctx *ctx, const uint32_t *keys, int n, uint32_t *keys_sorted_out,
                            int32_t *perm_out);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the sort of icempm_radix_sort_pairs on DEVICE arrays,
 * allocating nothing. keys_dev holds n keys; keys_sorted_dev receives them
 * sorted and perm_dev the stable order (keys_sorted_dev[i] =
 * keys_dev[perm_dev[i]]); keys_sorted_dev may equal keys_dev. The sort's
 * ping-pong arrays and digit counts live in one fixed device array of the
 * library per GPU (about 35 MB, set up by icempm_create), so
 * n <= ICEMPM_NOALLOC_MAX_PAIRS (larger n: ICEMPM_ERR_INVALID_ARG). The work
 * is enqueued on the context's stream and the call returns without waiting
 * (icempm_synchronize waits). Every context on one GPU shares that GPU's
 * array: calls that overlap in time on one GPU (two streams, two host
 * threads) race and must be serialised by the caller. */
#define ICEMPM_NOALLOC_MAX_PAIRS (1 << 22)
int icempm_radix_sort_pairs_noalloc(icempm_ctx *ctx, const uint32_t *keys_dev, int n, uint32_t *keys_sorted_dev,
                                    int32_t *perm_dev);
/* Copies the grid out in the reference's x-major node order
 * (x*GY*GZ + y*GZ + z), 4 lines of GX*GY*GZ doubles (m, vx|mvx, vy|mvy, vz|mvz). */
int icempm_download_grid(icempm_ctx *ctx, double *out);

/* Per-phase HIP-event timing of the kernels on the context's stream. */
int icempm_enable_timing(icempm_ctx *ctx, int on);
int icempm_reset_timings(icempm_ctx *ctx);
/* ms_out[ICEMPM_N_PHASES] accumulated ms, launches_out[ICEMPM_N_PHASES]
 * number of timed launches. Synchronizes. */
int icempm_read_timings(icempm_ctx *ctx, double *ms_out, long long *launches_out);

/* SnapshotManager::LoadRawPoints, file half (snapshotmanager.cpp:26-47): reads
 * an HDF5 raw point cloud (1-D datasets "x", "y", "z" of doubles and
 * "llGrainIDs" of 64-bit integers with a double attribute "volume"). With
 * hssoa == NULL only *n_points_out (and *volume_out) are set, so the caller can
 * size its buffer; otherwise positions go to lines 4-6 and the grain ids, as
 * stored, to line 0 of points [0, n). ICEMPM_ERR_CAPACITY if n > capacity,
 * ICEMPM_ERR_INVALID_ARG (with a message) for a missing file, dataset or
 * attribute, or an HDF5 feature outside the supported subset (version 0/1
 * superblock, version 1 object headers, symbol-table groups, contiguous or
 * chunked storage with deflate/shuffle/fletcher32 filters). Host only. */
int icempm_read_raw_points(const char *path, double *hssoa, unsigned capacity, int *n_points_out,
                           double *volume_out);
/* SnapshotManager::LoadRawPoints, geometry half (snapshotmanager.cpp:49-104)
 * on points [0, n) read by icempm_read_raw_points, prms->Volume set: block
 * boundaries into xmin..zmax, SetupType 2 liquid marking outside BlockCutout
 * (host_side_soa.cpp:33-53), the offset that centres the block in x and z with
 * its bottom at y = 2h (:20-31), a fresh state (Jp_inv = 1, Fe = I, other lines
 * zero; :88-97), the indenter start, ParticleVolume = Volume / n and
 * ComputeHelperVariables. The reference's host sort by cell
 * (RemoveDisabledAndSort) is left to icempm_remove_disabled_and_sort after
 * the upload. Host only. */
int icempm_prepare_raw_points(icempm_SimParams3D *prms, double *hssoa, unsigned capacity, int n);

/* SnapshotManager::SaveSnapshot (snapshotmanager.cpp:299-339): writes the
 * HDF5 snapshot file `path` ("d%05d.h5" in the reference) with the dataset
 * "Indenter" (IndenterSubdivisions*GridZ*3 doubles from indenter_bins) and the
 * dataset "Points" (the whole host buffer, capacity*28 doubles) carrying the
 * attributes HSSOA_capacity, HSSOA_size and the SimParams3D values of
 * SaveParametersAsAttributes (parameters_sim_3d.cpp:178-282). compress != 0
 * stores "Points" in chunks of min(2^20, size) elements, deflated at level 5,
 * as the reference does. The file is written to path.tmp and renamed. Host only. */
int icempm_save_snapshot(const char *path, const icempm_SimParams3D *prms, const double *hssoa, unsigned capacity,
                         unsigned size, const double *indenter_bins, int compress);
/* SnapshotManager::ReadSnapshot, file half (snapshotmanager.cpp:343-368, :396-399):
 * *capacity_out / *size_out from HSSOA_capacity / HSSOA_size; prms (if not
 * NULL) as SimParams3D::ReadParametersFromAttributes (parameters_sim_3d.cpp:
 * 285-372) leaves it, except that DP_threshold_p is restored too (the
 * reference opens that attribute but never reads it). With hssoa == NULL only
 * these are set; otherwise the points go to a buffer of `capacity` >=
 * HSSOA_capacity (lines re-pitched, the tail zeroed) and, if indenter_bins is
 * not NULL, the "Indenter" dataset to it (ICEMPM_ERR_CAPACITY if it holds more
 * than n_indenter_bins values). Host only. */
int icempm_read_snapshot(const char *path, icempm_SimParams3D *prms, double *hssoa, unsigned capacity,
                         unsigned *capacity_out, unsigned *size_out, double *indenter_bins, int n_indenter_bins);

/* Batched ensemble (no reference counterpart; the reference runs one scene
 * per process). n_scenes independent scenes, each with its own SimParams3D
 * (grid size, material, indenter) but a common InitialTimeStep and cellsize,
 * advance in the same kernel launches: every scene owns a slot of whole 4^3
 * node bricks in one stored grid, slots at least one empty brick apart, so no
 * point of one scene reaches a node of another. A scene's grid, walls,
 * indenter and accumulators behave as in a single-scene context allocated
 * with its parameters; a point leaving its scene's grid raises the device
 * error indicator as there. After this call the context takes
 * icempm_ensemble_transfer_points, icempm_ensemble_set_indenter,
 * icempm_step, icempm_reset_indenter_force_accumulator,
 * icempm_transfer_from_device (points back in scene-local coordinates, the
 * accumulators of all scenes concatenated in scene order) and
 * icempm_download_grid (the whole stored grid); icempm_update_constants,
 * icempm_set_indenter, icempm_transfer_points_to_device and
 * icempm_remove_disabled_and_sort return ICEMPM_ERR_INVALID_ARG. The output
 * substep for the p/q record follows scene 0's UpdateEveryNthStep.
 * ICEMPM_ERR_INVALID_ARG if the slots do not fit 1024 nodes per axis. */
int icempm_allocate_ensemble(icempm_ctx *ctx, const icempm_SimParams3D *scenes, int n_scenes, int n_points_capacity);
/* Replace scene `scene`'s indenter law; the center is in scene coordinates. */
int icempm_ensemble_set_indenter(icempm_ctx *ctx, int scene, const icempm_Indenter *ind);
/* Upload the points of every scene: scene i owns host points
 * [scene_first[i], scene_first[i+1]) (scene_first has n_scenes+1 entries,
 * scene_first[0] = 0), positions in scene coordinates. Each scene's node
 * offset inside its slot is chosen here (up to 3 nodes per axis, so that its
 * points' lowest base node starts a brick); icempm_ensemble_layout reports
 * it afterwards. */
int icempm_ensemble_transfer_points(icempm_ctx *ctx, const double *hssoa, unsigned capacity,
                                    const int *scene_first);
/* Where scene `scene` lives: node offset of its node (0,0,0) in the stored
 * grid and the offset of its IndenterSubdivisions*GridZ*3 accumulators in the
 * concatenated array. Either output may be NULL. */
int icempm_ensemble_layout(icempm_ctx *ctx, int scene, int origin_out[3], long long *bins_offset_out);

/* Autotuner (no reference counterpart). Times icempm_step for each
 * combination of G2P+P2G fusion (on/off), brick re-sort period (10, 20, 50,
 * 100, 200 substeps), workgroups per CU of the brick-segment kernels (8, 16,
 * 32) and, with fusion, points per chunk of the fused kernel (256 or 128);
 * where period 200 wins, periods 500 and then 1000 with the winning settings
 * while each gains. Each trial runs trial_substeps rounded up to whole
 * periods, each period opening with its re-sort, so the sort is paid in
 * every trial; each from the same device state,
 * then restores the points, the indenter accumulators and the error
 * indicator and keeps the fastest combination (written to *best if not
 * NULL). The context's simulation time is not advanced. Needs device
 * memory for one more copy of the points. */
/* Workgroups per CU of the brick-segment kernels (P2G, fused G2P+P2G; G2P
 * uses twice as many): 8 after icempm_create. */
int icempm_set_blocks_per_cu(icempm_ctx *ctx, int blocks_per_cu);
/* Points per chunk (= threads per workgroup) of the fused G2P+P2G kernel:
 * 256 (after icempm_create) or 128. 128-point chunks keep more of a
 * workgroup's lanes busy on brick segments of few points (spread or small
 * scenes) and fit three workgroups per CU instead of two. No reference
 * counterpart (the reference's fixed tpb_P2G/tpb_G2P, parameters_sim_3d.cpp:
 * 39-41, play this part). */
int icempm_set_points_per_chunk(icempm_ctx *ctx, int points_per_chunk);
typedef struct icempm_Tuning {
    int fuse;
    int sort_period;
    int blocks_per_cu;
    int points_per_chunk;
    double ms_per_substep;
} icempm_Tuning;
int icempm_autotune(icempm_ctx *ctx, int trial_substeps, icempm_Tuning *best);

/* Calibration adjoint (no reference counterpart: the reference has no
 * differentiation). A run of n_steps substeps from the context's current
 * points (which it leaves untouched), with the parameters and indenter law
 * of the context, producing force samples F_k: the total indenter force
 * (sum over contact nodes of (v_before - v_after) m / dt,
 * mpm_kernels.cuh:189-201) averaged over substeps [k S, (k+1) S),
 * S = sample_every. The indenter center at substep s is
 * center_initial + (t0 + (s+1) dt) velocity. checkpoint_every = K: the
 * forward keeps the point state every K substeps; the backward replays each
 * K-substep segment from its checkpoint (keeping that segment's states and
 * grids) and runs the reverse mode through it. n_steps must be a multiple
 * of S and of K. Device memory: about (n_steps/K + K + 4) * 28 * 8 bytes per
 * point plus (8 K + 8) * 8 bytes per grid node. The model differentiated is
 * the step of one scene (SetupType 0 or 1, solid points) with its plastic
 * return mapping (Wolper_Drucker_Prager, mpm_kernels.cuh:494-544); the crush
 * test and the return mapping's branch choices are held as the forward took
 * them. Liquid points end the forward with ICEMPM_ERR_INVALID_ARG. */
typedef struct icempm_AdjointRun {
    int n_steps;
    int sample_every;
    int checkpoint_every;
    double t0;
} icempm_AdjointRun;
/* Forward: writes 3 * (n_steps / sample_every) doubles to forces_out. */
int icempm_adjoint_forward(icempm_ctx *ctx, const icempm_AdjointRun *run, double *forces_out);
/* Reverse mode of the last icempm_adjoint_forward for a loss L with
 * dL/dF_k given (3 per sample, same layout as forces_out). grad_out[4]:
 * dL/d(YoungsModulus, PoissonsRatio, friction angle in degrees, cohesion),
 * with DP_tan_phi = tan(phi) and DP_threshold_p = cohesion / tan(phi); the
 * last two are exactly 0 when no point is crushed. raw_out[4] (may be NULL):
 * dL/d(mu, kappa, DP_threshold_p, DP_tan_phi). */
int icempm_adjoint_backward(icempm_ctx *ctx, const double *dL_dforces, double grad_out[4], double raw_out[4]);
/* The point positions at the end of the last icempm_adjoint_forward: 3 lines
 * of n doubles (x, y, z), points in upload order (as given to
 * icempm_transfer_points_to_device). Synchronizes. */
int icempm_adjoint_positions(icempm_ctx *ctx, double *positions_out);
/* icempm_adjoint_backward for a loss of the force samples and of those final
 * positions (a displacement misfit): dL_dpositions holds dL/dx_N in the
 * layout of icempm_adjoint_positions. Either input may be NULL (zero). */
int icempm_adjoint_backward_ex(icempm_ctx *ctx, const double *dL_dforces, const double *dL_dpositions,
                               double grad_out[4], double raw_out[4]);

/* Synthetic input (no reference counterpart; replaces the HDF5 point cloud of
 * SnapshotManager::LoadRawPoints for generated scenes): a block of
 * cells[0]*cells[1]*cells[2] lattice cells of edge `cell_size` starting at
 * `origin`, ppc_axis^3 points per cell on a sub-lattice of spacing
 * cell_size/ppc_axis, each displaced by U(-jitter, +jitter)*spacing per axis
 * and given a grain id U{0..n_grains-1}; draws come from SplitMix64 keyed by
 * (seed, point index). Writes points [0, n) of the host SoA with
 * Fe = I, Jp_inv = 1 and zeros elsewhere (HostSideSOA::InitializeBlock,
 * host_side_soa.cpp:88-97). Returns the point count in *n_out. */
int icempm_scene_block(double *hssoa, unsigned capacity, const int cells[3], int ppc_axis,
                       const double origin[3], double cell_size, double jitter,
                       unsigned long long seed, int n_grains, int *n_out);

#ifdef __cplusplus
}
#endif

#endif /* ICEMPM_3DM_H */
