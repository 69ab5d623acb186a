// kernels.h — host-side launchers of the HIP kernels (one definition per
// kernels_*.hip file). All launchers enqueue on `stream` and return the
// hipError_t of the launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.h"

namespace icempm {

// ---- sort (kernels_sort.hip) ----------------------------------------------
// Key of every device point: kind 0 reference nearest-node key
// (proxypoint3d.cpp:158-163), kind 1 brick key. Disabled points: 0xFFFFFFFF.
// keys[o] and vals[o] (when vals is not null) receive the key of device
// point i and i itself, o = order[i] when order is given (a permutation of
// [0, n)), else o = i.
hipError_t launch_cell_keys(const KParams &kp, PtsView pts, int n, int kind, uint32_t *keys,
                            int32_t *vals, hipStream_t stream, const int32_t *order = nullptr);
// Scratch words needed by radix_sort_pairs for n items.
size_t radix_sort_scratch_ints(int n);
// Stable LSD radix sort (8-bit digits, 4 passes) of (keys, vals); the result
// ends in keys/vals again (ping-pong through keys_tmp/vals_tmp).
hipError_t radix_sort_pairs(uint32_t *keys, int32_t *vals, uint32_t *keys_tmp, int32_t *vals_tmp,
                            int n, int32_t *scratch, hipStream_t stream);
// out line k [i] = in line k [perm[i]] for the 28 lines, slot_out[i] = slot_in[perm[i]].
hipError_t launch_permute(PtsView in, PtsView out, const int32_t *slot_in, int32_t *slot_out,
                          const int32_t *perm, int n, hipStream_t stream);
// out line k [slot[i]] = in line k [i] (back to upload order).
hipError_t launch_unpermute(PtsView in, PtsView out, const int32_t *slot, int n, hipStream_t stream);
// Brick segments of sorted brick keys: seg_start[s] (s < nseg) first point of
// the s-th non-empty brick, seg_start[nseg] = number of keyed points,
// seg_brick[s] its brick id, seg_coord[s] its brick coordinates packed
// bx | by << 8 | bz << 16, seg_slot[s] its ensemble slot (0 for one scene);
// *nseg_out on the device. The segment kernels read the coordinates and the
// slot instead of dividing the brick id by the grid's brick counts.
hipError_t launch_brick_segments(const KParams &kp, const uint32_t *sorted_keys, int n, int32_t *seg_start,
                                 int32_t *seg_brick, int32_t *seg_coord, int32_t *seg_slot, int32_t *nseg_out,
                                 int32_t *scratch, hipStream_t stream);
// *count_out = number of keys != 0xFFFFFFFF in a sorted key array.
hipError_t launch_count_valid(const uint32_t *sorted_keys, int n, int32_t *count_out, hipStream_t stream);
hipError_t launch_iota(int32_t *v, int n, hipStream_t stream);
// out[i] = in[i] (in == out allowed) and idx[i] = i.
hipError_t launch_copy_keys_iota(const uint32_t *in, uint32_t *out, int32_t *idx, int n, hipStream_t stream);
// Device-wide exclusive prefix sum of n int32 (in place allowed).
hipError_t exclusive_scan(const int32_t *in, int32_t *out, int n, int32_t *scratch, hipStream_t stream);
size_t exclusive_scan_scratch_ints(int n);

// ---- substep clock (kernels_grid.hip) --------------------------------------
// Simulation time and substep counter on the device, for step sequences that
// are captured once and replayed (a hipGraph cannot take new kernel arguments
// per replay). Two slots: the grid update of a substep reads slot `in`
// (t and step of the previous substep), uses t + dt (the host's own sequence
// of additions, so t is the same double) and writes slot `out`: t, step + 1
// and record = (step % every == 0), the flag the substep's G2P reads.
// Kernels given a clock read t / record from it instead of their t /
// record_pq arguments.
struct StepClock {
    double t;
    long long step;
    int every;
    int record;
};
hipError_t launch_clock_set(StepClock *clk, double t, long long step, int every, hipStream_t stream);

// ---- grid (kernels_grid.hip) ------------------------------------------------
struct GridState {
    double *grid;          // 4 lines of grid_pitch doubles
    int32_t *brick_flag;   // nbricks flags (1 = in a list)
    int32_t *brick_list[2];
    int32_t *brick_count;  // [2]
    // per brick of the stored grid: coordinates bx | by << 8 | bz << 16 and
    // the ensemble slot (nullptr for one scene); written once at allocation
    // (launch_brick_info), read by the grid update instead of dividing the
    // brick id by the grid's brick counts and slot sizes
    const int32_t *brick_coord;
    const int32_t *brick_slot;
};
// brick_coord[b] and, for ensembles, brick_slot[b] for every brick b of the
// stored grid.
hipError_t launch_brick_info(const KParams &kp, int32_t *brick_coord, int32_t *brick_slot, long long nbricks,
                             hipStream_t stream);
// Zeroes the bricks of list `cur`, clears their flags, zeroes count[cur^1].
hipError_t launch_reset_grid(const KParams &kp, GridState gs, int cur, hipStream_t stream);
// Zeroes the whole grid, all flags and both counts (after allocation).
hipError_t launch_clear_grid_all(const KParams &kp, GridState gs, long long nbricks, hipStream_t stream);
// Work the substep loop folds into the grid update's launch (all optional):
// the reset of another grid (launch_reset_grid of `reset` with list
// reset_cur), zeroing *zero (the G2P spill count), and the clock advance
// (clk_in -> clk_out; the update then takes t from the clock and the
// indenter center as ind_c0 + t * ind_vel).
struct UpdateExtras {
    GridState reset{};
    int reset_cur = 0;
    int *zero = nullptr;
    const StepClock *clk_in = nullptr;
    StepClock *clk_out = nullptr;
};
// partition_kernel_update_nodes over the bricks of list `cur`.
hipError_t launch_update_nodes(const KParams &kp, GridState gs, int cur, double cx, double cy, double cz, double t,
                               double *bins, const UpdateExtras &x, hipStream_t stream);
// Copy into the reference's x-major node order.
hipError_t launch_grid_to_xmajor(const KParams &kp, const double *grid, double *out, hipStream_t stream);

// ---- particle/grid transfers (kernels_p2g.hip, kernels_g2p.hip) ------------
struct SegInfo {
    const int32_t *seg_start; // nseg+1
    const int32_t *seg_brick;
    const int32_t *nseg;      // device int
    const int32_t *wg_start;  // first segment of each workgroup of the segment kernels (+1 entry), launch_work_split
    const int32_t *seg_coord; // brick coordinates bx | by << 8 | bz << 16 (at most 256 bricks per axis)
    const int32_t *seg_slot;  // ensembles: the scene slot of the brick (0 for one scene)
};
// Workgroups of a segment-kernel launch for a requested count (a multiple of
// the 8 XCDs).
inline int segment_grid(int n_blocks) { return (n_blocks + kXcd - 1) / kXcd * kXcd; }
// wg_start[b] for b in [0, nblocks]: the segment ranges of the nblocks
// workgroups, equal shares of the total cost (points + a fixed cost per
// segment). Recomputed whenever the segments or the workgroup count change.
hipError_t launch_work_split(const int32_t *seg_start, const int32_t *nseg, int32_t *wg_start, int nblocks,
                             hipStream_t stream);
hipError_t launch_p2g(const KParams &kp, PtsView pts, int n, SegInfo seg, GridState gs, int cur,
                      int *err, int n_blocks, hipStream_t stream);
// G2P including the return mapping of crushed particles. lists: 1 + n ints
// of scratch (spill count, spill list).
// src: read the point state from *src and write the new state to pts (every
// line G2P writes, and the utility line of disabled points); the default is
// in place.
hipError_t launch_g2p(const KParams &kp, PtsView pts, int n, SegInfo seg, const double *grid, int record_pq, int *err,
                      int *lists, int n_blocks, const StepClock *clk, bool count_zeroed, hipStream_t stream,
                      const PtsView *src = nullptr);
// G2P of one substep reading grid_a fused with P2G of the next substep into
// gs_b (kernels_g2p2g.hip). lists: 1 + n ints of scratch. store_vc = 0
// leaves the velocity and C lines unwritten (the next reader is a G2P).
// With clk (launch_g2p too), record_pq is clk->record. count_zeroed: the
// spill count (lists[0]) is already zero (cleared by the preceding update).
// chunk: points per chunk and threads per workgroup, 256 or 128 (128 not for
// the out-of-place passes of src).
hipError_t launch_g2p2g(const KParams &kp, PtsView pts, int n, SegInfo seg, const double *grid_a, GridState gs_b,
                        int cur_b, int record_pq, int store_vc, int *err, int *lists, int n_blocks,
                        const StepClock *clk, bool count_zeroed, hipStream_t stream,
                        const PtsView *src = nullptr, int chunk = 256);

// Reverse mode of the step for the calibration adjoint (kernels_adjoint.hip);
// grids in the step's bricked layout. theta: adjoints of (mu, kappa,
// DP_threshold_p, DP_tan_phi), added to. V: the grid after the update (m, v);
// gb: Vbar (v_bar in lines 1-3; lists marked, never reset); G: the grid after
// P2G (m, mv); Gbar: (m_bar, mv_bar) over every node.
hipError_t launch_adj_g2p_bwd(const KParams &kp, PtsView s0, PtsView s1, PtsView bar, PtsView part, int n,
                              SegInfo seg, const double *V, GridState gb, int cur_b, double *theta, bool plastic,
                              int n_blocks, hipStream_t stream);
// Vbar is read and left zeroed (every node).
hipError_t launch_adj_update_bwd(const KParams &kp, const double *G, double *Vbar, double *Gbar,
                                 const double c[3], const double fbar[3], hipStream_t stream);
hipError_t launch_adj_p2g_bwd(const KParams &kp, PtsView s0, const double *Gbar, PtsView part, PtsView bar,
                              double *theta, int n, hipStream_t stream);
// out[0..2] += scale * (sum of the indenter accumulators per component).
hipError_t launch_sum_bins(const double *bins, long long nbins, double scale, double *out, hipStream_t stream);
// `lines` point lines of n doubles between the device order and the upload
// order (slot: the upload index of each device point, a permutation of
// 0..n-1): to_upload: out[slot[i]] = in[i]; otherwise out[i] = in[slot[i]].
hipError_t launch_lines_by_slot(const double *in, const int32_t *slot, double *out, int n, int lines, bool to_upload,
                                hipStream_t stream);
// flags[0] |= 1 if a point is crushed, flags[1] |= 1 if a point is liquid.
hipError_t launch_adj_flags(PtsView st, int n, int *flags, hipStream_t stream);

} // namespace icempm
