// capi.hip — the C-ABI of include/icempm_3dm.h: one context per
// (device, partition), owning its stream, device particle storage, grid,
// indenter accumulators and sort scratch. Host code only; kernels live in
// kernels_*.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "icempm_3dm.h"
#include "kernels.h"

using namespace icempm;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

} // namespace

namespace icempm {
// Error reporting for the host-side translation units of the library.
int set_error(int code, const std::string &msg) { return fail(code, msg); }
} // namespace icempm

namespace {

#define ICEMPM_HIP(expr)                                                                                      \
    do {                                                                                                     \
        hipError_t _e = (expr);                                                                              \
        if (_e != hipSuccess)                                                                                \
            return fail(ICEMPM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));                  \
    } while (0)

#define ICEMPM_TRY(expr)                                                                                      \
    do {                                                                                                     \
        int _r = (expr);                                                                                     \
        if (_r != ICEMPM_OK) return _r;                                                                      \
    } while (0)

struct TimedLaunch {
    int phase;
    hipEvent_t start, stop;
};

// Everything a captured run of substeps has baked into its launches: the
// kernel parameter block, point count, launch sizes, and the host-side
// ping-pong state (grid, brick lists, point buffer) it starts from. Compared
// bytewise, so it is zero-filled before the fields are set.
struct GraphKeyFields {
    KParams kp;
    int n, p2g_blocks, g2p_blocks, fuse, chunk, sort_first, run;
    int g, cur0, cur1, cur_pts, clock_slot;
};
// A graph's key is the bytes of a GraphKeyFields whose padding is zero. It is
// kept and compared as a byte array: copies of a struct need not carry its
// padding, copies of an array of bytes do.
using GraphKey = std::array<unsigned char, sizeof(GraphKeyFields)>;

struct GraphEntry {
    GraphKey key;
    hipGraphExec_t exec;
    int g, cur0, cur1, cur_pts, clock_slot; // the ping-pong state after a replay
};

} // namespace

struct icempm_ctx {
    int device = -1;
    hipStream_t stream = nullptr;

    icempm_SimParams3D prms{};
    bool have_params = false;
    icempm_Indenter ind{};
    bool ind_custom = false;
    icempm_Damage damage{0, 1.0, 1.0};
    KParams kp{};
    double *rmc = nullptr; // kp.rmc: the return mapping's constants (device)

    // particles
    bool allocated = false;
    int capacity = 0;
    long long pitch = 0;
    int n = 0;
    double *pts[2] = {nullptr, nullptr};
    int cur_pts = 0;
    int32_t *slot[2] = {nullptr, nullptr};
    uint32_t *keys[2] = {nullptr, nullptr};
    int32_t *vals[2] = {nullptr, nullptr};
    int32_t *scratch = nullptr;
    int32_t *seg_start = nullptr;
    int32_t *seg_brick = nullptr;
    int32_t *seg_coord = nullptr; // packed brick coordinates of the segments
    int32_t *seg_slot = nullptr;  // ensemble slots of the segments
    int32_t *brick_coord = nullptr; // packed coordinates of every brick of the stored grid
    int32_t *brick_slot = nullptr;  // ensemble slot of every brick (ensembles only)
    int32_t *nseg = nullptr;
    int32_t *wg_start = nullptr; // segment ranges of the segment kernels' workgroups
    int split_blocks = 0;        // the grid wg_start was computed for (0: stale)
    int32_t *work = nullptr; // G2P worklists: 2 counts, crushed list, spill list
    int p2g_blocks = 2048;
    int g2p_blocks = 4096;
    int n_cu = 256;

    // grids: two buffers (the fused substep reads one and writes the other),
    // each with its own active-brick flags and ping-pong lists
    struct GridBuf {
        double *grid = nullptr;
        int32_t *flag = nullptr;
        int32_t *list[2] = {nullptr, nullptr};
        int32_t *count = nullptr;
        int cur = 0;
    } grids[2];
    int g = 0; // the grid the per-phase calls use, and the last one updated
    long long nbricks = 0;
    bool fuse = true;
    int chunk = 256; // points per chunk of the fused kernel (icempm_set_points_per_chunk)
    double *bins = nullptr;
    long long nbins = 0;
    int *err = nullptr;
    unsigned long long *stats = nullptr; // kStat* counters (KParams::stats)
    double *grid_xmajor = nullptr; // lazily allocated for download_grid

    int sort_period = 0;
    long long substeps_since_sort = 0;
    long long n_sorts = 0;

    // ensemble of independent scenes in one stored grid (ens.n == 0: one scene)
    struct Ensemble {
        int n = 0;
        int slot_b[3] = {0, 0, 0}, nslot[3] = {0, 0, 0};
        std::vector<icempm_SimParams3D> prms;
        std::vector<icempm_Indenter> ind; // scene-local centers
        std::vector<int> origin;          // 3 per scene, nodes of the stored grid
        std::vector<long long> bins_off;
        std::vector<int> first; // n + 1: each scene's points in upload order
        KParams *d_slots = nullptr;
        double *d_rmc = nullptr; // kRmcSize return-mapping constants per slot
    } ens;

    // calibration adjoint: checkpoints of a forward run, replayed per segment
    struct Adjoint {
        int n = 0;
        icempm_AdjointRun run{};
        int nseg = 0;
        double *ckpt = nullptr;  // nseg point states (28 lines of n)
        double *traj = nullptr;  // K + 1 point states of the segment being replayed
        double *G = nullptr;     // K grids after P2G (4 lines of grid_pitch)
        double *V = nullptr;     // K grids after the update
        double *Gbar = nullptr;  // m_bar, mv_bar (4 lines)
        double *vbar = nullptr;  // v_bar (4 lines; line 0 unused) ...
        int32_t *vbar_flag = nullptr, *vbar_list = nullptr, *vbar_count = nullptr; // ... and its brick lists
        double *bar = nullptr;   // point adjoints (28-line layout)
        double *part = nullptr;  // G2P part of the point adjoints
        double *forces = nullptr;
        double *theta = nullptr; // adjoints of mu, kappa, DP_threshold_p, DP_tan_phi (after the force samples)
        int *flags = nullptr;    // crushed / liquid points seen
        double *xfin = nullptr;  // final positions of the forward run (3 lines of n, device order)
        double *dx = nullptr;    // 3 lines of n: positions or their adjoints in upload order
        int32_t *slot = nullptr; // upload index of each device point at the forward run
        size_t x_cap = 0, dx_cap = 0, slot_cap = 0;
        std::vector<double> centers, times; // indenter center and time per substep
        size_t ckpt_cap = 0, traj_cap = 0, grid_cap = 0, pt_cap = 0, f_cap = 0;
        bool have_forward = false;
        bool any_crushed = false; // a point of the last forward run is crushed
    } adj;

    // captured runs of substeps (icempm_set_graph): the device clock that
    // gives them t and the p/q record flag, and the cache of executable graphs
    StepClock *clock = nullptr; // two slots
    int clock_slot = 0;         // the slot holding the current time
    bool graph = false;
    bool capturing = false;
    std::vector<GraphEntry> graphs;
    long long graph_captures = 0;

    // timing
    bool timing = false;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> event_pool;
    double ms[ICEMPM_N_PHASES] = {0};
    long long launches[ICEMPM_N_PHASES] = {0};

    PtsView view(int which) const { return PtsView{pts[which], pitch}; }
    SegInfo seginfo() const { return SegInfo{seg_start, seg_brick, nseg, wg_start, seg_coord, seg_slot}; }
    GridState gstate(int gi) const {
        GridState gs;
        gs.grid = grids[gi].grid;
        gs.brick_flag = grids[gi].flag;
        gs.brick_list[0] = grids[gi].list[0];
        gs.brick_list[1] = grids[gi].list[1];
        gs.brick_count = grids[gi].count;
        gs.brick_coord = brick_coord;
        gs.brick_slot = brick_slot;
        return gs;
    }
};

namespace {

int set_device(icempm_ctx *ctx) {
    ICEMPM_HIP(hipSetDevice(ctx->device));
    return ICEMPM_OK;
}

// ---------------------------------------------------------------- no-allocation sort
// Device memory of icempm_radix_sort_pairs_noalloc: keys and values of the
// ping-pong passes (ICEMPM_NOALLOC_MAX_PAIRS each), then the digit counts and
// scan partials (radix_sort_scratch_ints, checked against kNoAllocCountWords
// before use). One copy per GPU, held until the process ends. Host code never
// passes `scratch` itself to a kernel (on the host it is a stand-in, not a
// device address): it uses the per-device address from hipGetSymbolAddress.
constexpr size_t kNoAllocPairs = ICEMPM_NOALLOC_MAX_PAIRS;
constexpr size_t kNoAllocCountWords = kNoAllocPairs / 16 + 4096;
constexpr size_t kNoAllocWords = 2 * kNoAllocPairs + kNoAllocCountWords;
static __device__ __align__(256) int32_t scratch[kNoAllocWords];

constexpr int kMaxDevices = 64;
int32_t *g_noalloc_ptr[kMaxDevices] = {};
std::mutex g_noalloc_mutex;

// Looks up the array's address on `device` (current) once; icempm_create
// calls it for the device of every new context.
int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(ICEMPM_ERR_INVALID_ARG, "device index beyond the table");
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (g_noalloc_ptr[device]) return ICEMPM_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(ICEMPM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_noalloc_ptr[device] = static_cast<int32_t *>(ptr);
    return ICEMPM_OK;
}

hipEvent_t take_event(icempm_ctx *ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int resolve_timings(icempm_ctx *ctx) {
    if (ctx->pending.empty()) return ICEMPM_OK;
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    for (auto &tl : ctx->pending) {
        float msf = 0.f;
        ICEMPM_HIP(hipEventElapsedTime(&msf, tl.start, tl.stop));
        ctx->ms[tl.phase] += msf;
        ctx->launches[tl.phase] += 1;
        ctx->event_pool.push_back(tl.start);
        ctx->event_pool.push_back(tl.stop);
    }
    ctx->pending.clear();
    return ICEMPM_OK;
}

// Brackets one phase with events when timing is on.
struct PhaseTimer {
    icempm_ctx *ctx;
    int phase;
    hipEvent_t start = nullptr;
    PhaseTimer(icempm_ctx *c, int ph) : ctx(c), phase(ph) {
        if (ctx->timing && !ctx->capturing) {
            start = take_event(ctx);
            if (start && hipEventRecord(start, ctx->stream) != hipSuccess) start = nullptr;
        }
    }
    int finish() {
        if (!ctx->timing || ctx->capturing || !start) return ICEMPM_OK;
        hipEvent_t stop = take_event(ctx);
        if (!stop) return fail(ICEMPM_ERR_HIP, "hipEventCreate failed");
        ICEMPM_HIP(hipEventRecord(stop, ctx->stream));
        ctx->pending.push_back({phase, start, stop});
        if (ctx->pending.size() > 8192) return resolve_timings(ctx);
        return ICEMPM_OK;
    }
};

// The reference's indenter law for a parameter set (model_3d.cpp:38,
// mpm_kernels.cuh:154): z-cylinder from (indenter_x_initial,
// indenter_y_initial), moving along +x at IndVelocity.
icempm_Indenter reference_indenter(const icempm_SimParams3D &P) {
    icempm_Indenter ind{};
    ind.shape = ICEMPM_INDENTER_CYLINDER_Z;
    ind.center_initial[0] = P.indenter_x_initial;
    ind.center_initial[1] = P.indenter_y_initial;
    ind.center_initial[2] = 0.0;
    ind.velocity[0] = P.IndVelocity;
    return ind;
}

// Kernel parameters of one scene whose node (0,0,0) sits at node `origin` of a
// stored grid of nb bricks per axis, written into k. Every byte of k is
// cleared first, padding included: graph keys compare the block bytewise.
void scene_kparams(const icempm_SimParams3D &P, const icempm_Indenter &ind, const icempm_Damage &dmg,
                   const int origin[3], const int nb[3], KParams &k) {
    std::memset(&k, 0, sizeof(KParams));
    k.gx = P.GridXTotal;
    k.gy = P.GridY;
    k.gz = P.GridZ;
    k.ox = origin[0];
    k.oy = origin[1];
    k.oz = origin[2];
    k.nbx = nb[0];
    k.nby = nb[1];
    k.nbz = nb[2];
    k.grid_pitch = (long long)k.nbx * k.nby * k.nbz * kBrickNodes;
    k.h = P.cellsize;
    k.h_inv = P.cellsize_inv;
    k.dt = P.InitialTimeStep;
    k.mass_solid = P.ParticleMass;
    k.mass_liquid = P.LiquidParticleMass;
    k.dt_vol_Dpinv = P.dt_vol_Dpinv;
    k.dt_gravity = P.dt_Gravity;
    k.vmax = P.vmax;
    k.vmax_sq = P.vmax_squared;
    k.mu = P.mu;
    k.kappa = P.kappa;
    k.ice_comp = P.IceCompressiveStrength;
    k.ice_tens = P.IceTensileStrength;
    k.ice_shear = P.IceShearStrength;
    k.ice_tens2 = P.IceTensileStrength2;
    k.nacc_beta = P.NACC_beta;
    k.tan_phi = P.DP_tan_phi;
    k.dp_threshold_p = P.DP_threshold_p;
    k.grain_var = P.GrainVariability;
    k.rm_je_new = std::sqrt(-2. * (-P.DP_threshold_p) / P.kappa + 1.);
    k.rm_cbrt_je_new = std::cbrt(k.rm_je_new);
    k.drag = P.DragGlobal;
    k.setup_type = P.SetupType;
    k.dmg_on = dmg.enabled ? 1 : 0;
    k.dmg_rate = 1.0 / dmg.softening_strain;
    k.dmg_soft = 1.0 - dmg.residual_cohesion;
    k.ind_subdiv = P.IndenterSubdivisions;
    k.ind_rsq = P.IndRSq;
    k.ind_shape = ind.shape;
    for (int d = 0; d < 3; d++) {
        k.ind_vel[d] = ind.velocity[d];
        k.ind_c0[d] = ind.center_initial[d] + origin[d] * P.cellsize;
    }
}

// The return mapping's constants of a parameter block (KParams::rmc layout).
void rmc_values(const KParams &k, double out[kRmcSize]) {
    out[kRmcThresholdP] = k.dp_threshold_p;
    out[kRmcTanPhi] = k.tan_phi;
    out[kRmcJeNew] = k.rm_je_new;
    out[kRmcCbrtJeNew] = k.rm_cbrt_je_new;
    out[kRmcDmgRate] = k.dmg_rate;
    out[kRmcDmgSoft] = k.dmg_soft;
}

// ctx->kp from the context's parameters, indenter and damage model; once the
// device block exists, its return-mapping constants are written in stream
// order (after the work already queued, which keeps the old ones).
int fill_kparams(icempm_ctx *ctx) {
    const icempm_SimParams3D &P = ctx->prms;
    if (!ctx->ind_custom) ctx->ind = reference_indenter(P);
    const int origin[3] = {0, 0, 0};
    int nb[3] = {(P.GridXTotal + 3) / 4, (P.GridY + 3) / 4, (P.GridZ + 3) / 4};
    if (ctx->ens.n > 0)
        for (int d = 0; d < 3; d++) nb[d] = ctx->ens.slot_b[d] * ctx->ens.nslot[d];
    scene_kparams(P, ctx->ind, ctx->damage, origin, nb, ctx->kp);
    if (ctx->ens.n > 0) {
        ctx->kp.ens_n = ctx->ens.n;
        for (int d = 0; d < 3; d++) {
            ctx->kp.ens_slot_b[d] = ctx->ens.slot_b[d];
            ctx->kp.ens_nslot[d] = ctx->ens.nslot[d];
        }
        ctx->kp.ens_slots = ctx->ens.d_slots;
    }
    ctx->kp.rmc = ctx->rmc;
    ctx->kp.stats = ctx->stats;
    if (ctx->rmc) {
        double v[kRmcSize];
        rmc_values(ctx->kp, v);
        ICEMPM_HIP(hipSetDevice(ctx->device));
        ICEMPM_HIP(hipMemcpyAsync(ctx->rmc, v, sizeof(v), hipMemcpyHostToDevice, ctx->stream));
        ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    }
    return ICEMPM_OK;
}

template <typename T> int dev_alloc(T **p, size_t count, const char *what) {
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(ICEMPM_ERR_HIP, std::string("hipMalloc ") + what + " (" + std::to_string(count * sizeof(T)) +
                                        " bytes): " + hipGetErrorString(e));
    }
    return ICEMPM_OK;
}

// Destroys the cached graphs once the stream has finished their replays.
void drop_graphs(icempm_ctx *ctx) {
    if (ctx->graphs.empty()) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &e : ctx->graphs) (void)hipGraphExecDestroy(e.exec);
    ctx->graphs.clear();
}

void free_all(icempm_ctx *ctx) {
    auto f = [](void *p) {
        if (p) (void)hipFree(p);
    };
    drop_graphs(ctx);
    f(ctx->clock);
    ctx->clock = nullptr;
    for (int i = 0; i < 2; i++) {
        f(ctx->pts[i]);
        f(ctx->slot[i]);
        f(ctx->keys[i]);
        f(ctx->vals[i]);
        ctx->pts[i] = nullptr;
        ctx->slot[i] = nullptr;
        ctx->keys[i] = nullptr;
        ctx->vals[i] = nullptr;
        auto &gb = ctx->grids[i];
        f(gb.grid);
        f(gb.flag);
        f(gb.list[0]);
        f(gb.list[1]);
        f(gb.count);
        gb = icempm_ctx::GridBuf{};
    }
    f(ctx->ens.d_slots);
    ctx->ens.d_slots = nullptr;
    f(ctx->ens.d_rmc);
    ctx->ens.d_rmc = nullptr;
    f(ctx->rmc);
    ctx->rmc = nullptr;
    {
        auto &A = ctx->adj;
        // theta lives inside the forces allocation
        for (double *q : {A.ckpt, A.traj, A.G, A.V, A.Gbar, A.vbar, A.bar, A.part, A.forces, A.xfin, A.dx})
            if (q) (void)hipFree(q);
        if (A.slot) (void)hipFree(A.slot);
        for (int32_t *q : {A.vbar_flag, A.vbar_list, A.vbar_count})
            if (q) (void)hipFree(q);
        if (A.flags) (void)hipFree(A.flags);
        A = icempm_ctx::Adjoint{};
    }
    f(ctx->scratch);
    f(ctx->seg_start);
    f(ctx->seg_brick);
    f(ctx->seg_coord);
    f(ctx->seg_slot);
    f(ctx->brick_coord);
    f(ctx->brick_slot);
    f(ctx->nseg);
    f(ctx->wg_start);
    ctx->wg_start = nullptr;
    ctx->split_blocks = 0;
    f(ctx->work);
    f(ctx->bins);
    f(ctx->err);
    f(ctx->stats);
    ctx->stats = nullptr;
    f(ctx->grid_xmajor);
    ctx->scratch = ctx->seg_start = ctx->seg_brick = ctx->seg_coord = ctx->seg_slot = ctx->nseg = ctx->work = nullptr;
    ctx->brick_coord = ctx->brick_slot = nullptr;
    ctx->bins = nullptr;
    ctx->err = nullptr;
    ctx->grid_xmajor = nullptr;
    ctx->allocated = false;
}

int check_ready(icempm_ctx *ctx) {
    if (!ctx) return fail(ICEMPM_ERR_INVALID_ARG, "null context");
    if (!ctx->allocated) return fail(ICEMPM_ERR_NOT_READY, "icempm_allocate has not been called");
    if (!ctx->have_params) return fail(ICEMPM_ERR_NOT_READY, "icempm_update_constants has not been called");
    return set_device(ctx);
}

// The workgroup split of the segment kernels for the current segments and
// workgroup count (computed on the device after every change of either).
int ensure_split(icempm_ctx *ctx) {
    const int nb = segment_grid(ctx->p2g_blocks);
    if (ctx->split_blocks == nb) return ICEMPM_OK;
    ICEMPM_HIP(launch_work_split(ctx->seg_start, ctx->nseg, ctx->wg_start, nb, ctx->stream));
    ctx->split_blocks = nb;
    return ICEMPM_OK;
}

// Regroups device points by brick key and rebuilds the brick segments.
int brick_sort(icempm_ctx *ctx) {
    PhaseTimer tm(ctx, ICEMPM_PHASE_SORT);
    const int n = ctx->n;
    if (n > 0) {
        ICEMPM_HIP(launch_cell_keys(ctx->kp, ctx->view(ctx->cur_pts), n, 1, ctx->keys[0], ctx->vals[0], ctx->stream));
        ICEMPM_HIP(radix_sort_pairs(ctx->keys[0], ctx->vals[0], ctx->keys[1], ctx->vals[1], n, ctx->scratch,
                                    ctx->stream));
        const int nxt = ctx->cur_pts ^ 1;
        ICEMPM_HIP(launch_permute(ctx->view(ctx->cur_pts), ctx->view(nxt), ctx->slot[ctx->cur_pts], ctx->slot[nxt],
                                  ctx->vals[0], n, ctx->stream));
        ctx->cur_pts = nxt;
    }
    ICEMPM_HIP(launch_brick_segments(ctx->kp, ctx->keys[0], n, ctx->seg_start, ctx->seg_brick, ctx->seg_coord,
                                     ctx->seg_slot, ctx->nseg, ctx->scratch, ctx->stream));
    ctx->split_blocks = 0;
    ICEMPM_TRY(ensure_split(ctx));
    ctx->substeps_since_sort = 0;
    ctx->n_sorts++;
    return tm.finish();
}

int do_reset_grid(icempm_ctx *ctx, int gi) {
    PhaseTimer tm(ctx, ICEMPM_PHASE_RESET);
    auto &gb = ctx->grids[gi];
    ICEMPM_HIP(launch_reset_grid(ctx->kp, ctx->gstate(gi), gb.cur, ctx->stream));
    gb.cur ^= 1;
    return tm.finish();
}

int do_p2g(icempm_ctx *ctx, int gi) {
    PhaseTimer tm(ctx, ICEMPM_PHASE_P2G);
    ICEMPM_TRY(ensure_split(ctx));
    const SegInfo seg = ctx->seginfo();
    ICEMPM_HIP(launch_p2g(ctx->kp, ctx->view(ctx->cur_pts), ctx->n, seg, ctx->gstate(gi), ctx->grids[gi].cur, ctx->err,
                          ctx->p2g_blocks, ctx->stream));
    return tm.finish();
}

void indenter_center(const icempm_ctx *ctx, double t, double c[3]) {
    for (int d = 0; d < 3; d++) c[d] = ctx->ind.center_initial[d] + t * ctx->ind.velocity[d];
}

// Grid update of grid gi, with the reset of grid reset_gi (-1: none) and the
// clear of the G2P spill count in the same launch. clocked: t and the
// indenter center come from the device clock, which the launch advances
// from slot clock_slot to the other slot; otherwise from the arguments.
int do_update(icempm_ctx *ctx, int gi, double cx, double cy, double cz, double t, int reset_gi = -1,
              bool clocked = false) {
    PhaseTimer tm(ctx, ICEMPM_PHASE_UPDATE);
    UpdateExtras x;
    if (reset_gi >= 0) {
        x.reset = ctx->gstate(reset_gi);
        x.reset_cur = ctx->grids[reset_gi].cur;
    }
    x.zero = ctx->work;
    if (clocked) {
        x.clk_in = ctx->clock + ctx->clock_slot;
        x.clk_out = ctx->clock + (ctx->clock_slot ^ 1);
    }
    ICEMPM_HIP(launch_update_nodes(ctx->kp, ctx->gstate(gi), ctx->grids[gi].cur, cx, cy, cz, t, ctx->bins, x,
                                   ctx->stream));
    if (reset_gi >= 0) ctx->grids[reset_gi].cur ^= 1;
    if (clocked) ctx->clock_slot ^= 1;
    return tm.finish();
}

// G2P of the current points from grid gi. count_zeroed: the grid update just
// before cleared the spill count. clocked: the record flag is the clock's.
int do_g2p(icempm_ctx *ctx, int gi, int record_pq, bool count_zeroed = false, bool clocked = false) {
    PhaseTimer tm(ctx, ICEMPM_PHASE_G2P);
    ICEMPM_TRY(ensure_split(ctx));
    const SegInfo seg = ctx->seginfo();
    ICEMPM_HIP(launch_g2p(ctx->kp, ctx->view(ctx->cur_pts), ctx->n, seg, ctx->grids[gi].grid, record_pq, ctx->err,
                          ctx->work, ctx->g2p_blocks, clocked ? ctx->clock + ctx->clock_slot : nullptr, count_zeroed,
                          ctx->stream));
    ctx->substeps_since_sort++;
    return tm.finish();
}

// G2P of one substep from grid ga fused with P2G of the next into grid gb
// (reset, and the spill count cleared, by the grid update before it). Only
// icempm_step calls it, always followed by another fused pass or the closing
// G2P, which recompute velocity and C: those lines are not stored.
int do_g2p2g(icempm_ctx *ctx, int ga, int gb, int record_pq, bool clocked) {
    PhaseTimer tm(ctx, ICEMPM_PHASE_G2P2G);
    ICEMPM_TRY(ensure_split(ctx));
    const SegInfo seg = ctx->seginfo();
    ICEMPM_HIP(launch_g2p2g(ctx->kp, ctx->view(ctx->cur_pts), ctx->n, seg, ctx->grids[ga].grid, ctx->gstate(gb),
                            ctx->grids[gb].cur, record_pq, 0, ctx->err, ctx->work, ctx->p2g_blocks,
                            clocked ? ctx->clock + ctx->clock_slot : nullptr, true, ctx->stream, nullptr,
                            ctx->chunk));
    ctx->substeps_since_sort++;
    return tm.finish();
}

// One run of `run` substeps over the current brick segments (Model3D::Step's
// substep loop, model_3d.cpp:35-51): reset, P2G and grid update of the first
// substep, then G2P of substep k fused with P2G of substep k+1 on the other
// grid (or, unfused, run = 1), and the closing G2P. Each grid update also
// resets the grid the next fused pass scatters into. The host advances t;
// step0 is the substep number of the first substep (p/q recorded when it is
// a multiple of UpdateEveryNthStep). clocked: the launches take t and the
// record flags from the device clock instead, so that they can be captured
// once and replayed.
int step_run(icempm_ctx *ctx, int run, double &t, long long step0, bool clocked) {
    const int every = ctx->prms.UpdateEveryNthStep;
    auto rec = [&](int k) { return ((step0 + k) % every) == 0 ? 1 : 0; };
    double c[3];
    t += ctx->prms.InitialTimeStep;
    indenter_center(ctx, t, c);
    ICEMPM_TRY(do_reset_grid(ctx, ctx->g));
    ICEMPM_TRY(do_p2g(ctx, ctx->g));
    ICEMPM_TRY(do_update(ctx, ctx->g, c[0], c[1], c[2], t, run > 1 ? ctx->g ^ 1 : -1, clocked));
    for (int k = 0; k + 1 < run; k++) {
        const int nxt = ctx->g ^ 1;
        ICEMPM_TRY(do_g2p2g(ctx, ctx->g, nxt, rec(k), clocked));
        t += ctx->prms.InitialTimeStep;
        indenter_center(ctx, t, c);
        // grid g was last read by the pass above: reset it for the next one
        ICEMPM_TRY(do_update(ctx, nxt, c[0], c[1], c[2], t, k + 2 < run ? ctx->g : -1, clocked));
        ctx->g = nxt;
    }
    return do_g2p(ctx, ctx->g, rec(run - 1), true, clocked);
}

GraphKey graph_key(const icempm_ctx *ctx, bool sort_first, int run) {
    GraphKeyFields k;
    std::memset(&k, 0, sizeof(k));
    std::memcpy(&k.kp, &ctx->kp, sizeof(KParams));
    k.n = ctx->n;
    k.p2g_blocks = ctx->p2g_blocks;
    k.g2p_blocks = ctx->g2p_blocks;
    k.fuse = ctx->fuse ? 1 : 0;
    k.chunk = ctx->chunk;
    k.sort_first = sort_first ? 1 : 0;
    k.run = run;
    k.g = ctx->g;
    k.cur0 = ctx->grids[0].cur;
    k.cur1 = ctx->grids[1].cur;
    k.cur_pts = ctx->cur_pts;
    k.clock_slot = ctx->clock_slot;
    GraphKey bytes;
    std::memcpy(bytes.data(), &k, sizeof(k));
    return bytes;
}

// Captures [brick sort +] one run of substeps on the clock into an
// executable graph. The host-side state the capture walked through is put
// back; the entry records where a replay leaves it.
int capture_run(icempm_ctx *ctx, const GraphKey &key, bool sort_first, int run, GraphEntry **out) {
    if (ctx->graphs.size() >= 16) drop_graphs(ctx);
    const int g = ctx->g, cur0 = ctx->grids[0].cur, cur1 = ctx->grids[1].cur, cur_pts = ctx->cur_pts;
    const int slot = ctx->clock_slot;
    const long long since = ctx->substeps_since_sort, sorts = ctx->n_sorts;
    ICEMPM_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    double t = 0.0;
    int rc = sort_first ? brick_sort(ctx) : ICEMPM_OK;
    if (rc == ICEMPM_OK) rc = step_run(ctx, run, t, 0, true);
    ctx->capturing = false;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    GraphEntry ent;
    ent.key = key;
    ent.exec = nullptr;
    ent.g = ctx->g;
    ent.cur0 = ctx->grids[0].cur;
    ent.cur1 = ctx->grids[1].cur;
    ent.cur_pts = ctx->cur_pts;
    ent.clock_slot = ctx->clock_slot;
    ctx->g = g;
    ctx->clock_slot = slot;
    ctx->grids[0].cur = cur0;
    ctx->grids[1].cur = cur1;
    ctx->cur_pts = cur_pts;
    ctx->substeps_since_sort = since;
    ctx->n_sorts = sorts;
    if (rc != ICEMPM_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(ICEMPM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiateWithFlags(&ent.exec, graph, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(ICEMPM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    ctx->graphs.push_back(ent);
    ctx->graph_captures++;
    *out = &ctx->graphs.back();
    return ICEMPM_OK;
}

GraphEntry *find_graph(icempm_ctx *ctx, const GraphKey &key) {
    for (auto &e : ctx->graphs)
        if (e.key == key) return &e;
    return nullptr;
}

__global__ void k_fill_double(double *p, long long n, double v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_int(int *p, long long n, int v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int zero_bins(icempm_ctx *ctx) {
    if (ctx->nbins > 0) {
        k_fill_double<<<(unsigned)((ctx->nbins + 255) / 256), 256, 0, ctx->stream>>>(ctx->bins, ctx->nbins, 0.0);
        ICEMPM_HIP(hipGetLastError());
    }
    return ICEMPM_OK;
}

int clear_err(icempm_ctx *ctx) {
    k_fill_int<<<1, 1, 0, ctx->stream>>>(ctx->err, 1, 0);
    ICEMPM_HIP(hipGetLastError());
    return ICEMPM_OK;
}

} // namespace

extern "C" {

const char *icempm_last_error(void) { return g_last_error.c_str(); }

const char *icempm_version(void) { return "icempm_3dm_amd 0.1.0 (gfx950)"; }

int icempm_get_device_count(int *count) {
    if (!count) return fail(ICEMPM_ERR_INVALID_ARG, "null count");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return ICEMPM_OK;
}

int icempm_create(int device, icempm_ctx **out) {
    if (!out) return fail(ICEMPM_ERR_INVALID_ARG, "null output pointer");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(ICEMPM_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= count) return fail(ICEMPM_ERR_INVALID_ARG, "device index out of range");
    ICEMPM_HIP(hipSetDevice(device));
    ICEMPM_TRY(noalloc_setup(device));
    icempm_ctx *ctx = new icempm_ctx();
    ctx->device = device;
    hipError_t e = hipStreamCreate(&ctx->stream);
    if (e != hipSuccess) {
        delete ctx;
        return fail(ICEMPM_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->p2g_blocks = prop.multiProcessorCount * 8;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->n_cu = prop.multiProcessorCount;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->g2p_blocks = prop.multiProcessorCount * 16;
    *out = ctx;
    return ICEMPM_OK;
}

int icempm_destroy(icempm_ctx *ctx) {
    if (!ctx) return ICEMPM_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &tl : ctx->pending) {
        (void)hipEventDestroy(tl.start);
        (void)hipEventDestroy(tl.stop);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    free_all(ctx);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return ICEMPM_OK;
}

namespace {

int check_grid(const icempm_SimParams3D *prms) {
    if (prms->GridXTotal < 3 || prms->GridY < 3 || prms->GridZ < 3)
        return fail(ICEMPM_ERR_INVALID_ARG, "grid must be at least 3 nodes per axis");
    if (prms->IndenterSubdivisions < 0) return fail(ICEMPM_ERR_INVALID_ARG, "negative IndenterSubdivisions");
    if (prms->GridXTotal > 1024 || prms->GridY > 1024 || prms->GridZ > 1024)
        return fail(ICEMPM_ERR_INVALID_ARG, "grid is limited to 1024 nodes per axis");
    return ICEMPM_OK;
}

// Device storage for `prms` (the stored grid) and n_points_capacity points;
// ctx->ens must already describe the scenes of an ensemble (or be empty).
int alloc_common(icempm_ctx *ctx, const icempm_SimParams3D &prms, int n_points_capacity, long long nbins) {
    ctx->prms = prms;
    ICEMPM_TRY(dev_alloc(&ctx->rmc, kRmcSize, "return-mapping constants"));
    ICEMPM_TRY(dev_alloc(&ctx->stats, kStatCount, "step counters"));
    k_fill_double<<<1, 64, 0, ctx->stream>>>((double *)ctx->stats, kStatCount, 0.0); // all-zero bits
    ICEMPM_HIP(hipGetLastError());
    ICEMPM_TRY(fill_kparams(ctx));
    const long long nb = (long long)ctx->kp.nbx * ctx->kp.nby * ctx->kp.nbz;
    if (nb * kBrickNodes > (1ll << 40) || nb > 0x7FFFFFF)
        return fail(ICEMPM_ERR_INVALID_ARG, "grid too large");
    ctx->nbricks = nb;
    // the kernels address a point line with a 32-bit byte offset
    if ((long long)n_points_capacity > kMaxPointPitch - 64)
        return fail(ICEMPM_ERR_CAPACITY, "n_points_capacity above 2^29 - 64");
    ctx->capacity = n_points_capacity;
    ctx->pitch = ((long long)std::max(n_points_capacity, 1) + 63) / 64 * 64;
    const size_t cap = (size_t)std::max(n_points_capacity, 1);
    for (int i = 0; i < 2; i++) {
        ICEMPM_TRY(dev_alloc(&ctx->pts[i], (size_t)ctx->pitch * ICEMPM_N_PTS_ARRAYS, "points"));
        ICEMPM_TRY(dev_alloc(&ctx->slot[i], cap, "slots"));
        ICEMPM_TRY(dev_alloc(&ctx->keys[i], cap, "keys"));
        ICEMPM_TRY(dev_alloc(&ctx->vals[i], cap, "values"));
        auto &gb = ctx->grids[i];
        ICEMPM_TRY(dev_alloc(&gb.grid, (size_t)ctx->kp.grid_pitch * 4, "grid"));
        ICEMPM_TRY(dev_alloc(&gb.flag, (size_t)nb, "brick flags"));
        ICEMPM_TRY(dev_alloc(&gb.list[0], (size_t)nb, "brick list"));
        ICEMPM_TRY(dev_alloc(&gb.list[1], (size_t)nb, "brick list"));
        ICEMPM_TRY(dev_alloc(&gb.count, 2, "brick counts"));
        gb.cur = 0;
    }
    size_t scr = std::max(radix_sort_scratch_ints((int)cap), 2 * cap + 1 + exclusive_scan_scratch_ints((int)cap));
    ICEMPM_TRY(dev_alloc(&ctx->scratch, scr, "sort scratch"));
    ICEMPM_TRY(dev_alloc(&ctx->seg_start, cap + 1, "segments"));
    ICEMPM_TRY(dev_alloc(&ctx->seg_brick, cap, "segment bricks"));
    ICEMPM_TRY(dev_alloc(&ctx->seg_coord, cap, "segment brick coordinates"));
    ICEMPM_TRY(dev_alloc(&ctx->seg_slot, cap, "segment scene slots"));
    ICEMPM_TRY(dev_alloc(&ctx->nseg, 1, "segment count"));
    // at most 64 workgroups per CU (icempm_set_blocks_per_cu), rounded to the XCDs, plus one
    ICEMPM_TRY(dev_alloc(&ctx->wg_start, (size_t)segment_grid(64 * ctx->n_cu) + 1, "workgroup split"));
    ctx->split_blocks = 0;
    ICEMPM_TRY(dev_alloc(&ctx->work, 2 * cap + 2, "G2P worklists"));
    ctx->nbins = nbins;
    ICEMPM_TRY(dev_alloc(&ctx->bins, (size_t)ctx->nbins, "indenter bins"));
    ICEMPM_TRY(dev_alloc(&ctx->err, 1, "error indicator"));
    ICEMPM_TRY(dev_alloc(&ctx->clock, 2, "substep clock"));
    ICEMPM_HIP(launch_clock_set(ctx->clock, 0.0, 0, 1, ctx->stream));
    ctx->clock_slot = 0;
    ICEMPM_TRY(dev_alloc(&ctx->brick_coord, (size_t)nb, "brick coordinates"));
    if (ctx->kp.ens_n > 0) ICEMPM_TRY(dev_alloc(&ctx->brick_slot, (size_t)nb, "brick scene slots"));
    ICEMPM_HIP(launch_brick_info(ctx->kp, ctx->brick_coord, ctx->brick_slot, nb, ctx->stream));
    for (int i = 0; i < 2; i++) ICEMPM_HIP(launch_clear_grid_all(ctx->kp, ctx->gstate(i), nb, ctx->stream));
    ctx->g = 0;
    ICEMPM_TRY(zero_bins(ctx));
    ICEMPM_TRY(clear_err(ctx));
    k_fill_int<<<1, 1, 0, ctx->stream>>>(ctx->nseg, 1, 0);
    ICEMPM_HIP(hipGetLastError());
    k_fill_int<<<1, 1, 0, ctx->stream>>>(ctx->seg_start, 1, 0);
    ICEMPM_HIP(hipGetLastError());
    ctx->cur_pts = 0;
    ctx->n = 0;
    ctx->allocated = true;
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    return ICEMPM_OK;
}

} // namespace

int icempm_allocate(icempm_ctx *ctx, const icempm_SimParams3D *prms, int n_points_capacity) {
    if (!ctx || !prms) return fail(ICEMPM_ERR_INVALID_ARG, "null argument");
    if (n_points_capacity < 0) return fail(ICEMPM_ERR_INVALID_ARG, "negative capacity");
    ICEMPM_TRY(check_grid(prms));
    ICEMPM_TRY(set_device(ctx));
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    free_all(ctx);
    ctx->ens = icempm_ctx::Ensemble{};
    return alloc_common(ctx, *prms, n_points_capacity, (long long)prms->IndenterSubdivisions * prms->GridZ * 3);
}

namespace {

// Per-slot parameter blocks of an ensemble, to the device (empty slots get
// scene 0's block: no point or active brick ever lies there).
int ens_upload_slots(icempm_ctx *ctx) {
    auto &E = ctx->ens;
    const int nslots = E.nslot[0] * E.nslot[1] * E.nslot[2];
    const int nb[3] = {E.slot_b[0] * E.nslot[0], E.slot_b[1] * E.nslot[1], E.slot_b[2] * E.nslot[2]};
    std::vector<KParams> table((size_t)nslots);
    std::vector<double> rmc((size_t)nslots * kRmcSize);
    for (int k = 0; k < nslots; k++) {
        const int sc = k < E.n ? k : 0;
        KParams &e = table[(size_t)k];
        scene_kparams(E.prms[sc], E.ind[sc], ctx->damage, &E.origin[3 * sc], nb, e);
        rmc_values(e, &rmc[(size_t)k * kRmcSize]);
        e.rmc = E.d_rmc + (size_t)k * kRmcSize;
        e.bins_off = E.bins_off[sc];
        e.ens_n = E.n;
        for (int d = 0; d < 3; d++) {
            e.ens_slot_b[d] = E.slot_b[d];
            e.ens_nslot[d] = E.nslot[d];
        }
        e.ens_slots = E.d_slots;
        e.stats = ctx->stats;
    }
    ICEMPM_HIP(hipMemcpyAsync(E.d_slots, table.data(), sizeof(KParams) * nslots, hipMemcpyHostToDevice,
                              ctx->stream));
    ICEMPM_HIP(hipMemcpyAsync(E.d_rmc, rmc.data(), sizeof(double) * rmc.size(), hipMemcpyHostToDevice, ctx->stream));
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    return ICEMPM_OK;
}

} // namespace

int icempm_allocate_ensemble(icempm_ctx *ctx, const icempm_SimParams3D *scenes, int n_scenes, int n_points_capacity) {
    if (!ctx || !scenes) return fail(ICEMPM_ERR_INVALID_ARG, "null argument");
    if (n_scenes < 1 || n_scenes > 65536) return fail(ICEMPM_ERR_INVALID_ARG, "n_scenes must be in [1, 65536]");
    if (n_points_capacity < 0) return fail(ICEMPM_ERR_INVALID_ARG, "negative capacity");
    int maxg[3] = {0, 0, 0};
    for (int i = 0; i < n_scenes; i++) {
        const icempm_SimParams3D &P = scenes[i];
        ICEMPM_TRY(check_grid(&P));
        if (P.InitialTimeStep != scenes[0].InitialTimeStep || P.cellsize != scenes[0].cellsize)
            return fail(ICEMPM_ERR_INVALID_ARG, "ensemble scenes must share InitialTimeStep and cellsize");
        maxg[0] = std::max(maxg[0], P.GridXTotal);
        maxg[1] = std::max(maxg[1], P.GridY);
        maxg[2] = std::max(maxg[2], P.GridZ);
    }
    ICEMPM_TRY(set_device(ctx));
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    free_all(ctx);
    icempm_ctx::Ensemble E;
    E.n = n_scenes;
    int maxslots[3];
    for (int d = 0; d < 3; d++) {
        // at least one empty brick between scene boxes
        E.slot_b[d] = (maxg[d] + 4 + 3) / 4;
        maxslots[d] = 1024 / (E.slot_b[d] * 4);
        if (maxslots[d] < 1) return fail(ICEMPM_ERR_INVALID_ARG, "scene grid too large for an ensemble slot");
    }
    E.nslot[2] = std::min(n_scenes, maxslots[2]);
    E.nslot[1] = std::min((n_scenes + E.nslot[2] - 1) / E.nslot[2], maxslots[1]);
    E.nslot[0] = (n_scenes + E.nslot[1] * E.nslot[2] - 1) / (E.nslot[1] * E.nslot[2]);
    if (E.nslot[0] > maxslots[0])
        return fail(ICEMPM_ERR_INVALID_ARG, "too many scenes for 1024 nodes per axis of the stored grid");
    long long nbins = 0;
    for (int i = 0; i < n_scenes; i++) {
        E.prms.push_back(scenes[i]);
        E.ind.push_back(reference_indenter(scenes[i]));
        const int sl[3] = {i / (E.nslot[1] * E.nslot[2]), (i / E.nslot[2]) % E.nslot[1], i % E.nslot[2]};
        for (int d = 0; d < 3; d++) E.origin.push_back(sl[d] * E.slot_b[d] * 4);
        E.bins_off.push_back(nbins);
        nbins += (long long)scenes[i].IndenterSubdivisions * scenes[i].GridZ * 3;
    }
    ctx->ens = E;
    icempm_SimParams3D stored = scenes[0];
    stored.GridXTotal = E.slot_b[0] * E.nslot[0] * 4;
    stored.GridY = E.slot_b[1] * E.nslot[1] * 4;
    stored.GridZ = E.slot_b[2] * E.nslot[2] * 4;
    stored.DomainDimensionX = stored.GridXTotal * stored.cellsize;
    ICEMPM_TRY(alloc_common(ctx, stored, n_points_capacity, nbins));
    ICEMPM_TRY(dev_alloc(&ctx->ens.d_slots, (size_t)E.nslot[0] * E.nslot[1] * E.nslot[2], "ensemble slots"));
    ICEMPM_TRY(dev_alloc(&ctx->ens.d_rmc, (size_t)E.nslot[0] * E.nslot[1] * E.nslot[2] * kRmcSize,
                         "ensemble return-mapping constants"));
    ICEMPM_TRY(ens_upload_slots(ctx));
    ICEMPM_TRY(fill_kparams(ctx));
    ctx->have_params = true;
    return clear_err(ctx);
}

int icempm_ensemble_set_indenter(icempm_ctx *ctx, int scene, const icempm_Indenter *ind) {
    if (!ctx || !ind) return fail(ICEMPM_ERR_INVALID_ARG, "null argument");
    if (ctx->ens.n == 0) return fail(ICEMPM_ERR_NOT_READY, "icempm_allocate_ensemble has not been called");
    if (scene < 0 || scene >= ctx->ens.n) return fail(ICEMPM_ERR_INVALID_ARG, "scene out of range");
    if (ind->shape != ICEMPM_INDENTER_CYLINDER_Z && ind->shape != ICEMPM_INDENTER_SPHERE)
        return fail(ICEMPM_ERR_INVALID_ARG, "unknown indenter shape");
    ICEMPM_TRY(set_device(ctx));
    ctx->ens.ind[scene] = *ind;
    return ens_upload_slots(ctx);
}

int icempm_ensemble_layout(icempm_ctx *ctx, int scene, int origin_out[3], long long *bins_offset_out) {
    if (!ctx) return fail(ICEMPM_ERR_INVALID_ARG, "null context");
    if (ctx->ens.n == 0) return fail(ICEMPM_ERR_NOT_READY, "icempm_allocate_ensemble has not been called");
    if (scene < 0 || scene >= ctx->ens.n) return fail(ICEMPM_ERR_INVALID_ARG, "scene out of range");
    for (int d = 0; d < 3; d++)
        if (origin_out) origin_out[d] = ctx->ens.origin[3 * scene + d];
    if (bins_offset_out) *bins_offset_out = ctx->ens.bins_off[scene];
    return ICEMPM_OK;
}

int icempm_update_constants(icempm_ctx *ctx, const icempm_SimParams3D *prms) {
    if (!ctx || !prms) return fail(ICEMPM_ERR_INVALID_ARG, "null argument");
    if (!ctx->allocated) return fail(ICEMPM_ERR_NOT_READY, "icempm_allocate has not been called");
    if (ctx->ens.n > 0) return fail(ICEMPM_ERR_INVALID_ARG, "an ensemble takes its parameters at allocation");
    ICEMPM_TRY(set_device(ctx));
    const icempm_SimParams3D &old = ctx->prms;
    if (prms->GridXTotal != old.GridXTotal || prms->GridY != old.GridY || prms->GridZ != old.GridZ ||
        (long long)prms->IndenterSubdivisions * prms->GridZ * 3 != ctx->nbins)
        return fail(ICEMPM_ERR_INVALID_ARG, "grid or indenter sizes differ from the allocation");
    ctx->prms = *prms;
    ICEMPM_TRY(fill_kparams(ctx));
    ctx->have_params = true;
    return clear_err(ctx);
}

int icempm_set_indenter(icempm_ctx *ctx, const icempm_Indenter *ind) {
    if (!ctx || !ind) return fail(ICEMPM_ERR_INVALID_ARG, "null argument");
    if (ctx->ens.n > 0) return fail(ICEMPM_ERR_INVALID_ARG, "ensembles set indenters per scene");
    if (ind->shape != ICEMPM_INDENTER_CYLINDER_Z && ind->shape != ICEMPM_INDENTER_SPHERE)
        return fail(ICEMPM_ERR_INVALID_ARG, "unknown indenter shape");
    ctx->ind = *ind;
    ctx->ind_custom = true;
    return fill_kparams(ctx);
}

int icempm_set_damage(icempm_ctx *ctx, const icempm_Damage *damage) {
    if (!ctx || !damage) return fail(ICEMPM_ERR_INVALID_ARG, "null argument");
    if (damage->enabled && !(damage->softening_strain > 0))
        return fail(ICEMPM_ERR_INVALID_ARG, "softening_strain must be positive");
    if (damage->enabled && !(damage->residual_cohesion >= 0 && damage->residual_cohesion <= 1))
        return fail(ICEMPM_ERR_INVALID_ARG, "residual_cohesion must be in [0, 1]");
    ctx->damage = damage->enabled ? *damage : icempm_Damage{0, 1.0, 1.0};
    ICEMPM_TRY(fill_kparams(ctx));
    if (ctx->ens.n > 0 && ctx->ens.d_slots) {
        ICEMPM_TRY(set_device(ctx));
        ICEMPM_TRY(ens_upload_slots(ctx));
    }
    return ICEMPM_OK;
}

namespace {
int upload_points(icempm_ctx *ctx, const double *hssoa, unsigned capacity, unsigned offset, int n_points);
}

int icempm_transfer_points_to_device(icempm_ctx *ctx, const double *hssoa, unsigned capacity, unsigned offset,
                                     int n_points) {
    ICEMPM_TRY(check_ready(ctx));
    if (ctx->ens.n > 0) return fail(ICEMPM_ERR_INVALID_ARG, "ensembles upload with icempm_ensemble_transfer_points");
    return upload_points(ctx, hssoa, capacity, offset, n_points);
}

int icempm_ensemble_transfer_points(icempm_ctx *ctx, const double *hssoa, unsigned capacity,
                                    const int *scene_first) {
    ICEMPM_TRY(check_ready(ctx));
    auto &E = ctx->ens;
    if (E.n == 0) return fail(ICEMPM_ERR_NOT_READY, "icempm_allocate_ensemble has not been called");
    if (!scene_first || scene_first[0] != 0) return fail(ICEMPM_ERR_INVALID_ARG, "scene_first must start at 0");
    for (int i = 0; i < E.n; i++)
        if (scene_first[i + 1] < scene_first[i]) return fail(ICEMPM_ERR_INVALID_ARG, "scene_first must not decrease");
    const int n = scene_first[E.n];
    if ((unsigned)n > capacity) return fail(ICEMPM_ERR_CAPACITY, "scene_first exceeds the host capacity");
    if (n > 0 && !hssoa) return fail(ICEMPM_ERR_INVALID_ARG, "bad host buffer");
    // every point must start inside its own scene's grid box: a point placed
    // in another scene's slot would be advanced as a point of that scene
    for (int i = 0; i < E.n; i++) {
        const double ext[3] = {E.prms[i].GridXTotal * E.prms[i].cellsize, E.prms[i].GridY * E.prms[i].cellsize,
                               E.prms[i].GridZ * E.prms[i].cellsize};
        for (int d = 0; d < 3; d++) {
            const double *line = hssoa + (size_t)(ICEMPM_POSX + d) * capacity;
            for (int k = scene_first[i]; k < scene_first[i + 1]; k++)
                if (!(line[k] >= 0.0 && line[k] <= ext[d]))
                    return fail(ICEMPM_ERR_INVALID_ARG, "point " + std::to_string(k) + " lies outside the grid of scene " +
                                                            std::to_string(i));
        }
    }
    // Each scene's node offset inside its slot: the slot is whole bricks and
    // leaves at least 4 spare nodes per axis, so the scene box can move by
    // up to 3 nodes and still own every brick it touches. The offset puts the
    // lowest base node of the scene's points on a brick boundary, which
    // minimises the bricks (and so the segments) the block occupies.
    for (int i = 0; i < E.n; i++) {
        const int sl[3] = {i / (E.nslot[1] * E.nslot[2]), (i / E.nslot[2]) % E.nslot[1], i % E.nslot[2]};
        const int g[3] = {E.prms[i].GridXTotal, E.prms[i].GridY, E.prms[i].GridZ};
        for (int d = 0; d < 3; d++) {
            const double *line = hssoa + (size_t)(ICEMPM_POSX + d) * capacity;
            long long lo = 0;
            bool any = false;
            for (int k = scene_first[i]; k < scene_first[i + 1]; k++) {
                const long long b = (long long)std::floor(line[k] * E.prms[i].cellsize_inv - 0.5);
                if (!any || b < lo) lo = b;
                any = true;
            }
            int shift = any ? (int)(((-lo) % 4 + 4) % 4) : 0;
            if (shift > E.slot_b[d] * 4 - g[d]) shift = 0;
            E.origin[3 * i + d] = sl[d] * E.slot_b[d] * 4 + shift;
        }
    }
    ICEMPM_TRY(ens_upload_slots(ctx));
    // scene-local positions -> stored-grid positions
    std::vector<double> tmp((size_t)std::max(n, 1) * ICEMPM_N_PTS_ARRAYS);
    for (int l = 0; l < ICEMPM_N_PTS_ARRAYS; l++)
        std::memcpy(tmp.data() + (size_t)l * n, hssoa + (size_t)l * capacity, sizeof(double) * n);
    for (int i = 0; i < E.n; i++)
        for (int d = 0; d < 3; d++) {
            const double off = E.origin[3 * i + d] * E.prms[i].cellsize;
            double *line = tmp.data() + (size_t)(ICEMPM_POSX + d) * n;
            for (int k = scene_first[i]; k < scene_first[i + 1]; k++) line[k] += off;
        }
    E.first.assign(scene_first, scene_first + E.n + 1);
    return upload_points(ctx, tmp.data(), (unsigned)n, 0, n);
}

namespace {
int upload_points(icempm_ctx *ctx, const double *hssoa, unsigned capacity, unsigned offset, int n_points) {
    if (n_points < 0 || (n_points > 0 && !hssoa)) return fail(ICEMPM_ERR_INVALID_ARG, "bad host buffer");
    if (n_points > ctx->capacity) return fail(ICEMPM_ERR_CAPACITY, "more points than the device capacity");
    if ((unsigned long long)offset + (unsigned)n_points > capacity)
        return fail(ICEMPM_ERR_CAPACITY, "offset + n_points exceeds the host capacity");
    const int cur = ctx->cur_pts;
    for (int l = 0; l < ICEMPM_N_PTS_ARRAYS && n_points > 0; l++) {
        ICEMPM_HIP(hipMemcpyAsync(ctx->pts[cur] + (size_t)l * ctx->pitch, hssoa + (size_t)l * capacity + offset,
                                  sizeof(double) * n_points, hipMemcpyHostToDevice, ctx->stream));
    }
    ICEMPM_HIP(launch_iota(ctx->slot[cur], n_points, ctx->stream));
    ctx->n = n_points;
    ICEMPM_TRY(brick_sort(ctx));
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    return ICEMPM_OK;
}
} // namespace

int icempm_transfer_from_device(icempm_ctx *ctx, double *hssoa, unsigned capacity, unsigned offset,
                                int *n_points_out, double *indenter_bins) {
    ICEMPM_TRY(check_ready(ctx));
    const int n = ctx->n;
    if (n > 0 && !hssoa) return fail(ICEMPM_ERR_INVALID_ARG, "null host buffer");
    if ((unsigned long long)offset + (unsigned)n > capacity)
        return fail(ICEMPM_ERR_CAPACITY, "transfer_from_device: host capacity exceeded");
    const int cur = ctx->cur_pts, spare = cur ^ 1;
    ICEMPM_HIP(launch_unpermute(ctx->view(cur), ctx->view(spare), ctx->slot[cur], n, ctx->stream));
    for (int l = 0; l < ICEMPM_N_PTS_ARRAYS && n > 0; l++) {
        ICEMPM_HIP(hipMemcpyAsync(hssoa + (size_t)l * capacity + offset, ctx->pts[spare] + (size_t)l * ctx->pitch,
                                  sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    }
    if (indenter_bins && ctx->nbins > 0)
        ICEMPM_HIP(hipMemcpyAsync(indenter_bins, ctx->bins, sizeof(double) * ctx->nbins, hipMemcpyDeviceToHost,
                                  ctx->stream));
    int code = 0;
    ICEMPM_HIP(hipMemcpyAsync(&code, ctx->err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    const auto &E = ctx->ens;
    if (E.n > 0 && (int)E.first.size() == E.n + 1 && E.first[E.n] == n) {
        for (int i = 0; i < E.n; i++)
            for (int d = 0; d < 3; d++) {
                const double off = E.origin[3 * i + d] * E.prms[i].cellsize;
                double *line = hssoa + (size_t)(ICEMPM_POSX + d) * capacity + offset;
                for (int k = E.first[i]; k < E.first[i + 1]; k++) line[k] -= off;
            }
    }
    if (n_points_out) *n_points_out = n;
    if (code) return fail(ICEMPM_ERR_DEVICE_FLAG, "device error indicator " + std::to_string(code));
    return ICEMPM_OK;
}

int icempm_reset_indenter_force_accumulator(icempm_ctx *ctx) {
    ICEMPM_TRY(check_ready(ctx));
    return zero_bins(ctx);
}

int icempm_read_indenter_bins(icempm_ctx *ctx, double *indenter_bins) {
    ICEMPM_TRY(check_ready(ctx));
    if (!indenter_bins) return fail(ICEMPM_ERR_INVALID_ARG, "null output");
    if (ctx->nbins > 0)
        ICEMPM_HIP(hipMemcpyAsync(indenter_bins, ctx->bins, sizeof(double) * ctx->nbins, hipMemcpyDeviceToHost,
                                  ctx->stream));
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    return ICEMPM_OK;
}

int icempm_reset_grid(icempm_ctx *ctx) {
    ICEMPM_TRY(check_ready(ctx));
    return do_reset_grid(ctx, ctx->g);
}

int icempm_p2g(icempm_ctx *ctx) {
    ICEMPM_TRY(check_ready(ctx));
    return do_p2g(ctx, ctx->g);
}

int icempm_update_nodes(icempm_ctx *ctx, double cx, double cy, double cz) {
    ICEMPM_TRY(check_ready(ctx));
    return do_update(ctx, ctx->g, cx, cy, cz, 0.0);
}

int icempm_g2p(icempm_ctx *ctx, int record_pq) {
    ICEMPM_TRY(check_ready(ctx));
    return do_g2p(ctx, ctx->g, record_pq);
}

int icempm_step(icempm_ctx *ctx, int n_substeps, double *simulation_time, int simulation_step) {
    ICEMPM_TRY(check_ready(ctx));
    if (!simulation_time || n_substeps < 0) return fail(ICEMPM_ERR_INVALID_ARG, "bad step arguments");
    const int every = ctx->prms.UpdateEveryNthStep;
    if (every <= 0) return fail(ICEMPM_ERR_INVALID_ARG, "UpdateEveryNthStep must be positive");
    double t = *simulation_time;
    if (ctx->graph && n_substeps > 0) {
        ICEMPM_HIP(launch_clock_set(ctx->clock, t, simulation_step, every, ctx->stream));
        ctx->clock_slot = 0;
    }
    int s = 0;
    while (s < n_substeps) {
        const bool sort_now = ctx->sort_period > 0 && ctx->substeps_since_sort >= ctx->sort_period;
        // substeps until the next regrouping
        int run = n_substeps - s;
        if (ctx->sort_period > 0)
            run = (int)std::min<long long>(run, ctx->sort_period - (sort_now ? 0 : ctx->substeps_since_sort));
        if (!ctx->fuse) run = 1;
        if (ctx->graph) {
            const GraphKey key = graph_key(ctx, sort_now, run);
            GraphEntry *ent = find_graph(ctx, key);
            if (!ent) ICEMPM_TRY(capture_run(ctx, key, sort_now, run, &ent));
            PhaseTimer tm(ctx, ICEMPM_PHASE_GRAPH);
            ICEMPM_HIP(hipGraphLaunch(ent->exec, ctx->stream));
            ctx->g = ent->g;
            ctx->grids[0].cur = ent->cur0;
            ctx->grids[1].cur = ent->cur1;
            ctx->cur_pts = ent->cur_pts;
            ctx->clock_slot = ent->clock_slot;
            if (sort_now) {
                ctx->substeps_since_sort = 0;
                ctx->n_sorts++;
            }
            ctx->substeps_since_sort += run;
            for (int k = 0; k < run; k++) t += ctx->prms.InitialTimeStep; // the clock's additions
            ICEMPM_TRY(tm.finish());
        } else {
            if (sort_now) ICEMPM_TRY(brick_sort(ctx));
            ICEMPM_TRY(step_run(ctx, run, t, (long long)simulation_step + s, false));
        }
        s += run;
    }
    *simulation_time = t;
    return ICEMPM_OK;
}

int icempm_set_graph(icempm_ctx *ctx, int on) {
    if (!ctx) return fail(ICEMPM_ERR_INVALID_ARG, "null context");
    ctx->graph = on != 0;
    if (!ctx->graph) {
        ICEMPM_TRY(set_device(ctx));
        ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
        drop_graphs(ctx);
    }
    return ICEMPM_OK;
}

int icempm_graph_captures(icempm_ctx *ctx, long long *captures_out, int *cached_out) {
    if (!ctx) return fail(ICEMPM_ERR_INVALID_ARG, "null context");
    if (captures_out) *captures_out = ctx->graph_captures;
    if (cached_out) *cached_out = (int)ctx->graphs.size();
    return ICEMPM_OK;
}

int icempm_set_fusion(icempm_ctx *ctx, int on) {
    if (!ctx) return fail(ICEMPM_ERR_INVALID_ARG, "null context");
    ctx->fuse = on != 0;
    return ICEMPM_OK;
}

int icempm_synchronize(icempm_ctx *ctx) {
    if (!ctx) return fail(ICEMPM_ERR_INVALID_ARG, "null context");
    ICEMPM_TRY(set_device(ctx));
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    return ICEMPM_OK;
}

int icempm_error_code(icempm_ctx *ctx, int *code) {
    ICEMPM_TRY(check_ready(ctx));
    if (!code) return fail(ICEMPM_ERR_INVALID_ARG, "null code");
    int c = 0;
    ICEMPM_HIP(hipMemcpyAsync(&c, ctx->err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    *code = c;
    return ICEMPM_OK;
}

int icempm_point_count(icempm_ctx *ctx, int *n) {
    if (!ctx || !n) return fail(ICEMPM_ERR_INVALID_ARG, "null argument");
    *n = ctx->n;
    return ICEMPM_OK;
}

int icempm_remove_disabled_and_sort(icempm_ctx *ctx, int *n_points_out) {
    ICEMPM_TRY(check_ready(ctx));
    if (ctx->ens.n > 0) return fail(ICEMPM_ERR_INVALID_ARG, "an ensemble keeps its scenes' point ranges");
    const int n = ctx->n;
    if (n > 0) {
        PhaseTimer tm(ctx, ICEMPM_PHASE_SORT);
        // Keys in upload order (the host buffer's order: slot maps device
        // point -> upload index), so that the stable sort breaks ties between
        // points of one cell by their place in the host buffer, as a stable
        // sort of the host buffer does, whatever order the brick sorts left
        // the device points in.
        ICEMPM_HIP(launch_cell_keys(ctx->kp, ctx->view(ctx->cur_pts), n, 0, ctx->keys[0], ctx->vals[0], ctx->stream,
                                    ctx->slot[ctx->cur_pts]));
        ICEMPM_HIP(radix_sort_pairs(ctx->keys[0], ctx->vals[0], ctx->keys[1], ctx->vals[1], n, ctx->scratch,
                                    ctx->stream));
        const int nxt = ctx->cur_pts ^ 1;
        ICEMPM_HIP(launch_permute(ctx->view(ctx->cur_pts), ctx->view(nxt), ctx->slot[ctx->cur_pts], ctx->slot[nxt],
                                  ctx->vals[0], n, ctx->stream));
        ctx->cur_pts = nxt;
        ICEMPM_HIP(launch_count_valid(ctx->keys[0], n, ctx->nseg, ctx->stream));
        int n_new = 0;
        ICEMPM_HIP(hipMemcpyAsync(&n_new, ctx->nseg, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
        ICEMPM_TRY(tm.finish());
        ctx->n = n_new;
        // the sorted set becomes the upload order
        ICEMPM_HIP(launch_iota(ctx->slot[ctx->cur_pts], ctx->n, ctx->stream));
        ICEMPM_TRY(brick_sort(ctx));
    }
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    if (n_points_out) *n_points_out = ctx->n;
    return ICEMPM_OK;
}

int icempm_sort_by_brick(icempm_ctx *ctx) {
    ICEMPM_TRY(check_ready(ctx));
    return brick_sort(ctx);
}

int icempm_autotune(icempm_ctx *ctx, int trial_substeps, icempm_Tuning *best) {
    ICEMPM_TRY(check_ready(ctx));
    if (trial_substeps < 1) return fail(ICEMPM_ERR_INVALID_ARG, "trial_substeps must be positive");
    const int n = ctx->n;
    icempm_Tuning keep{ctx->fuse ? 1 : 0, ctx->sort_period, std::max(1, ctx->p2g_blocks / std::max(ctx->n_cu, 1)),
                       ctx->chunk, 0.0};
    if (n == 0) {
        if (best) *best = keep;
        return ICEMPM_OK;
    }
    // the state every trial starts from
    const size_t lines = (size_t)ICEMPM_N_PTS_ARRAYS * ctx->pitch;
    double *save_pts = nullptr, *save_bins = nullptr;
    int32_t *save_slot = nullptr;
    int save_err = 0;
    auto release = [&]() {
        if (save_pts) (void)hipFree(save_pts);
        if (save_bins) (void)hipFree(save_bins);
        if (save_slot) (void)hipFree(save_slot);
    };
    int rc = dev_alloc(&save_pts, lines, "autotune point copy");
    if (rc == ICEMPM_OK) rc = dev_alloc(&save_slot, (size_t)n, "autotune slot copy");
    if (rc == ICEMPM_OK) rc = dev_alloc(&save_bins, (size_t)std::max(ctx->nbins, 1LL), "autotune bins copy");
    if (rc != ICEMPM_OK) {
        release();
        return rc;
    }
    const int cur = ctx->cur_pts;
    auto copy = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream);
    };
    hipError_t e = copy(save_pts, ctx->pts[cur], sizeof(double) * lines);
    if (e == hipSuccess) e = copy(save_slot, ctx->slot[cur], sizeof(int32_t) * n);
    if (e == hipSuccess && ctx->nbins > 0) e = copy(save_bins, ctx->bins, sizeof(double) * ctx->nbins);
    if (e == hipSuccess) e = hipMemcpyAsync(&save_err, ctx->err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        release();
        return fail(ICEMPM_ERR_HIP, std::string("autotune save: ") + hipGetErrorString(e));
    }
    // trials launch directly: a capture per trial would be timed with it
    const bool timing = ctx->timing, graph = ctx->graph;
    ctx->timing = false;
    ctx->graph = false;
    auto restore = [&]() -> int {
        ctx->cur_pts = cur;
        ICEMPM_HIP(copy(ctx->pts[cur], save_pts, sizeof(double) * lines));
        ICEMPM_HIP(copy(ctx->slot[cur], save_slot, sizeof(int32_t) * n));
        if (ctx->nbins > 0) ICEMPM_HIP(copy(ctx->bins, save_bins, sizeof(double) * ctx->nbins));
        k_fill_int<<<1, 1, 0, ctx->stream>>>(ctx->err, 1, save_err);
        ICEMPM_HIP(hipGetLastError());
        return ICEMPM_OK;
    };
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    (void)hipEventCreate(&ev0);
    (void)hipEventCreate(&ev1);
    icempm_Tuning top{keep.fuse, keep.sort_period, keep.blocks_per_cu, keep.points_per_chunk, 1e300};
    // one trial: ms per substep of whole periods from the saved state, each
    // period opening with its re-sort (the sort is amortised over the period)
    auto trial = [&](int fuse, int chunk, int period, int bpc, double *ms_per_substep) -> int {
        ctx->fuse = fuse != 0;
        ctx->chunk = chunk;
        ctx->sort_period = period;
        ctx->p2g_blocks = bpc * ctx->n_cu;
        ctx->g2p_blocks = 2 * bpc * ctx->n_cu;
        ICEMPM_TRY(restore());
        ICEMPM_TRY(brick_sort(ctx));
        double t = 0.0;
        ICEMPM_TRY(icempm_step(ctx, 2, &t, 0));
        const int len = (trial_substeps + period - 1) / period * period;
        ctx->substeps_since_sort = period;
        float ms = 0.f;
        if (hipEventRecord(ev0, ctx->stream) != hipSuccess) return fail(ICEMPM_ERR_HIP, "hipEventRecord");
        ICEMPM_TRY(icempm_step(ctx, len, &t, 2));
        if (hipEventRecord(ev1, ctx->stream) != hipSuccess) return fail(ICEMPM_ERR_HIP, "hipEventRecord");
        if (hipEventSynchronize(ev1) != hipSuccess || hipEventElapsedTime(&ms, ev0, ev1) != hipSuccess)
            return fail(ICEMPM_ERR_HIP, "autotune timing");
        *ms_per_substep = (double)ms / len;
        return ICEMPM_OK;
    };
    rc = ICEMPM_OK;
    for (int fuse = 1; fuse >= 0 && rc == ICEMPM_OK; fuse--)
        for (int chunk : {256, 128})
            for (int period : {10, 20, 50, 100, 200})
            for (int bpc : {8, 16, 32}) {
                if (rc != ICEMPM_OK) break;
                if (!fuse && chunk != 256) continue; // the chunk size is the fused kernel's
                double ms = 0.0;
                rc = trial(fuse, chunk, period, bpc, &ms);
                if (rc == ICEMPM_OK && ms < top.ms_per_substep) top = icempm_Tuning{fuse, period, bpc, chunk, ms};
            }
    // Where the longest period tried wins (points that keep to their tiles),
    // longer ones are tried with the winning settings, 500 and then 1000
    // substeps, as long as each one gains.
    for (int period : {500, 1000}) {
        if (rc != ICEMPM_OK || top.sort_period != (period == 500 ? 200 : 500)) break;
        double ms = 0.0;
        rc = trial(top.fuse, top.points_per_chunk, period, top.blocks_per_cu, &ms);
        if (rc == ICEMPM_OK && ms < top.ms_per_substep)
            top = icempm_Tuning{top.fuse, period, top.blocks_per_cu, top.points_per_chunk, ms};
    }
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    ctx->timing = timing;
    ctx->graph = graph;
    const icempm_Tuning use = rc == ICEMPM_OK ? top : keep;
    ctx->fuse = use.fuse != 0;
    ctx->chunk = use.points_per_chunk;
    ctx->sort_period = use.sort_period;
    ctx->p2g_blocks = use.blocks_per_cu * ctx->n_cu;
    ctx->g2p_blocks = 2 * use.blocks_per_cu * ctx->n_cu;
    int rc2 = restore();
    if (rc2 == ICEMPM_OK) rc2 = brick_sort(ctx);
    if (rc2 == ICEMPM_OK) {
        hipError_t e2 = hipStreamSynchronize(ctx->stream);
        if (e2 != hipSuccess) rc2 = fail(ICEMPM_ERR_HIP, std::string("autotune restore: ") + hipGetErrorString(e2));
    }
    release();
    if (rc != ICEMPM_OK) return rc;
    if (rc2 != ICEMPM_OK) return rc2;
    ctx->substeps_since_sort = 0;
    if (best) *best = use;
    return ICEMPM_OK;
}

int icempm_set_points_per_chunk(icempm_ctx *ctx, int points_per_chunk) {
    if (!ctx || (points_per_chunk != 128 && points_per_chunk != 256))
        return fail(ICEMPM_ERR_INVALID_ARG, "points_per_chunk must be 128 or 256");
    ctx->chunk = points_per_chunk;
    return ICEMPM_OK;
}

int icempm_set_blocks_per_cu(icempm_ctx *ctx, int blocks_per_cu) {
    if (!ctx || blocks_per_cu < 1 || blocks_per_cu > 64)
        return fail(ICEMPM_ERR_INVALID_ARG, "blocks_per_cu must be in [1, 64]");
    ctx->p2g_blocks = blocks_per_cu * ctx->n_cu;
    ctx->g2p_blocks = 2 * blocks_per_cu * ctx->n_cu;
    return ICEMPM_OK;
}

int icempm_set_sort_period(icempm_ctx *ctx, int period) {
    if (!ctx || period < 0) return fail(ICEMPM_ERR_INVALID_ARG, "bad sort period");
    ctx->sort_period = period;
    return ICEMPM_OK;
}

int icempm_read_stats(icempm_ctx *ctx, long long *out, int n_out) {
    ICEMPM_TRY(check_ready(ctx));
    if (!out || n_out < 0) return fail(ICEMPM_ERR_INVALID_ARG, "bad stats buffer");
    int dev[2] = {0, 0};
    unsigned long long cum[kStatCount] = {0, 0};
    ICEMPM_HIP(hipMemcpyAsync(&dev[0], ctx->work, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ICEMPM_HIP(hipMemcpyAsync(&dev[1], ctx->nseg, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ICEMPM_HIP(hipMemcpyAsync(cum, ctx->stats, sizeof(cum), hipMemcpyDeviceToHost, ctx->stream));
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    static_assert(ICEMPM_N_STATS == 5, "stats layout");
    const long long v[ICEMPM_N_STATS] = {dev[0], ctx->n_sorts, dev[1], (long long)cum[kStatP2GOutOfTile],
                                         (long long)cum[kStatG2PSpilled]};
    for (int i = 0; i < n_out && i < ICEMPM_N_STATS; i++) out[i] = v[i];
    return ICEMPM_OK;
}

int icempm_cell_keys(icempm_ctx *ctx, int kind, uint32_t *keys_out) {
    ICEMPM_TRY(check_ready(ctx));
    if (kind != 0 && kind != 1) return fail(ICEMPM_ERR_INVALID_ARG, "kind must be 0 or 1");
    if (ctx->n > 0 && !keys_out) return fail(ICEMPM_ERR_INVALID_ARG, "null output");
    if (ctx->n == 0) return ICEMPM_OK;
    // keys[1] is free outside a sort
    ICEMPM_HIP(launch_cell_keys(ctx->kp, ctx->view(ctx->cur_pts), ctx->n, kind, ctx->keys[1], nullptr, ctx->stream));
    ICEMPM_HIP(hipMemcpyAsync(keys_out, ctx->keys[1], sizeof(uint32_t) * ctx->n, hipMemcpyDeviceToHost, ctx->stream));
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    return ICEMPM_OK;
}

int icempm_device_order(icempm_ctx *ctx, int32_t *slot_out) {
    ICEMPM_TRY(check_ready(ctx));
    if (ctx->n > 0 && !slot_out) return fail(ICEMPM_ERR_INVALID_ARG, "null output");
    if (ctx->n == 0) return ICEMPM_OK;
    ICEMPM_HIP(hipMemcpyAsync(slot_out, ctx->slot[ctx->cur_pts], sizeof(int32_t) * ctx->n, hipMemcpyDeviceToHost,
                              ctx->stream));
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    return ICEMPM_OK;
}

int icempm_radix_sort_pairs(icempm_ctx *ctx, const uint32_t *keys, int n, uint32_t *keys_sorted_out,
                            int32_t *perm_out) {
    if (!ctx) return fail(ICEMPM_ERR_INVALID_ARG, "null context");
    if (n < 0 || (n > 0 && (!keys || !keys_sorted_out || !perm_out)))
        return fail(ICEMPM_ERR_INVALID_ARG, "bad arguments");
    ICEMPM_TRY(set_device(ctx));
    if (n == 0) return ICEMPM_OK;
    uint32_t *k0 = nullptr, *k1 = nullptr;
    int32_t *v0 = nullptr, *v1 = nullptr, *scr = nullptr;
    int rc = ICEMPM_OK;
    auto cleanup = [&]() {
        (void)hipFree(k0);
        (void)hipFree(k1);
        (void)hipFree(v0);
        (void)hipFree(v1);
        (void)hipFree(scr);
    };
    if ((rc = dev_alloc(&k0, n, "keys")) || (rc = dev_alloc(&k1, n, "keys")) || (rc = dev_alloc(&v0, n, "vals")) ||
        (rc = dev_alloc(&v1, n, "vals")) || (rc = dev_alloc(&scr, radix_sort_scratch_ints(n), "scratch"))) {
        cleanup();
        return rc;
    }
    hipError_t e = hipMemcpyAsync(k0, keys, sizeof(uint32_t) * n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = launch_iota(v0, n, ctx->stream);
    if (e == hipSuccess) e = radix_sort_pairs(k0, v0, k1, v1, n, scr, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(keys_sorted_out, k0, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(perm_out, v0, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    cleanup();
    if (e != hipSuccess) return fail(ICEMPM_ERR_HIP, std::string("radix_sort_pairs: ") + hipGetErrorString(e));
    return ICEMPM_OK;
}

int icempm_radix_sort_pairs_noalloc(icempm_ctx *ctx, const uint32_t *keys_dev, int n, uint32_t *keys_sorted_dev,
                                    int32_t *perm_dev) {
    if (!ctx) return fail(ICEMPM_ERR_INVALID_ARG, "null context");
    if (n < 0 || (n > 0 && (!keys_dev || !keys_sorted_dev || !perm_dev)))
        return fail(ICEMPM_ERR_INVALID_ARG, "bad arguments");
    if ((size_t)n > kNoAllocPairs)
        return fail(ICEMPM_ERR_INVALID_ARG, "radix_sort_pairs_noalloc: n = " + std::to_string(n) +
                                                " exceeds ICEMPM_NOALLOC_MAX_PAIRS = " + std::to_string(kNoAllocPairs));
    if (radix_sort_scratch_ints(n) > kNoAllocCountWords)
        return fail(ICEMPM_ERR_INVALID_ARG, "radix_sort_pairs_noalloc: digit counts exceed the fixed array");
    ICEMPM_TRY(set_device(ctx));
    if (n == 0) return ICEMPM_OK;
    int32_t *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mutex);
        base = g_noalloc_ptr[ctx->device];
    }
    if (!base) return fail(ICEMPM_ERR_NOT_READY, "radix_sort_pairs_noalloc: no fixed array on this device");
    uint32_t *keys_tmp = reinterpret_cast<uint32_t *>(base);
    int32_t *vals_tmp = base + kNoAllocPairs;
    int32_t *counts = base + 2 * kNoAllocPairs;
    ICEMPM_HIP(launch_copy_keys_iota(keys_dev, keys_sorted_dev, perm_dev, n, ctx->stream));
    ICEMPM_HIP(radix_sort_pairs(keys_sorted_dev, perm_dev, keys_tmp, vals_tmp, n, counts, ctx->stream));
    return ICEMPM_OK;
}

int icempm_download_grid(icempm_ctx *ctx, double *out) {
    ICEMPM_TRY(check_ready(ctx));
    if (!out) return fail(ICEMPM_ERR_INVALID_ARG, "null output");
    const long long nn = (long long)ctx->kp.gx * ctx->kp.gy * ctx->kp.gz;
    if (!ctx->grid_xmajor) ICEMPM_TRY(dev_alloc(&ctx->grid_xmajor, (size_t)nn * 4, "grid copy"));
    ICEMPM_HIP(launch_grid_to_xmajor(ctx->kp, ctx->grids[ctx->g].grid, ctx->grid_xmajor, ctx->stream));
    ICEMPM_HIP(hipMemcpyAsync(out, ctx->grid_xmajor, sizeof(double) * nn * 4, hipMemcpyDeviceToHost, ctx->stream));
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    return ICEMPM_OK;
}

int icempm_enable_timing(icempm_ctx *ctx, int on) {
    if (!ctx) return fail(ICEMPM_ERR_INVALID_ARG, "null context");
    ctx->timing = on != 0;
    return ICEMPM_OK;
}

int icempm_reset_timings(icempm_ctx *ctx) {
    if (!ctx) return fail(ICEMPM_ERR_INVALID_ARG, "null context");
    ICEMPM_TRY(set_device(ctx));
    ICEMPM_TRY(resolve_timings(ctx));
    for (int i = 0; i < ICEMPM_N_PHASES; i++) {
        ctx->ms[i] = 0;
        ctx->launches[i] = 0;
    }
    return ICEMPM_OK;
}

int icempm_read_timings(icempm_ctx *ctx, double *ms_out, long long *launches_out) {
    if (!ctx) return fail(ICEMPM_ERR_INVALID_ARG, "null context");
    ICEMPM_TRY(set_device(ctx));
    ICEMPM_TRY(resolve_timings(ctx));
    for (int i = 0; i < ICEMPM_N_PHASES; i++) {
        if (ms_out) ms_out[i] = ctx->ms[i];
        if (launches_out) launches_out[i] = ctx->launches[i];
    }
    return ICEMPM_OK;
}

extern "C++" {
namespace {

// (Re)allocates a device array of at least `count` doubles.
int adj_reserve(double **p, size_t &cap, size_t count, const char *what) {
    if (*p && cap >= count) return ICEMPM_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    cap = 0;
    ICEMPM_TRY(dev_alloc(p, count, what));
    cap = count;
    return ICEMPM_OK;
}

PtsView adj_view(double *base, int n) { return PtsView{base, (long long)n}; }

// Substeps s0 .. s0 + count - 1 of the step's own kernels from the state in
// slot(0); the state after substep s0 + j goes to slot(j + 1). P2G of each
// substep after the first is fused with the G2P before it (velocity and C
// stored), on the context's two grids. With keep, the grid after P2G and
// after the update of substep s0 + j is copied to G + 4 gp j / V + 4 gp j;
// with record, the force samples are accumulated.
template <typename Slot>
int adj_run(icempm_ctx *ctx, int s0, int count, Slot slot, bool keep, bool record) {
    auto &A = ctx->adj;
    const long long gp = ctx->kp.grid_pitch;
    const int S = A.run.sample_every;
    ICEMPM_TRY(ensure_split(ctx));
    const SegInfo seg = ctx->seginfo();
    int ga = 0, gb = 1;
    ICEMPM_TRY(do_reset_grid(ctx, ga));
    ICEMPM_HIP(launch_p2g(ctx->kp, adj_view(slot(0), A.n), A.n, seg, ctx->gstate(ga), ctx->grids[ga].cur, ctx->err,
                          ctx->p2g_blocks, ctx->stream));
    for (int j = 0; j < count; j++) {
        const int s = s0 + j;
        const double *c = &A.centers[3 * s];
        if (record && s % S == 0) ICEMPM_TRY(zero_bins(ctx));
        if (keep)
            ICEMPM_HIP(hipMemcpyAsync(A.G + (size_t)4 * gp * j, ctx->grids[ga].grid, sizeof(double) * 4 * gp,
                                      hipMemcpyDeviceToDevice, ctx->stream));
        // the update's launch also resets grid gb for the fused pass and
        // clears the spill count
        UpdateExtras x;
        const bool fused = j + 1 < count;
        if (fused) {
            x.reset = ctx->gstate(gb);
            x.reset_cur = ctx->grids[gb].cur;
        }
        x.zero = ctx->work;
        ICEMPM_HIP(launch_update_nodes(ctx->kp, ctx->gstate(ga), ctx->grids[ga].cur, c[0], c[1], c[2], A.times[s],
                                       ctx->bins, x, ctx->stream));
        if (fused) ctx->grids[gb].cur ^= 1;
        if (record && (s + 1) % S == 0)
            ICEMPM_HIP(launch_sum_bins(ctx->bins, ctx->nbins, 1.0 / S, A.forces + 3 * (s / S), ctx->stream));
        if (keep)
            ICEMPM_HIP(hipMemcpyAsync(A.V + (size_t)4 * gp * j, ctx->grids[ga].grid, sizeof(double) * 4 * gp,
                                      hipMemcpyDeviceToDevice, ctx->stream));
        // G2P reads slot(j) and writes the new state to slot(j + 1)
        const PtsView in = adj_view(slot(j), A.n), out = adj_view(slot(j + 1), A.n);
        if (fused) {
            ICEMPM_HIP(launch_g2p2g(ctx->kp, out, A.n, seg, ctx->grids[ga].grid, ctx->gstate(gb), ctx->grids[gb].cur, 0,
                                    1, ctx->err, ctx->work, ctx->p2g_blocks, nullptr, true, ctx->stream, &in));
            std::swap(ga, gb);
        } else {
            ICEMPM_HIP(launch_g2p(ctx->kp, out, A.n, seg, ctx->grids[ga].grid, 0, ctx->err, ctx->work,
                                  ctx->g2p_blocks, nullptr, true, ctx->stream, &in));
        }
    }
    return ICEMPM_OK;
}

int adj_check(icempm_ctx *ctx) {
    int flags[2] = {0, 0}, code = 0;
    ICEMPM_HIP(hipMemcpyAsync(flags, ctx->adj.flags, sizeof(flags), hipMemcpyDeviceToHost, ctx->stream));
    ICEMPM_HIP(hipMemcpyAsync(&code, ctx->err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    if (code) return fail(ICEMPM_ERR_DEVICE_FLAG, "adjoint run: device error indicator " + std::to_string(code));
    if (flags[1]) return fail(ICEMPM_ERR_INVALID_ARG, "adjoint run: liquid points are not differentiated");
    ctx->adj.any_crushed = flags[0] != 0; // sticky bit: the final state shows every crushed point
    return ICEMPM_OK;
}

} // namespace
} // extern "C++"

int icempm_adjoint_forward(icempm_ctx *ctx, const icempm_AdjointRun *run, double *forces_out) {
    ICEMPM_TRY(check_ready(ctx));
    if (!run || !forces_out) return fail(ICEMPM_ERR_INVALID_ARG, "null argument");
    if (ctx->ens.n > 0) return fail(ICEMPM_ERR_INVALID_ARG, "the adjoint runs one scene, not an ensemble");
    if (ctx->prms.SetupType == 2) return fail(ICEMPM_ERR_INVALID_ARG, "the adjoint does not cover SetupType 2");
    if (ctx->nbins <= 0) return fail(ICEMPM_ERR_INVALID_ARG, "the force samples need IndenterSubdivisions > 0");
    if (ctx->damage.enabled)
        return fail(ICEMPM_ADJ_UNSUPPORTED, "the adjoint does not differentiate the damage model");
    const icempm_AdjointRun R = *run;
    if (R.n_steps <= 0 || R.sample_every <= 0 || R.checkpoint_every <= 0 || R.n_steps % R.sample_every ||
        R.n_steps % R.checkpoint_every)
        return fail(ICEMPM_ERR_INVALID_ARG, "n_steps must be a positive multiple of sample_every and checkpoint_every");
    auto &A = ctx->adj;
    A.have_forward = false;
    A.run = R;
    A.n = ctx->n;
    A.nseg = R.n_steps / R.checkpoint_every;
    const int K = R.checkpoint_every, n = A.n;
    const long long gp = ctx->kp.grid_pitch;
    const size_t st = (size_t)ICEMPM_N_PTS_ARRAYS * std::max(n, 1);
    ICEMPM_TRY(adj_reserve(&A.ckpt, A.ckpt_cap, st * A.nseg, "adjoint checkpoints"));
    ICEMPM_TRY(adj_reserve(&A.traj, A.traj_cap, st * (K + 1), "adjoint segment states"));
    const size_t need_grid = (size_t)gp * 4 * (2 * K + 2);
    if (A.grid_cap < need_grid) {
        for (double **q : {&A.G, &A.V, &A.Gbar, &A.vbar}) {
            if (*q) (void)hipFree(*q);
            *q = nullptr;
        }
        for (int32_t **q : {&A.vbar_flag, &A.vbar_list, &A.vbar_count}) {
            if (*q) (void)hipFree(*q);
            *q = nullptr;
        }
        A.grid_cap = 0;
        ICEMPM_TRY(dev_alloc(&A.G, (size_t)gp * 4 * K, "adjoint grids"));
        ICEMPM_TRY(dev_alloc(&A.V, (size_t)gp * 4 * K, "adjoint grids"));
        ICEMPM_TRY(dev_alloc(&A.Gbar, (size_t)gp * 4, "adjoint grid adjoints"));
        ICEMPM_TRY(dev_alloc(&A.vbar, (size_t)gp * 4, "adjoint velocity adjoints"));
        ICEMPM_TRY(dev_alloc(&A.vbar_flag, (size_t)ctx->nbricks, "adjoint brick flags"));
        ICEMPM_TRY(dev_alloc(&A.vbar_list, (size_t)ctx->nbricks, "adjoint brick list"));
        ICEMPM_TRY(dev_alloc(&A.vbar_count, 2, "adjoint brick count"));
        ICEMPM_HIP(hipMemsetAsync(A.vbar_flag, 0, sizeof(int32_t) * ctx->nbricks, ctx->stream));
        ICEMPM_HIP(hipMemsetAsync(A.vbar_count, 0, sizeof(int32_t) * 2, ctx->stream));
        A.grid_cap = need_grid;
    }
    if (A.pt_cap < st) {
        for (double **q : {&A.bar, &A.part}) {
            if (*q) (void)hipFree(*q);
            *q = nullptr;
        }
        A.pt_cap = 0;
        ICEMPM_TRY(dev_alloc(&A.bar, st, "point adjoints"));
        ICEMPM_TRY(dev_alloc(&A.part, st, "point adjoints"));
        A.pt_cap = st;
    }
    const int nsamp = R.n_steps / R.sample_every;
    ICEMPM_TRY(adj_reserve(&A.forces, A.f_cap, 3 * (size_t)nsamp + 4, "force samples"));
    A.theta = A.forces + 3 * nsamp;
    if (!A.flags) ICEMPM_TRY(dev_alloc(&A.flags, 2, "adjoint flags"));
    ICEMPM_HIP(hipMemsetAsync(A.flags, 0, sizeof(int) * 2, ctx->stream));
    ICEMPM_TRY(clear_err(ctx));
    ICEMPM_HIP(hipMemsetAsync(A.forces, 0, sizeof(double) * 3 * nsamp, ctx->stream));
    // the context's points, in the order the device keeps them, are the initial state
    for (int l = 0; l < ICEMPM_N_PTS_ARRAYS && n > 0; l++)
        ICEMPM_HIP(hipMemcpyAsync(A.traj + (size_t)l * n, ctx->pts[ctx->cur_pts] + (size_t)l * ctx->pitch,
                                  sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    ICEMPM_HIP(launch_adj_flags(adj_view(A.traj, n), n, A.flags, ctx->stream));
    A.centers.assign(3 * (size_t)R.n_steps, 0.0);
    A.times.assign((size_t)R.n_steps, 0.0);
    double t = R.t0;
    for (int s = 0; s < R.n_steps; s++) {
        t += ctx->prms.InitialTimeStep;
        A.times[s] = t;
        for (int d = 0; d < 3; d++) A.centers[3 * s + d] = ctx->ind.center_initial[d] + t * ctx->ind.velocity[d];
    }
    // the grids' brick lists must be consistent for the step's reset
    for (int g = 0; g < 2; g++) {
        ICEMPM_HIP(launch_clear_grid_all(ctx->kp, ctx->gstate(g), ctx->nbricks, ctx->stream));
        ctx->grids[g].cur = 0;
    }
    // segment by segment, keeping each start; states ping-pong between two slots
    double *cur = A.traj, *nxt = A.traj + st;
    for (int seg_i = 0; seg_i < A.nseg; seg_i++) {
        ICEMPM_HIP(hipMemcpyAsync(A.ckpt + st * seg_i, cur, sizeof(double) * st, hipMemcpyDeviceToDevice,
                                  ctx->stream));
        double *a = cur, *b = nxt;
        ICEMPM_TRY(adj_run(ctx, seg_i * K, K, [&](int j) { return (j & 1) ? b : a; }, false, true));
        if (K & 1) std::swap(cur, nxt);
    }
    ICEMPM_HIP(launch_adj_flags(adj_view(cur, n), n, A.flags, ctx->stream));
    // the final positions and the point order, for position losses
    ICEMPM_TRY(adj_reserve(&A.xfin, A.x_cap, 3 * (size_t)std::max(n, 1), "final positions"));
    ICEMPM_TRY(adj_reserve(&A.dx, A.dx_cap, 3 * (size_t)std::max(n, 1), "position adjoints"));
    if (A.slot_cap < (size_t)std::max(n, 1)) {
        if (A.slot) (void)hipFree(A.slot);
        A.slot = nullptr;
        A.slot_cap = 0;
        ICEMPM_TRY(dev_alloc(&A.slot, (size_t)std::max(n, 1), "point order"));
        A.slot_cap = (size_t)std::max(n, 1);
    }
    if (n > 0) {
        ICEMPM_HIP(hipMemcpyAsync(A.xfin, cur + (size_t)ICEMPM_POSX * n, sizeof(double) * 3 * n,
                                  hipMemcpyDeviceToDevice, ctx->stream));
        ICEMPM_HIP(hipMemcpyAsync(A.slot, ctx->slot[ctx->cur_pts], sizeof(int32_t) * n, hipMemcpyDeviceToDevice,
                                  ctx->stream));
    }
    ICEMPM_HIP(hipMemcpyAsync(forces_out, A.forces, sizeof(double) * 3 * nsamp, hipMemcpyDeviceToHost, ctx->stream));
    ICEMPM_TRY(adj_check(ctx));
    A.have_forward = true;
    return ICEMPM_OK;
}

int icempm_adjoint_positions(icempm_ctx *ctx, double *positions_out) {
    ICEMPM_TRY(check_ready(ctx));
    auto &A = ctx->adj;
    if (!A.have_forward) return fail(ICEMPM_ERR_NOT_READY, "icempm_adjoint_forward has not completed");
    if (!positions_out) return fail(ICEMPM_ERR_INVALID_ARG, "null argument");
    const int n = A.n;
    if (n > 0) {
        ICEMPM_HIP(launch_lines_by_slot(A.xfin, A.slot, A.dx, n, 3, true, ctx->stream));
        ICEMPM_HIP(hipMemcpyAsync(positions_out, A.dx, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
    }
    ICEMPM_HIP(hipStreamSynchronize(ctx->stream));
    return ICEMPM_OK;
}

int icempm_adjoint_backward(icempm_ctx *ctx, const double *dL_dforces, double grad_out[4], double raw_out[4]) {
    if (!dL_dforces) return fail(ICEMPM_ERR_INVALID_ARG, "null argument");
    return icempm_adjoint_backward_ex(ctx, dL_dforces, nullptr, grad_out, raw_out);
}

int icempm_adjoint_backward_ex(icempm_ctx *ctx, const double *dL_dforces, const double *dL_dpositions,
                               double grad_out[4], double raw_out[4]) {
    ICEMPM_TRY(check_ready(ctx));
    auto &A = ctx->adj;
    if (!A.have_forward) return fail(ICEMPM_ERR_NOT_READY, "icempm_adjoint_forward has not completed");
    if (!grad_out) return fail(ICEMPM_ERR_INVALID_ARG, "null argument");
    if (A.n != ctx->n) return fail(ICEMPM_ERR_NOT_READY, "the points changed since icempm_adjoint_forward");
    const int K = A.run.checkpoint_every, n = A.n, S = A.run.sample_every;
    const long long gp = ctx->kp.grid_pitch;
    const size_t st = (size_t)ICEMPM_N_PTS_ARRAYS * std::max(n, 1);
    ICEMPM_HIP(hipMemsetAsync(A.bar, 0, sizeof(double) * st, ctx->stream));
    ICEMPM_HIP(hipMemsetAsync(A.part, 0, sizeof(double) * st, ctx->stream));
    ICEMPM_HIP(hipMemsetAsync(A.theta, 0, sizeof(double) * 4, ctx->stream));
    if (dL_dpositions && n > 0) {
        // the adjoint of the final state: dL/dx_N, from upload order into the device order
        ICEMPM_HIP(hipMemcpyAsync(A.dx, dL_dpositions, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
        ICEMPM_HIP(launch_lines_by_slot(A.dx, A.slot, A.bar + (size_t)ICEMPM_POSX * n, n, 3, false, ctx->stream));
    }
    const PtsView bar = adj_view(A.bar, n), part = adj_view(A.part, n);
    ICEMPM_TRY(ensure_split(ctx));
    const SegInfo seg = ctx->seginfo();
    GridState gb;
    gb.grid = A.vbar;
    gb.brick_flag = A.vbar_flag;
    gb.brick_list[0] = gb.brick_list[1] = A.vbar_list;
    gb.brick_count = A.vbar_count;
    gb.brick_coord = ctx->brick_coord;
    gb.brick_slot = ctx->brick_slot;
    for (int g = 0; g < 2; g++) {
        ICEMPM_HIP(launch_clear_grid_all(ctx->kp, ctx->gstate(g), ctx->nbricks, ctx->stream));
        ctx->grids[g].cur = 0;
    }
    // v_bar starts at zero; the reverse grid update leaves it zeroed after each substep
    ICEMPM_HIP(hipMemsetAsync(A.vbar, 0, sizeof(double) * 4 * gp, ctx->stream));
    for (int seg_i = A.nseg - 1; seg_i >= 0; seg_i--) {
        const int s0 = seg_i * K;
        ICEMPM_HIP(hipMemcpyAsync(A.traj, A.ckpt + st * seg_i, sizeof(double) * st, hipMemcpyDeviceToDevice,
                                  ctx->stream));
        ICEMPM_TRY(adj_run(ctx, s0, K, [&](int j) { return A.traj + st * j; }, true, false));
        for (int j = K - 1; j >= 0; j--) {
            const int s = s0 + j;
            double fbar[3];
            for (int i = 0; i < 3; i++) fbar[i] = dL_dforces ? dL_dforces[3 * (s / S) + i] / S : 0.0;
            const PtsView s_0 = adj_view(A.traj + st * j, n), s_1 = adj_view(A.traj + st * (j + 1), n);
            const double *G = A.G + (size_t)4 * gp * j, *V = A.V + (size_t)4 * gp * j;
            ICEMPM_HIP(launch_adj_g2p_bwd(ctx->kp, s_0, s_1, bar, part, n, seg, V, gb, 0, A.theta, A.any_crushed,
                                          ctx->p2g_blocks, ctx->stream));
            ICEMPM_HIP(launch_adj_update_bwd(ctx->kp, G, A.vbar, A.Gbar, &A.centers[3 * s], fbar, ctx->stream));
            ICEMPM_HIP(launch_adj_p2g_bwd(ctx->kp, s_0, A.Gbar, part, bar, A.theta, n, ctx->stream));
        }
    }
    double th[4] = {0.0, 0.0, 0.0, 0.0};
    ICEMPM_HIP(hipMemcpyAsync(th, A.theta, sizeof(th), hipMemcpyDeviceToHost, ctx->stream));
    ICEMPM_TRY(adj_check(ctx));
    // Lame chain rule (ComputeLame, parameters_sim_3d.cpp:164-169):
    // mu = E / (2 (1 + nu)), lambda = E nu / ((1 + nu)(1 - 2 nu)), kappa = 2 mu / 3 + lambda
    const double E = ctx->prms.YoungsModulus, nu = ctx->prms.PoissonsRatio;
    const double D = (1 + nu) * (1 - 2 * nu);
    const double dmu_dE = 1.0 / (2 * (1 + nu)), dmu_dnu = -E / (2 * (1 + nu) * (1 + nu));
    const double dl_dE = nu / D, dl_dnu = E * (1 + 2 * nu * nu) / (D * D);
    const double dk_dE = 2.0 / 3.0 * dmu_dE + dl_dE, dk_dnu = 2.0 / 3.0 * dmu_dnu + dl_dnu;
    grad_out[0] = th[0] * dmu_dE + th[1] * dk_dE;
    grad_out[1] = th[0] * dmu_dnu + th[1] * dk_dnu;
    // friction angle phi (degrees) and cohesion c: DP_tan_phi = tan(phi pi / 180)
    // (SetPhi, parameters_sim_3d.h:100), DP_threshold_p = c / DP_tan_phi
    const double tphi = ctx->prms.DP_tan_phi, pt = ctx->prms.DP_threshold_p, c = pt * tphi;
    const double dt_dphi = ICEMPM_PI / 180.0 * (1 + tphi * tphi);
    grad_out[2] = (th[3] - th[2] * c / (tphi * tphi)) * dt_dphi;
    grad_out[3] = th[2] / tphi;
    if (raw_out)
        for (int k = 0; k < 4; k++) raw_out[k] = th[k];
    return ICEMPM_OK;
}

} // extern "C"
