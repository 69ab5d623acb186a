// kernels_sort.hip — cell keys, device-wide scan, stable LSD radix sort of
// (key, index) pairs, particle permutation and brick segmentation.
//
// The reference sorts on the host only (std::sort over proxy iterators,
// host_side_soa.cpp:56-67, key proxypoint3d.cpp:158-163). Here the sort is a
// 4-pass, 8-bit-digit LSD radix sort whose scatter ranks keys in input order,
// so the result equals std::stable_sort by key.
#include "kernels.h"

namespace icempm {

namespace {

constexpr int kWave = 64;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems; // 2048

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / kWave;     // 4
constexpr int kSortRounds = 16;                      // rounds of 64 items per wave
constexpr int kSortTile = kSortThreads * kSortRounds; // 4096 items per block
constexpr int kRadix = 256;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Inclusive scan across a 64-lane wave.
__device__ __forceinline__ int wave_inclusive_scan(int x) {
    int lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        int y = __shfl_up(x, d, kWave);
        if (lane >= d) x += y;
    }
    return x;
}

// ---------------------------------------------------------------- keys
// Key of device point i, written at position order[i] (i itself when order is
// null) together with i in vals (when not null).
__global__ void k_cell_keys(KParams kp, PtsView pts, int n, int kind, uint32_t *keys, int32_t *vals,
                            const int32_t *order) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long long util = double_to_bits(pts.at(ICEMPM_IDX_UTILITY, i));
    uint32_t key;
    if (util & ICEMPM_STATUS_DISABLED) {
        key = 0xFFFFFFFFu;
    } else if (kind == 0) {
        // ProxyPoint3D::getCellIndex (proxypoint3d.cpp:158-163)
        int ix = (int)(pts.at(ICEMPM_POSX, i) * kp.h_inv + 0.5);
        int iy = (int)(pts.at(ICEMPM_POSX + 1, i) * kp.h_inv + 0.5);
        int iz = (int)(pts.at(ICEMPM_POSX + 2, i) * kp.h_inv + 0.5);
        key = (uint32_t)(ix * kp.gy * kp.gz + iy * kp.gz + iz);
    } else {
        // Same nearest-node cell, clamped into the grid, grouped in 4^3 bricks
        // (ensembles: the stored grid that holds every scene).
        const int g[3] = {kp.ens_n > 0 ? kp.nbx * 4 : kp.gx, kp.ens_n > 0 ? kp.nby * 4 : kp.gy,
                          kp.ens_n > 0 ? kp.nbz * 4 : kp.gz};
        int c[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double x = pts.at(ICEMPM_POSX + d, i) * kp.h_inv + 0.5;
            if (!(x >= 0.0)) x = 0.0;
            if (x > (double)(g[d] - 1)) x = (double)(g[d] - 1);
            c[d] = (int)x;
        }
        uint32_t brick = (uint32_t)(((c[0] >> 2) * kp.nby + (c[1] >> 2)) * kp.nbz + (c[2] >> 2));
        uint32_t local = (uint32_t)(((c[0] & 3) << 4) | ((c[1] & 3) << 2) | (c[2] & 3));
        key = brick * 64u + local;
    }
    const int o = order ? order[i] : i;
    keys[o] = key;
    if (vals) vals[o] = i;
}

__global__ void k_copy_keys_iota(const uint32_t *in, uint32_t *out, int32_t *idx, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        out[i] = in[i];
        idx[i] = i;
    }
}

__global__ void k_iota(int32_t *v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

__global__ void k_set_int(int32_t *p, int value) { *p = value; }

// ---------------------------------------------------------------- scan
// Block-local exclusive scan of kScanTile items; block totals to sums.
__global__ __launch_bounds__(kScanThreads) void k_scan_block(const int32_t *in, int32_t *out, int n,
                                                             int32_t *sums) {
    __shared__ int s_wave[kScanThreads / kWave];
    const int t = threadIdx.x;
    const long long base = (long long)blockIdx.x * kScanTile + (long long)t * kScanItems;
    int v[kScanItems];
    int acc = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        long long i = base + k;
        v[k] = (i < n) ? in[i] : 0;
        acc += v[k];
    }
    int incl = wave_inclusive_scan(acc);
    int wave = t / kWave;
    if (lane_id() == kWave - 1) s_wave[wave] = incl;
    __syncthreads();
    int wave_off = 0;
    for (int w = 0; w < wave; w++) wave_off += s_wave[w];
    int run = wave_off + incl - acc; // exclusive prefix of this thread
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        long long i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (sums && t == kScanThreads - 1) sums[blockIdx.x] = run;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(int32_t *out, int n, const int32_t *sums) {
    const int add = sums[blockIdx.x];
    const long long base = (long long)blockIdx.x * kScanTile;
    for (int k = threadIdx.x; k < kScanTile; k += kScanThreads) {
        long long i = base + k;
        if (i < n) out[i] += add;
    }
}

// ---------------------------------------------------------------- radix
// Per-wave digit histogram of the wave's 1024 consecutive items of the tile.
__device__ __forceinline__ void wave_histogram(const uint32_t *keys, int n, long long wave_start, int shift,
                                               int *hist) {
    const int lane = lane_id();
    for (int r = 0; r < kSortRounds; r++) {
        long long i = wave_start + (long long)r * kWave + lane;
        bool valid = i < n;
        uint32_t d = valid ? ((keys[i] >> shift) & (kRadix - 1)) : 0u;
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            unsigned long long bb = __ballot((d >> b) & 1u);
            m &= ((d >> b) & 1u) ? bb : ~bb;
        }
        if (valid) {
            unsigned long long below = m & ((1ull << lane) - 1ull);
            if (below == 0ull) hist[d] += __popcll(m);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ __launch_bounds__(kSortThreads) void k_radix_upsweep(const uint32_t *keys, int n, int shift,
                                                                int32_t *hist, int nblocks) {
    __shared__ int s_hist[kSortWaves][kRadix];
    const int t = threadIdx.x;
    const int wave = t / kWave;
    for (int k = t; k < kSortWaves * kRadix; k += kSortThreads) (&s_hist[0][0])[k] = 0;
    __syncthreads();
    long long wave_start = (long long)blockIdx.x * kSortTile + (long long)wave * kSortRounds * kWave;
    wave_histogram(keys, n, wave_start, shift, s_hist[wave]);
    __syncthreads();
    int tot = 0;
    for (int w = 0; w < kSortWaves; w++) tot += s_hist[w][t];
    hist[(long long)t * nblocks + blockIdx.x] = tot;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_downsweep(const uint32_t *keys_in,
                                                                  const int32_t *vals_in, uint32_t *keys_out,
                                                                  int32_t *vals_out, int n, int shift,
                                                                  const int32_t *offsets, int nblocks) {
    __shared__ int s_hist[kSortWaves][kRadix];
    __shared__ int s_gbase[kRadix];
    const int t = threadIdx.x;
    const int wave = t / kWave;
    const int lane = lane_id();
    for (int k = t; k < kSortWaves * kRadix; k += kSortThreads) (&s_hist[0][0])[k] = 0;
    __syncthreads();
    const long long wave_start = (long long)blockIdx.x * kSortTile + (long long)wave * kSortRounds * kWave;
    wave_histogram(keys_in, n, wave_start, shift, s_hist[wave]);
    __syncthreads();
    {
        // exclusive prefix over the waves of this tile, per digit
        int run = 0;
        for (int w = 0; w < kSortWaves; w++) {
            int c = s_hist[w][t];
            s_hist[w][t] = run;
            run += c;
        }
        s_gbase[t] = offsets[(long long)t * nblocks + blockIdx.x];
    }
    __syncthreads();
    int *run = s_hist[wave];
    for (int r = 0; r < kSortRounds; r++) {
        long long i = wave_start + (long long)r * kWave + lane;
        bool valid = i < n;
        uint32_t key = valid ? keys_in[i] : 0u;
        uint32_t d = (key >> shift) & (kRadix - 1);
        unsigned long long m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            unsigned long long bb = __ballot((d >> b) & 1u);
            m &= ((d >> b) & 1u) ? bb : ~bb;
        }
        int rank = __popcll(m & ((1ull << lane) - 1ull));
        int before = valid ? run[d] : 0;
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            long long pos = (long long)s_gbase[d] + before + rank;
            keys_out[pos] = key;
            vals_out[pos] = vals_in[i];
            if (rank == 0) run[d] = before + __popcll(m);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------- permute
__global__ void k_permute(PtsView in, PtsView out, const int32_t *slot_in, int32_t *slot_out,
                          const int32_t *perm, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int j = perm[i];
    const long long *src = reinterpret_cast<const long long *>(in.lines);
    long long *dst = reinterpret_cast<long long *>(out.lines);
#pragma unroll 4
    for (int l = 0; l < ICEMPM_N_PTS_ARRAYS; l++) dst[(long long)l * out.pitch + i] = src[(long long)l * in.pitch + j];
    slot_out[i] = slot_in[j];
}

__global__ void k_unpermute(PtsView in, PtsView out, const int32_t *slot, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = slot[i];
    const long long *src = reinterpret_cast<const long long *>(in.lines);
    long long *dst = reinterpret_cast<long long *>(out.lines);
#pragma unroll 4
    for (int l = 0; l < ICEMPM_N_PTS_ARRAYS; l++) dst[(long long)l * out.pitch + s] = src[(long long)l * in.pitch + i];
}

// ---------------------------------------------------------------- segments
__global__ void k_brick_heads(const uint32_t *keys, int n, int32_t *flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = keys[i];
    int head = (k != 0xFFFFFFFFu) && (i == 0 || (keys[i - 1] >> 6) != (k >> 6));
    flags[i] = head;
}

__global__ void k_brick_write(KParams kp, const uint32_t *keys, int n, const int32_t *flags, const int32_t *pos,
                              int32_t *seg_start, int32_t *seg_brick, int32_t *seg_coord, int32_t *seg_slot) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flags[i]) {
        const int b = (int)(keys[i] >> 6);
        const int bz = b % kp.nbz, by = (b / kp.nbz) % kp.nby, bx = b / (kp.nbz * kp.nby);
        seg_start[pos[i]] = i;
        seg_brick[pos[i]] = b;
        seg_coord[pos[i]] = bx | (by << 8) | (bz << 16);
        seg_slot[pos[i]] = kp.ens_n > 0 ? ensemble_slot(kp, bx, by, bz) : 0;
    }
}

__global__ void k_count_valid(const uint32_t *keys, int n, int32_t *count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (keys[i] != 0xFFFFFFFFu && (i == n - 1 || keys[i + 1] == 0xFFFFFFFFu)) *count = i + 1;
}

__global__ void k_brick_finish(const int32_t *flags, const int32_t *pos, int n, const int32_t *n_valid,
                               int32_t *seg_start, int32_t *nseg) {
    int s = (n > 0) ? pos[n - 1] + flags[n - 1] : 0;
    *nseg = s;
    seg_start[s] = *n_valid;
}

inline int blocks_for(long long n, int threads) { return (int)((n + threads - 1) / threads); }

} // namespace

hipError_t launch_cell_keys(const KParams &kp, PtsView pts, int n, int kind, uint32_t *keys, int32_t *vals,
                            hipStream_t stream, const int32_t *order) {
    if (n <= 0) return hipSuccess;
    k_cell_keys<<<blocks_for(n, 256), 256, 0, stream>>>(kp, pts, n, kind, keys, vals, order);
    return hipGetLastError();
}

hipError_t launch_copy_keys_iota(const uint32_t *in, uint32_t *out, int32_t *idx, int n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_copy_keys_iota<<<blocks_for(n, 256), 256, 0, stream>>>(in, out, idx, n);
    return hipGetLastError();
}

hipError_t launch_iota(int32_t *v, int n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_iota<<<blocks_for(n, 256), 256, 0, stream>>>(v, n);
    return hipGetLastError();
}

size_t exclusive_scan_scratch_ints(int n) {
    size_t total = 0;
    long long m = n;
    while (m > kScanTile) {
        m = (m + kScanTile - 1) / kScanTile;
        total += (size_t)m;
    }
    return total + 1;
}

hipError_t exclusive_scan(const int32_t *in, int32_t *out, int n, int32_t *scratch, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    int nb = blocks_for(n, kScanTile);
    if (nb == 1) {
        k_scan_block<<<1, kScanThreads, 0, stream>>>(in, out, n, nullptr);
        return hipGetLastError();
    }
    int32_t *sums = scratch;
    k_scan_block<<<nb, kScanThreads, 0, stream>>>(in, out, n, sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = exclusive_scan(sums, sums, nb, scratch + nb, stream);
    if (e != hipSuccess) return e;
    k_scan_add<<<nb, kScanThreads, 0, stream>>>(out, n, sums);
    return hipGetLastError();
}

size_t radix_sort_scratch_ints(int n) {
    size_t nblocks = (size_t)blocks_for(n > 0 ? n : 1, kSortTile);
    size_t h = nblocks * kRadix;
    return h + exclusive_scan_scratch_ints((int)h);
}

hipError_t radix_sort_pairs(uint32_t *keys, int32_t *vals, uint32_t *keys_tmp, int32_t *vals_tmp, int n,
                            int32_t *scratch, hipStream_t stream) {
    if (n <= 1) return hipSuccess;
    const int nblocks = blocks_for(n, kSortTile);
    int32_t *hist = scratch;
    int32_t *scan_scratch = scratch + (size_t)nblocks * kRadix;
    uint32_t *ka = keys, *kb = keys_tmp;
    int32_t *va = vals, *vb = vals_tmp;
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 8 * pass;
        k_radix_upsweep<<<nblocks, kSortThreads, 0, stream>>>(ka, n, shift, hist, nblocks);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = exclusive_scan(hist, hist, nblocks * kRadix, scan_scratch, stream);
        if (e != hipSuccess) return e;
        k_radix_downsweep<<<nblocks, kSortThreads, 0, stream>>>(ka, va, kb, vb, n, shift, hist, nblocks);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        uint32_t *tk = ka;
        ka = kb;
        kb = tk;
        int32_t *tv = va;
        va = vb;
        vb = tv;
    }
    // four passes: the result is back in keys/vals
    return hipSuccess;
}

hipError_t launch_permute(PtsView in, PtsView out, const int32_t *slot_in, int32_t *slot_out,
                          const int32_t *perm, int n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_permute<<<blocks_for(n, 256), 256, 0, stream>>>(in, out, slot_in, slot_out, perm, n);
    return hipGetLastError();
}

hipError_t launch_unpermute(PtsView in, PtsView out, const int32_t *slot, int n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    k_unpermute<<<blocks_for(n, 256), 256, 0, stream>>>(in, out, slot, n);
    return hipGetLastError();
}

hipError_t launch_count_valid(const uint32_t *sorted_keys, int n, int32_t *count_out, hipStream_t stream) {
    k_set_int<<<1, 1, 0, stream>>>(count_out, 0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || n <= 0) return e;
    k_count_valid<<<blocks_for(n, 256), 256, 0, stream>>>(sorted_keys, n, count_out);
    return hipGetLastError();
}

hipError_t launch_brick_segments(const KParams &kp, const uint32_t *sorted_keys, int n, int32_t *seg_start,
                                 int32_t *seg_brick, int32_t *seg_coord, int32_t *seg_slot, int32_t *nseg_out,
                                 int32_t *scratch, hipStream_t stream) {
    // scratch layout: flags[n] | pos[n] | n_valid[1] | scan scratch
    int32_t *flags = scratch;
    int32_t *pos = scratch + n;
    int32_t *n_valid = scratch + 2 * (size_t)n;
    int32_t *scan_scratch = n_valid + 1;
    hipError_t e = launch_count_valid(sorted_keys, n, n_valid, stream);
    if (e != hipSuccess) return e;
    if (n > 0) {
        k_brick_heads<<<blocks_for(n, 256), 256, 0, stream>>>(sorted_keys, n, flags);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = exclusive_scan(flags, pos, n, scan_scratch, stream)) != hipSuccess) return e;
        k_brick_write<<<blocks_for(n, 256), 256, 0, stream>>>(kp, sorted_keys, n, flags, pos, seg_start, seg_brick,
                                                               seg_coord, seg_slot);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    k_brick_finish<<<1, 1, 0, stream>>>(flags, pos, n, n_valid, seg_start, nseg_out);
    return hipGetLastError();
}

} // namespace icempm
