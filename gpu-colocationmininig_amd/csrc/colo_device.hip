// colo_device.hip -- HIP kernels for gfx950 (MI355X) and the device half of
// the C ABI: binning, filter, neighbour-cell join (count / write), PI.
//
// Reference (spatialdatasciencegroup/GPU-ColocationMininig):
//   Codes/DeviceProcess_Kernel.h (K), Codes/DeviceProcess.cu (cu),
//   Codes/Utility_Inline.h (U), Codes/colocationFinder.cpp (cpp).
//
// Numerics: all coordinate arithmetic is binary32 with round-to-nearest and
// no contraction (this file is compiled with -ffp-contract=off), matching
// the reference's Real = float (Codes/DataTypes.h:11):
//   cell  = gx * trunc((y - minY) / r) + trunc((x - minX) / r)   (K:615-625)
//   d2    = (dx*dx) + (dy*dy), neighbour iff d2 <= r2               (U:25-33)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "colo_internal.h"

#pragma clang fp contract(off)

using namespace colo;

namespace colo {

int set_error(colo_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

hipStream_t pick_stream(colo_ctx *c, void *s) { return s ? (hipStream_t)s : c->stream; }

int num_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) == hipSuccess) cus = p.multiProcessorCount;
        if (cus <= 0) cus = 256;
    }
    return cus;
}

}  // namespace colo

colo::DevData colo_ctx::dev() const {
    DevData d;
    d.g = DevGrid{minX, minY, r, r2, gx, gy, cells, F};
    d.x = d_x; d.y = d_y; d.cell_of = d_cell_of; d.sidx = d_sidx; d.spos = d_spos;
    d.sxy = d_sxy; d.scell = d_scell; d.dir = d_dir;
    d.fstart = d_fstart; d.fcount = d_fcount;
    d.mask3 = d_mask3;
    d.mask_w = mask_words(F);
    return d;
}

// ------------------------------------------------------------ utilities

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

static unsigned grid_for(uint64_t n, unsigned cap = 65535u * 4) {
    uint64_t b = (n + kBlock - 1) / kBlock;
    if (b == 0) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

int colo::fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t s) {
    if (!n) return COLO_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 8192)), dim3(kBlock), 0, s, p, n, v);
    return hipGetLastError() == hipSuccess ? COLO_OK : COLO_EHIP;
}

int colo::scratch(colo_ctx *c, size_t bytes, void **p) {
    colo_ctx *ctx_for_err = c;
    if (bytes > c->scratch_bytes) {
        if (c->d_scratch) COLO_HIP(hipFree(c->d_scratch));
        c->d_scratch = nullptr;
        size_t nb = std::max(bytes, c->scratch_bytes * 2);
        if (hipMalloc(&c->d_scratch, nb) != hipSuccess) {
            c->scratch_bytes = 0;
            return set_error(c, COLO_ENOMEM, "scratch allocation failed");
        }
        c->scratch_bytes = nb;
    }
    *p = c->d_scratch;
    return COLO_OK;
}

// ------------------------------------------------------------------ scan
// Three-phase exclusive scan: tile sums, scan of tile sums (recursive),
// tile scan + carry.  Tile = 256 threads x 8 items.

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__device__ inline uint64_t wave_incl_scan(uint64_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint64_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ inline uint64_t block_excl_scan(uint64_t v, uint64_t &excl) {
    __shared__ uint64_t wsum[kBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t inc = wave_incl_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint64_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; i++) {
        if (i < w) base += wsum[i];
        tot += wsum[i];
    }
    excl = base + inc - v;
    __syncthreads();
    return tot;
}

template <typename Tin>
__global__ void k_scan_reduce(const Tin *in, uint64_t n, uint64_t *partial) {
    uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++)
        if (base + i < n) s += in[base + i];
    uint64_t ex;
    uint64_t tot = block_excl_scan(s, ex);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

template <typename Tin, typename Tout>
__global__ void k_scan_apply(const Tin *in, uint64_t n, const uint64_t *carry, Tout *out, uint64_t *total) {
    uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint64_t v[kScanItems];
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        v[i] = (base + i < n) ? (uint64_t)in[base + i] : 0;
        s += v[i];
    }
    uint64_t ex;
    uint64_t tot = block_excl_scan(s, ex);
    uint64_t run = ex + (carry ? carry[blockIdx.x] : 0);
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        if (base + i < n) out[base + i] = (Tout)run;
        run += v[i];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1)
        *total = (carry ? carry[blockIdx.x] : 0) + tot;
}

// Segmented exclusive scans of many u32 arrays in three launches: tile
// sums, per-segment carries (one workgroup per segment), tile scans.
// Segment i writes out_i[0..n_i) and its total at out_i[n_i].
struct ScanSegDev {
    const uint32_t *in;
    uint64_t *out;
    uint64_t n, tile0, tile1;
    uint32_t chunked, pad;  // chunked: out[c] = sum of in[0 .. 64c)
};
struct ScanTile {
    uint32_t seg, pad;
    uint64_t start;
};

// Global loads and stores of a tile are striped (item r * kBlock + thread,
// consecutive lanes on consecutive items); each thread's kScanItems
// consecutive items are exchanged through LDS.  The padded index keeps the
// contiguous reads off one bank.
__device__ inline uint32_t scan_pad(uint32_t i) { return i + (i >> 5); }

__global__ __launch_bounds__(kBlock) void k_segscan_reduce(const ScanSegDev *segs, const ScanTile *tiles,
                                                           uint64_t *partial) {
    const ScanTile t = tiles[blockIdx.x];
    const ScanSegDev sg = segs[t.seg];
    const uint64_t end = sg.n < t.start + kScanTile ? sg.n : t.start + kScanTile;
    uint64_t v = 0;
    if (sg.chunked) {
        // sums of 64 rows: chunk r * 4 + wave is row r of the striped tile,
        // summed over the wave; kept in out until k_segscan_apply_chunks
        // turns them into prefixes
#pragma unroll
        for (int r = 0; r < kScanItems; r++) {
            const uint64_t idx = t.start + (uint64_t)r * kBlock + threadIdx.x;
            uint64_t g = idx < end ? sg.in[idx] : 0u;
            v += g;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) g += __shfl_xor(g, d, 64);
            const uint64_t ch = (t.start >> 6) + (uint64_t)r * (kBlock / 64) + (threadIdx.x >> 6);
            if ((threadIdx.x & 63) == 0 && ch < (sg.n + 63) >> 6) sg.out[ch] = g;
        }
    } else {
#pragma unroll
        for (int r = 0; r < kScanItems; r++) {
            const uint64_t idx = t.start + (uint64_t)r * kBlock + threadIdx.x;
            if (idx < end) v += sg.in[idx];
        }
    }
    uint64_t ex;
    const uint64_t tot = block_excl_scan(v, ex);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void k_segscan_carry(const ScanSegDev *segs, uint64_t *partial) {
    const ScanSegDev sg = segs[blockIdx.x];
    uint64_t run = 0;
    for (uint64_t t0 = sg.tile0; t0 < sg.tile1; t0 += kBlock) {
        const uint64_t t = t0 + threadIdx.x;
        const uint64_t v = t < sg.tile1 ? partial[t] : 0;
        uint64_t ex;
        const uint64_t tot = block_excl_scan(v, ex);
        if (t < sg.tile1) partial[t] = run + ex;
        run += tot;
    }
    if (threadIdx.x == 0) sg.out[sg.chunked ? (sg.n + 63) >> 6 : sg.n] = run;
}

// Chunked segments: the 32 chunk sums of each tile (kScanTile = 32 x 64
// rows) become the tile's carry plus their exclusive prefix.  One thread
// per chunk, eight tiles per workgroup.
__global__ __launch_bounds__(kBlock) void k_segscan_apply_chunks(const ScanSegDev *segs, const ScanTile *tiles,
                                                                 uint64_t nt, const uint64_t *carry) {
    static_assert(kScanTile == 32 * 64, "a tile holds 32 chunks of 64 rows");
    const uint64_t ti = (uint64_t)blockIdx.x * (kBlock / 32) + (threadIdx.x >> 5);
    const uint32_t q = threadIdx.x & 31;
    uint64_t v = 0, ch = 0;
    bool own = false;
    ScanSegDev sg{};
    if (ti < nt) {
        const ScanTile t = tiles[ti];
        sg = segs[t.seg];
        ch = (t.start >> 6) + q;
        own = sg.chunked && ch < (sg.n + 63) >> 6;
        if (own) v = sg.out[ch];
    }
    uint64_t x = v;
#pragma unroll
    for (int d = 1; d < 32; d <<= 1) {
        const uint64_t y = __shfl_up(x, d, 32);
        if (q >= (uint32_t)d) x += y;
    }
    if (own) sg.out[ch] = carry[ti] + x - v;
}

__global__ __launch_bounds__(kBlock) void k_segscan_apply(const ScanSegDev *segs, const ScanTile *tiles,
                                                          const uint64_t *carry) {
    __shared__ uint32_t s_in[kScanTile + kScanTile / 32];
    __shared__ uint64_t s_out[kScanTile + kScanTile / 32];
    const ScanTile t = tiles[blockIdx.x];
    const ScanSegDev sg = segs[t.seg];
    if (sg.chunked) return;  // the whole workgroup: one tile, one segment
    const uint64_t end = sg.n < t.start + kScanTile ? sg.n : t.start + kScanTile;
#pragma unroll
    for (int r = 0; r < kScanItems; r++) {
        const uint32_t e = (uint32_t)r * kBlock + threadIdx.x;
        const uint64_t idx = t.start + e;
        s_in[scan_pad(e)] = idx < end ? sg.in[idx] : 0u;
    }
    __syncthreads();
    uint64_t v[kScanItems];
    uint64_t sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        v[i] = s_in[scan_pad(threadIdx.x * kScanItems + i)];
        sum += v[i];
    }
    uint64_t ex;
    block_excl_scan(sum, ex);
    uint64_t run = ex + carry[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        s_out[scan_pad(threadIdx.x * kScanItems + i)] = run;
        run += v[i];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kScanItems; r++) {
        const uint32_t e = (uint32_t)r * kBlock + threadIdx.x;
        const uint64_t idx = t.start + e;
        if (idx < end) sg.out[idx] = s_out[scan_pad(e)];
    }
}

// Exclusive scan in -> out (may alias), 64-bit carries; *total (device) =
// sum.  tmp must hold scan_tmp_words(n) uint64.
static uint64_t scan_tmp_words(uint64_t n) {
    uint64_t w = 0;
    uint64_t b = (n + kScanTile - 1) / kScanTile;
    while (b > 1) {
        w += b + 1;
        b = (b + kScanTile - 1) / kScanTile;
    }
    return w + 2;
}

template <typename Tin, typename Tout>
static void scan_rec(const Tin *in, uint64_t n, Tout *out, uint64_t *total, uint64_t *tmp, hipStream_t s) {
    uint64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb <= 1) {
        hipLaunchKernelGGL((k_scan_apply<Tin, Tout>), dim3(1), dim3(kBlock), 0, s, in, n,
                           (const uint64_t *)nullptr, out, total);
        return;
    }
    uint64_t *partial = tmp;
    hipLaunchKernelGGL((k_scan_reduce<Tin>), dim3((unsigned)nb), dim3(kBlock), 0, s, in, n, partial);
    // exclusive scan of the partials in place, then the total of the last
    // tile is carry[nb-1] + its own sum (written by the apply kernel)
    scan_rec<uint64_t, uint64_t>(partial, nb, partial, nullptr, tmp + nb + 1, s);
    hipLaunchKernelGGL((k_scan_apply<Tin, Tout>), dim3((unsigned)nb), dim3(kBlock), 0, s, in, n,
                       (const uint64_t *)partial, out, total);
}

int colo::scan_u32_u64(colo_ctx *c, const uint32_t *in, uint64_t n, uint64_t *out, uint64_t *total_dev,
                       hipStream_t s) {
    void *tmp;
    int rc = scratch(c, scan_tmp_words(n) * 8, &tmp);
    if (rc) return rc;
    if (n == 0) {
        if (total_dev) return fill_u32((uint32_t *)total_dev, 2, 0, s);
        return COLO_OK;
    }
    scan_rec<uint32_t, uint64_t>(in, n, out, total_dev, (uint64_t *)tmp, s);
    return hipGetLastError() == hipSuccess ? COLO_OK : COLO_EHIP;
}

uint64_t colo::scan_ws_words(uint64_t n) { return scan_tmp_words(n); }

int colo::scan_slots_ws(const uint32_t *in, uint64_t n, uint64_t *out, uint64_t *ws, hipStream_t s) {
    if (n == 0) return fill_u32((uint32_t *)out, 2, 0, s);
    scan_rec<uint32_t, uint64_t>(in, n, out, out + n, ws, s);
    return hipGetLastError() == hipSuccess ? COLO_OK : COLO_EHIP;
}

// ------------------------------------------------------------- binning

// getCellID (K:615-625) in binary32; key = fid << cellbits | cell.
template <typename K>
__global__ void k_cell_keys(const float *x, const float *y, uint32_t n, const uint32_t *fstart, uint32_t F,
                            DevGrid g, uint32_t cellbits, uint32_t *cell_of, K *keys, uint32_t *vals,
                            uint32_t *flag) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float rx = (x[p] - g.minX) / g.r;
    float ry = (y[p] - g.minY) / g.r;
    uint32_t i = (uint32_t)rx;
    uint32_t j = (uint32_t)ry;
    uint32_t cell = g.gx * j + i;
    if (cell >= g.cells || !(rx > -1.0f) || !(ry > -1.0f) || !(rx < 4294967296.0f) || !(ry < 4294967296.0f)) {
        atomicOr(flag, 1u);
        cell = 0;
    }
    // feature of input index p: last f with fstart[f] <= p
    uint32_t lo = 0, hi = F;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (fstart[mid] <= p) lo = mid; else hi = mid;
    }
    cell_of[p] = cell;
    keys[p] = ((K)lo << cellbits) | (K)cell;
    vals[p] = p;
}

// LSD radix sort, 8 bits per pass, stable.  Tile = 256 threads x 16 keys.
constexpr int kRadixItems = 16;
constexpr int kRadixTile = kBlock * kRadixItems;
constexpr int kWaveKeys = kRadixTile / (kBlock / 64);

template <typename K>
__global__ void k_radix_hist(const K *keys, uint64_t n, int shift, uint32_t *hist, uint32_t ntiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * kRadixTile;
#pragma unroll 4
    for (int i = 0; i < kRadixItems; i++) {
        uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
        if (idx < n) atomicAdd(&h[(uint32_t)(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Scatter with wave-ballot digit ranking: each wave ranks its contiguous
// quarter of the tile in rounds of 64 keys (8 ballots give the lanes sharing
// a digit); per-wave digit counters in LDS make ranks stable across rounds;
// an exclusive scan over the 4 waves orders the waves.
template <typename K>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const K *kin, const uint32_t *vin, K *kout,
                                                           uint32_t *vout, uint64_t n, int shift,
                                                           const uint32_t *hscan, uint32_t ntiles) {
    __shared__ uint32_t wcnt[kBlock / 64][256];
    __shared__ uint32_t tbase[256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = 0; i < kBlock / 64; i++) wcnt[i][threadIdx.x] = 0;
    tbase[threadIdx.x] = hscan[(uint64_t)threadIdx.x * ntiles + blockIdx.x];
    __syncthreads();
    constexpr int R = kWaveKeys / 64;
    K key[R];
    uint32_t val[R], rank[R], dig[R];
    const uint64_t wbase = (uint64_t)blockIdx.x * kRadixTile + (uint64_t)w * kWaveKeys;
    const uint64_t lt = __lanemask_lt();
#pragma unroll
    for (int r = 0; r < R; r++) {
        uint64_t idx = wbase + (uint64_t)r * 64 + lane;
        bool valid = idx < n;
        key[r] = valid ? kin[idx] : (K)0;
        val[r] = valid ? vin[idx] : 0u;
        uint32_t d = (uint32_t)(key[r] >> shift) & 255u;
        dig[r] = d;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint64_t bb = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        uint32_t before = (uint32_t)__popcll(peers & lt);
        uint32_t prev = valid ? wcnt[w][d] : 0u;
        rank[r] = prev + before;
        __builtin_amdgcn_wave_barrier();
        if (valid && before == 0) wcnt[w][d] = prev + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    {
        uint32_t s = 0;
        for (int i = 0; i < kBlock / 64; i++) {
            uint32_t t = wcnt[i][threadIdx.x];
            wcnt[i][threadIdx.x] = s;
            s += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
        uint64_t idx = wbase + (uint64_t)r * 64 + lane;
        if (idx < n) {
            uint32_t dst = tbase[dig[r]] + wcnt[w][dig[r]] + rank[r];
            kout[dst] = key[r];
            vout[dst] = val[r];
        }
    }
}

template <typename K>
static int radix_sort(colo_ctx *c, K *keys, uint32_t *vals, K *keys2, uint32_t *vals2, uint64_t n, int bits,
                      uint32_t *hist, uint64_t *scan_tmp, K **kres, uint32_t **vres, hipStream_t s) {
    uint32_t ntiles = (uint32_t)((n + kRadixTile - 1) / kRadixTile);
    K *ki = keys, *ko = keys2;
    uint32_t *vi = vals, *vo = vals2;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL((k_radix_hist<K>), dim3(ntiles), dim3(kBlock), 0, s, ki, n, shift, hist, ntiles);
        scan_rec<uint32_t, uint32_t>(hist, (uint64_t)ntiles * 256, hist, nullptr, scan_tmp, s);
        hipLaunchKernelGGL((k_radix_scatter<K>), dim3(ntiles), dim3(kBlock), 0, s, ki, vi, ko, vo, n, shift,
                           hist, ntiles);
        std::swap(ki, ko);
        std::swap(vi, vo);
    }
    *kres = ki;
    *vres = vi;
    (void)c;
    return hipGetLastError() == hipSuccess ? COLO_OK : COLO_EHIP;
}

// Sorted-order gathers and the inverse permutation.
template <typename K>
__global__ void k_gather_sorted(const K *skeys, const uint32_t *sidx, uint32_t n, uint32_t cellbits,
                                const float *x, const float *y, float2 *sxy, uint32_t *spos,
                                uint32_t *scell) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint32_t i = sidx[p];
    sxy[p] = make_float2(x[i], y[i]);
    spos[i] = p;
    scell[p] = (uint32_t)(skeys[p] & (((K)1 << cellbits) - 1));
}

// Directory: dir[f * (cells + 1) + c] = first sorted position whose
// (feature, cell) index is at least (f, c), i.e. the exclusive prefix count
// of the per-(feature, cell) counts.  It is written tile by tile
// (kDirTile entries): k_dir_tiles flags the run starts (first element of
// each distinct key) and finds each tile's first element, k_dir_fill
// places the tile's run starts in LDS and fills the entries between them by
// a suffix minimum.  Every entry is written once and the keys are read
// twice; no counts, atomics or scan over the F x (cells + 1) entries.
constexpr int kDirItems = 16;
constexpr uint32_t kDirTile = kBlock * kDirItems;

template <typename K>
__device__ inline uint64_t dir_index(K key, uint32_t cellbits, uint32_t cells) {
    return (uint64_t)(key >> cellbits) * (cells + 1ull) + (uint64_t)(key & (((K)1 << cellbits) - 1));
}

// run_flag[p] = 1 at run starts; tile_lo[T] (T <= ntiles) = first sorted
// position whose directory index is >= T * kDirTile (n past the last).
template <typename K>
__global__ void k_dir_tiles(const K *skeys, uint32_t n, uint32_t cellbits, uint32_t cells, uint64_t ntiles,
                            uint32_t *run_flag, uint32_t *tile_lo) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const K k = skeys[p];
    const bool start = p == 0 || skeys[p - 1] != k;
    run_flag[p] = start ? 1u : 0u;
    const uint64_t d = dir_index(k, cellbits, cells);
    if (start) {
        // the tiles whose first entry lies in (index of p - 1, index of p]
        const uint64_t t0 = p ? dir_index(skeys[p - 1], cellbits, cells) / kDirTile + 1 : 0;
        for (uint64_t t = t0; t <= d / kDirTile; t++) tile_lo[t] = p;
    }
    if (p + 1 == n)
        for (uint64_t t = d / kDirTile + 1; t <= ntiles; t++) tile_lo[t] = n;
}

// With cellmask, each run start also sets its feature's bit in its cell's
// presence mask (W words per cell, cleared beforehand).
template <typename K>
__global__ __launch_bounds__(kBlock) void k_dir_fill(const K *skeys, uint32_t cellbits, uint32_t cells,
                                                     uint64_t dir_len, uint64_t ntiles, const uint32_t *tile_lo,
                                                     uint32_t *dir, unsigned long long *cellmask, uint32_t W) {
    __shared__ uint32_t v[kDirTile];
    __shared__ uint32_t wmin[kBlock / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint64_t e0 = t * kDirTile;
        const uint32_t lo = tile_lo[t], hi = tile_lo[t + 1];
        for (uint32_t i = tid; i < kDirTile; i += kBlock) v[i] = hi;
        __syncthreads();
        for (uint32_t p = lo + tid; p < hi; p += kBlock) {
            const K k = skeys[p];
            if (p == 0 || skeys[p - 1] != k) {
                v[dir_index(k, cellbits, cells) - e0] = p;
                if (cellmask) {
                    const uint32_t f = (uint32_t)(k >> cellbits), cell = (uint32_t)(k & (((K)1 << cellbits) - 1));
                    atomicOr(&cellmask[(uint64_t)cell * W + (f >> 6)], 1ull << (f & 63));
                }
            }
        }
        __syncthreads();
        // suffix minimum: thread tid owns entries [kDirItems * tid, + kDirItems)
        uint32_t r[kDirItems];
        uint32_t m = ~0u;
#pragma unroll
        for (int i = kDirItems - 1; i >= 0; i--) {
            m = min(m, v[kDirItems * tid + i]);
            r[i] = m;
        }
        uint32_t x = m;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_down(x, d);
            if (lane + d < 64) x = min(x, y);
        }
        uint32_t carry = __shfl_down(x, 1);
        if (lane == 63) carry = ~0u;
        if (lane == 0) wmin[wave] = x;
        __syncthreads();
        for (uint32_t w = wave + 1; w < kBlock / 64; w++) carry = min(carry, wmin[w]);
#pragma unroll
        for (int i = 0; i < kDirItems; i++) v[kDirItems * tid + i] = min(r[i], carry);
        __syncthreads();
        for (uint32_t i = tid; i < kDirTile && e0 + i < dir_len; i += kBlock) dir[e0 + i] = v[i];
        __syncthreads();
    }
}

__global__ void k_runs(const uint32_t *run_flag, const uint64_t *run_idx, const uint32_t *scell, uint32_t n,
                       uint32_t *run_cell) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !run_flag[p]) return;
    run_cell[run_idx[p]] = scell[p];
}

// Quadrant masks: OR of the 4 cells of quadrant (qx, qy), qx < gx-1, qy < gy-1
// (one thread per quadrant and mask word).
__global__ void k_qmask(const uint64_t *cellmask, uint32_t gx, uint32_t gy, uint32_t W, uint64_t *qmask) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nq = (uint64_t)(gx - 1) * (gy - 1);
    if (i >= nq * W) return;
    const uint64_t q = i / W;
    const uint32_t w = (uint32_t)(i - q * W);
    uint32_t qy = (uint32_t)(q / (gx - 1)), qx = (uint32_t)(q - (uint64_t)qy * (gx - 1));
    uint64_t c0 = (uint64_t)qy * gx + qx;
    qmask[i] = cellmask[c0 * W + w] | cellmask[(c0 + 1) * W + w] | cellmask[(c0 + gx) * W + w] |
               cellmask[(c0 + gx + 1) * W + w];
}

// Neighbourhood masks: OR of the cell masks over the clamped 3 x 3 block
// around each cell, a superset of the cells whose directory spans
// row_spans reads for an anchor there (one thread per cell and mask word).
__global__ void k_mask3(const uint64_t *cellmask, uint32_t gx, uint32_t gy, uint32_t W, uint64_t *mask3) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)gx * gy * W) return;
    const uint64_t c = i / W;
    const uint32_t w = (uint32_t)(i - c * W);
    const uint32_t ay = (uint32_t)(c / gx), ax = (uint32_t)(c - (uint64_t)ay * gx);
    const uint32_t x0 = ax ? ax - 1 : 0, x1 = ax + 1 < gx ? ax + 1 : gx - 1;
    const uint32_t y0 = ay ? ay - 1 : 0, y1 = ay + 1 < gy ? ay + 1 : gy - 1;
    uint64_t m = 0;
    for (uint32_t yy = y0; yy <= y1; yy++)
        for (uint32_t xx = x0; xx <= x1; xx++) m |= cellmask[((uint64_t)yy * gx + xx) * W + w];
    mask3[i] = m;
}

// Binning results for the host, in one block (u64 words): [0] bounds flag,
// [1] number of runs H, [2 .. 2+F] run offsets per feature (H last),
// [3+F .. 3+2F) sorted rank of each feature's first input instance,
// [3+2F .. 3+3F) of its last.
__global__ void k_bin_tail(const uint32_t *flag, const uint64_t *run_idx, const uint32_t *fstart,
                           const uint32_t *fcount, uint32_t F, const uint32_t *spos, uint64_t *out) {
    uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t H = out[1];  // written by the run scan
    if (f == 0) {
        out[0] = *flag;
        out[2 + F] = H;
    }
    if (f >= F) return;
    const uint32_t s0 = fstart[f], e = s0 + fcount[f] - 1;
    out[2 + f] = run_idx[s0];
    out[3 + F + f] = spos[s0] - s0;
    out[3 + 2 * F + f] = spos[e] - s0;
}

// ------------------------------------------------------------ filter
// A.6 / filter2_Kernel (K:469-509) + threshold_Kernel (K:545-573) as a
// segmented reduction: for candidate P and member t, sum the counts of the
// non-empty cells c of feature p_t for which one of the <= 4 quadrants
// containing c holds every feature of P (qmask superset of mask(P)).

struct FilterItem {
    uint32_t cand, t;
    uint64_t run_begin, run_end;
};

constexpr int kFilterRuns = kBlock * 8;

// One block = (feature f, up to kBlock candidate members (P, t) with
// P[t] = f, a tile of kFilterRuns non-empty cells of f).  The block stages
// each cell's <= 4 quadrant masks and count in LDS once; every thread then
// tests its candidate's mask against all staged cells (LDS broadcast), so
// a cell is read from memory once per group of candidates instead of once
// per candidate.
struct FilterGroup {
    uint32_t f, n_memb;
    uint64_t memb_off;  // first entry in the member list
    uint64_t run_begin, run_end;
};

// W: 64-bit words of a feature mask (features <= 64 W).
template <int W>
__global__ __launch_bounds__(kBlock) void k_filter(const FilterGroup *groups, const uint64_t *chunk_prefix,
                                                   uint64_t n_groups, const uint32_t *memb, const uint64_t *cmask,
                                                   uint32_t k, const uint32_t *run_cell, const uint32_t *dir,
                                                   const uint64_t *qmask, DevGrid g, uint32_t *counts) {
    __shared__ uint64_t s_q[W][4][kBlock];
    __shared__ uint32_t s_n[kBlock];
    uint64_t b = blockIdx.x;
    uint64_t lo = 0, hi = n_groups;
    while (hi - lo > 1) {
        uint64_t mid = (lo + hi) >> 1;
        if (chunk_prefix[mid] <= b) lo = mid; else hi = mid;
    }
    const FilterGroup G = groups[lo];
    const uint64_t r0 = G.run_begin + (b - chunk_prefix[lo]) * kFilterRuns;
    const uint64_t r1 = r0 + kFilterRuns < G.run_end ? r0 + kFilterRuns : G.run_end;
    const uint32_t *dirf = dir + (uint64_t)G.f * (g.cells + 1);
    // with at most 64 (128) members, four (two) copies of the member list
    // each test a quarter (half) of the staged cells, so that every wave
    // works; each copy adds its partial count
    const uint32_t nw = G.n_memb > 64 ? (G.n_memb + 63) / 64 : 1;
    const uint32_t copies = nw == 1 ? 4 : nw == 2 ? 2 : 1;
    const uint32_t span = 64 * nw, copy = threadIdx.x / span, mi = threadIdx.x - copy * span;
    const bool mine = copy < copies && mi < G.n_memb;
    const uint32_t me = mine ? memb[G.memb_off + mi] : 0;
    const uint32_t j0 = copy * (kBlock / copies);
    uint64_t M[W];
#pragma unroll
    for (int w = 0; w < W; w++) M[w] = mine ? cmask[(uint64_t)(me / k) * W + w] : 0;
    uint32_t sum = 0;
    const bool quads = g.gx >= 2 && g.gy >= 2;
    for (uint64_t base = r0; base < r1; base += kBlock) {
        const uint64_t h = base + threadIdx.x;
        uint64_t q[W][4];
#pragma unroll
        for (int w = 0; w < W; w++)
#pragma unroll
            for (int dq = 0; dq < 4; dq++) q[w][dq] = 0;
        uint32_t n = 0;
        if (h < r1 && quads) {
            const uint32_t c = run_cell[h];
            const uint32_t cy = c / g.gx, cx = c - cy * g.gx;
#pragma unroll
            for (int dq = 0; dq < 4; dq++) {
                const uint32_t dqy = dq >> 1, dqx = dq & 1;
                if (cy < dqy || cx < dqx) continue;
                const uint32_t qy = cy - dqy, qx = cx - dqx;
                if (qy + 1 >= g.gy || qx + 1 >= g.gx) continue;
#pragma unroll
                for (int w = 0; w < W; w++) q[w][dq] = qmask[((uint64_t)qy * (g.gx - 1) + qx) * W + w];
            }
            n = dirf[c + 1] - dirf[c];
        }
#pragma unroll
        for (int w = 0; w < W; w++)
#pragma unroll
            for (int dq = 0; dq < 4; dq++) s_q[w][dq][threadIdx.x] = q[w][dq];
        s_n[threadIdx.x] = n;
        __syncthreads();
        if (mine) {
            const uint32_t cnt = (uint32_t)(r1 - base < kBlock ? r1 - base : kBlock);
            const uint32_t j1 = j0 + kBlock / copies < cnt ? j0 + kBlock / copies : cnt;
            for (uint32_t j = j0; j < j1; j++) {
                bool ok = false;
#pragma unroll
                for (int dq = 0; dq < 4; dq++) {
                    bool sup = true;
#pragma unroll
                    for (int w = 0; w < W; w++) sup &= (s_q[w][dq][j] & M[w]) == M[w];
                    ok |= sup;
                }
                sum += ok ? s_n[j] : 0u;
            }
        }
        __syncthreads();
    }
    if (mine && sum) atomicAdd(&counts[me], sum);
}

// threshold_Kernel (K:545-573): ub_t = fl32(double(cnt)/double(fl32(N))),
// pass iff min_t ub_t >= theta.  Without a pass array the flag replaces
// counts[i*k], as the reference leaves it in countMap (K:570).
__global__ void k_threshold(uint32_t *counts, const uint32_t *cands, uint32_t C, uint32_t k,
                            const uint32_t *fcount, float theta, uint8_t *pass) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    float PI = 1.0f;
    for (uint32_t t = 0; t < k; t++) {
        double q = (double)counts[(uint64_t)i * k + t] / (double)(float)fcount[cands[(uint64_t)i * k + t]];
        float ub = (float)q;
        if (t == 0 || PI > ub) PI = ub;
    }
    if (pass)
        pass[i] = PI >= theta ? 1 : 0;
    else
        counts[(uint64_t)i * k] = PI >= theta ? 1u : 0u;
}

int colo::filter_batch(colo_ctx *c, const uint32_t *d_cands, uint32_t C, uint32_t k, float theta,
                       uint32_t *d_counts, uint8_t *d_pass, hipStream_t s) {
    colo_ctx *ctx_for_err = c;
    if (C == 0) return COLO_OK;
    const uint32_t W = mask_words(c->F);
    if (W > kMaxMaskWords)
        return set_error(c, COLO_EUNSUPPORTED, "the cell-aggregate filter supports at most 256 features (mine with the "
                                               "filter off for more)");
    std::vector<uint32_t> hc((uint64_t)C * k);
    COLO_HIP(hipMemcpyAsync(hc.data(), d_cands, hc.size() * 4, hipMemcpyDeviceToHost, s));
    COLO_HIP(hipStreamSynchronize(s));
    // members (P, t) grouped by feature; candidate masks
    std::vector<uint64_t> cmask((uint64_t)C * W, 0);
    std::vector<std::vector<uint32_t>> by_f(c->F);
    for (uint32_t i = 0; i < C; i++)
        for (uint32_t t = 0; t < k; t++) {
            uint32_t f = hc[(uint64_t)i * k + t];
            if (f >= c->F) return set_error(c, COLO_EARG, "candidate feature out of range");
            cmask[(uint64_t)i * W + (f >> 6)] |= 1ull << (f & 63);
            by_f[f].push_back(i * k + t);
        }
    std::vector<uint32_t> memb;
    std::vector<FilterGroup> groups;
    std::vector<uint64_t> prefix;
    uint64_t chunks = 0;
    for (uint32_t f = 0; f < c->F; f++) {
        const uint64_t rb = c->run_off[f], re = c->run_off[f + 1];
        const std::vector<uint32_t> &L = by_f[f];
        if (re == rb || L.empty()) continue;
        for (size_t g0 = 0; g0 < L.size(); g0 += kBlock) {
            const uint32_t n = (uint32_t)std::min<size_t>(kBlock, L.size() - g0);
            groups.push_back(FilterGroup{f, n, memb.size(), rb, re});
            memb.insert(memb.end(), L.begin() + g0, L.begin() + g0 + n);
            prefix.push_back(chunks);
            chunks += (re - rb + kFilterRuns - 1) / kFilterRuns;
        }
    }
    if ((uint64_t)C * k > 0xffffffffull) return set_error(c, COLO_EUNSUPPORTED, "too many candidates for the filter");
    int rc = fill_u32(d_counts, (uint64_t)C * k, 0, s);
    if (rc) return rc;
    if (!groups.empty()) {
        size_t bytes = groups.size() * sizeof(FilterGroup) + prefix.size() * 8 + cmask.size() * 8 + memb.size() * 4;
        void *tmp;
        rc = scratch(c, bytes, &tmp);
        if (rc) return rc;
        char *q = (char *)tmp;
        FilterGroup *d_groups = (FilterGroup *)q;
        q += groups.size() * sizeof(FilterGroup);
        uint64_t *d_prefix = (uint64_t *)q;
        q += prefix.size() * 8;
        uint64_t *d_cmask = (uint64_t *)q;
        q += cmask.size() * 8;
        uint32_t *d_memb = (uint32_t *)q;
        COLO_HIP(hipMemcpyAsync(d_groups, groups.data(), groups.size() * sizeof(FilterGroup), hipMemcpyHostToDevice,
                                s));
        COLO_HIP(hipMemcpyAsync(d_prefix, prefix.data(), prefix.size() * 8, hipMemcpyHostToDevice, s));
        COLO_HIP(hipMemcpyAsync(d_cmask, cmask.data(), cmask.size() * 8, hipMemcpyHostToDevice, s));
        COLO_HIP(hipMemcpyAsync(d_memb, memb.data(), memb.size() * 4, hipMemcpyHostToDevice, s));
        DevGrid g = c->dev().g;
        hipLaunchKernelGGL(W == 1 ? k_filter<1> : W == 2 ? k_filter<2> : W == 3 ? k_filter<3> : k_filter<4>,
                           dim3((unsigned)chunks), dim3(kBlock), 0, s, d_groups, d_prefix,
                           (uint64_t)groups.size(), d_memb, d_cmask, k, c->d_run_cell, c->d_dir, c->d_qmask, g,
                           d_counts);
        COLO_HIP(hipGetLastError());
        // the group arrays live in scratch: finish before the next scratch use
        COLO_HIP(hipStreamSynchronize(s));
    }
    return threshold_batch(c, d_cands, C, k, theta, d_counts, d_pass, s);
}

int colo::threshold_batch(colo_ctx *c, const uint32_t *d_cands, uint32_t C, uint32_t k, float theta,
                          uint32_t *d_counts, uint8_t *d_pass, hipStream_t s) {
    colo_ctx *ctx_for_err = c;
    if (C == 0) return COLO_OK;
    hipLaunchKernelGGL(k_threshold, dim3(grid_for(C)), dim3(kBlock), 0, s, d_counts, d_cands, C, k,
                       c->d_fcount, theta, d_pass);
    COLO_HIP(hipGetLastError());
    return COLO_OK;
}

// ------------------------------------------------------------------ join
// Neighbour cells of anchor cell (ax, ay) per getNeighborCellID (U:35-53)
// and the range check of K:28-29: rows ay-1..ay+1 clipped to [0, gy);
// columns ax-1..ax+1 clipped to [0, gx) when gx >= 3; only column ax when
// gx == 2 (the reference drops the other column as a wrap); none when
// gx == 1.  Scan order: rows outer, columns inner, then sorted position,
// which in the feature-major order is (cell, input index) -- the order of
// the reference's 3x3 loop (K:24-38).  For a fixed row the columns are
// adjacent cells, so one row is one contiguous range of sorted positions.

struct RowSpan {
    uint32_t lo[3], hi[3];
};

__device__ inline int row_spans(const DevGrid &g, const uint32_t *dirf, uint32_t cell, RowSpan &rs) {
    if (g.gx == 1) return 0;
    uint32_t ay = cell / g.gx, ax = cell - ay * g.gx;
    uint32_t cl, ch;
    if (g.gx == 2) {
        cl = ch = ax;
    } else {
        cl = ax ? ax - 1 : 0;
        ch = ax + 1 < g.gx ? ax + 1 : g.gx - 1;
    }
    uint32_t r0 = ay ? ay - 1 : 0;
    uint32_t r1 = ay + 1 < g.gy ? ay + 1 : g.gy - 1;
    // static indices keep rs in registers (rows past r1 repeat row r1)
#pragma unroll
    for (uint32_t i = 0; i < 3; i++) {
        const uint32_t ry = r0 + i <= r1 ? r0 + i : r1;
        const uint64_t base = (uint64_t)ry * g.gx;
        rs.lo[i] = dirf[base + cl];
        rs.hi[i] = dirf[base + ch + 1];
    }
    return (int)(r1 - r0 + 1);
}

__device__ inline float sqdist(float x1, float y1, float x2, float y2) {
    float dx = x2 - x1;
    float dy = y2 - y1;
    float sx = dx * dx;
    float sy = dy * dy;
    return sx + sy;
}

constexpr int kJoinUnroll = 4;   // list entries whose loads are issued together (counts)
constexpr int kWriteUnroll = 8;  // the same for the level-k write, which runs fewer rows

// A pointer read from a descriptor in memory, as a global-memory pointer
// (global rather than flat loads).
typedef __attribute__((address_space(1))) const uint32_t gcu32;
__device__ inline const gcu32 *as_global(const uint32_t *p) { return (const gcu32 *)p; }
// table outputs taken from descriptors, stored with global (not flat)
// instructions
typedef __attribute__((address_space(1))) uint32_t gu32;
__device__ inline gu32 *as_global_out(uint32_t *p) { return (gu32 *)p; }
typedef __attribute__((address_space(1))) const uint64_t gcu64;
__device__ inline const gcu64 *as_global(const uint64_t *p) { return (const gcu64 *)p; }

// List bounds [b0, b1) of anchor rank t from a u64 row-offsets array:
// entries t and t + 1 in one 16-byte load (8-byte aligned).
typedef uint64_t u64x2a8 __attribute__((ext_vector_type(2), aligned(8)));
typedef __attribute__((address_space(1))) const u64x2a8 gcu64x2;
__device__ inline void nb_bounds(const uint64_t *off, uint64_t base, uint64_t t, uint64_t &b0, uint64_t &b1) {
    const u64x2a8 v = *(const gcu64x2 *)(off + t);
    b0 = v.x - base;
    b1 = v.y - base;
}

__device__ inline void set_bit(uint32_t *bits, uint64_t i) {
    uint32_t m = 1u << (i & 31);
    uint32_t *w = bits + (i >> 5);
    if (!(__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & m)) atomicOr(w, m);
}

// Collects the bits a lane sets in ascending index order (candidates come
// in sorted position order) and issues one atomic per 32-bit word.
struct BitAcc {
    uint32_t *bits;
    uint64_t word = ~0ull;
    uint32_t mask = 0;
    __device__ explicit BitAcc(uint32_t *b) : bits(b) {}
    __device__ inline void flush() {
        if (mask) {
            uint32_t *w = bits + word;
            if ((__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & mask) != mask) atomicOr(w, mask);
        }
        mask = 0;
    }
    __device__ inline void set(uint64_t i) {
        uint64_t w = i >> 5;
        if (w != word) {
            flush();
            word = w;
        }
        mask |= 1u << (i & 31);
    }
};

// BitAcc whose words are first looked up in a workgroup's LDS table of
// words known to hold bits (cache, kBitCache entries of (arena word, mask),
// direct-mapped): bits found there need no check and no atomic.  An entry
// only ever holds bits that were seen set or whose atomicOr was issued, so a
// stale or overwritten entry loses hits, never bits; 64-bit LDS stores keep
// each (word, mask) pair whole.
constexpr uint32_t kBitCache = 256;
struct BitAccC {
    uint32_t *bits;
    const uint32_t *arena;
    unsigned long long *cache;
    uint64_t word = ~0ull;
    uint32_t mask = 0;
    __device__ BitAccC(uint32_t *b, const uint32_t *a, unsigned long long *c) : bits(b), arena(a), cache(c) {}
    __device__ inline void flush() {
        if (mask) {
            uint32_t *w = bits + word;
            const uint64_t off = (uint64_t)(w - arena);
            const uint32_t aw = (uint32_t)off;
            // words past 2^32 - 1 of the arena bypass the table (keys are 32-bit)
            unsigned long long *e = cache && off < 0xffffffffull ? &cache[aw % kBitCache] : nullptr;
            const unsigned long long cur = e ? *e : 0ull;
            if (!e || (uint32_t)cur != aw || ((uint32_t)(cur >> 32) & mask) != mask) {
                const uint32_t seen = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((seen & mask) != mask) atomicOr(w, mask);
                if (e) *e = ((unsigned long long)(seen | mask) << 32) | aw;
            }
        }
        mask = 0;
    }
    __device__ inline void set(uint64_t i) {
        uint64_t w = i >> 5;
        if (w != word) {
            flush();
            word = w;
        }
        mask |= 1u << (i & 31);
    }
};

// Level-2 work estimate of a (first, last) feature pair: over the non-empty
// cells c of `first`, count(first, c) x (instances of `last` in the 3x3
// cells around c) -- the number of distance tests of the 3x3 join.  Cells
// whose 3 x 3 mask lacks `last` add nothing and read no directory.
__global__ __launch_bounds__(kBlock) void k_pairwork(const FilterItem *items, const uint64_t *chunk_prefix,
                                                     uint64_t n_items, const uint32_t *pairs, const uint32_t *run_cell,
                                                     const uint32_t *dir, DevGrid g, const uint64_t *mask3,
                                                     uint32_t mask_w, unsigned long long *out) {
    uint64_t b = blockIdx.x;
    uint64_t lo = 0, hi = n_items;
    while (hi - lo > 1) {
        uint64_t mid = (lo + hi) >> 1;
        if (chunk_prefix[mid] <= b) lo = mid; else hi = mid;
    }
    const FilterItem it = items[lo];
    uint64_t r0 = it.run_begin + (b - chunk_prefix[lo]) * kFilterRuns;
    uint64_t r1 = r0 + kFilterRuns < it.run_end ? r0 + kFilterRuns : it.run_end;
    const uint32_t f0 = pairs[2 * it.cand], fl = pairs[2 * it.cand + 1];
    const uint32_t *dir0 = dir + (uint64_t)f0 * (g.cells + 1);
    const uint32_t *dirl = dir + (uint64_t)fl * (g.cells + 1);
    uint64_t sum = 0;
    for (uint64_t h = r0 + threadIdx.x; h < r1; h += kBlock) {
        uint32_t c = run_cell[h];
        if (mask3 && !((mask3[(uint64_t)c * mask_w + (fl >> 6)] >> (fl & 63)) & 1)) continue;
        RowSpan rs;
        int m = row_spans(g, dirl, c, rs);
        uint64_t nb = 0;
        for (int r = 0; r < m; r++) nb += rs.hi[r] - rs.lo[r];
        sum += (uint64_t)(dir0[c + 1] - dir0[c]) * nb;
    }
    uint64_t ex;
    uint64_t tot = block_excl_scan(sum, ex);
    if (threadIdx.x == 0 && tot) atomicAdd(&out[it.cand], (unsigned long long)tot);
}

// Refinement tests of size-k candidates: one workgroup per chunk of
// kFilterRuns parent rows of one candidate; each row adds the length of
// its first instance's neighbour list.
__global__ __launch_bounds__(kBlock) void k_refine_tests(const colo::RefineItem *items, const uint64_t *chunk_prefix,
                                                         uint64_t n_items, unsigned long long *out) {
    uint64_t b = blockIdx.x;
    uint64_t lo = 0, hi = n_items;
    while (hi - lo > 1) {
        uint64_t mid = (lo + hi) >> 1;
        if (chunk_prefix[mid] <= b) lo = mid; else hi = mid;
    }
    const colo::RefineItem it = items[lo];
    const uint64_t r0 = (b - chunk_prefix[lo]) * kFilterRuns;
    const uint64_t r1 = r0 + kFilterRuns < it.rows ? r0 + kFilterRuns : it.rows;
    uint64_t sum = 0;
    for (uint64_t h = r0 + threadIdx.x; h < r1; h += kBlock) {
        const uint32_t t = it.parent[h * it.km1] - it.fstart0;
        sum += it.nb_off[t + 1] - it.nb_off[t];
    }
    uint64_t ex;
    uint64_t tot = block_excl_scan(sum, ex);
    if (threadIdx.x == 0 && tot) atomicAdd(&out[lo], (unsigned long long)tot);
}

int colo::refine_tests_batch(colo_ctx *c, const RefineItem *items, uint32_t n, uint64_t *out, hipStream_t s) {
    colo_ctx *ctx_for_err = c;
    if (!n) return COLO_OK;
    std::vector<uint64_t> prefix(n);
    uint64_t chunks = 0;
    for (uint32_t i = 0; i < n; i++) {
        prefix[i] = chunks;
        chunks += (items[i].rows + kFilterRuns - 1) / kFilterRuns;
    }
    const size_t ib = ((size_t)n * sizeof(RefineItem) + 7) & ~(size_t)7;
    void *tmp;
    int rc = scratch(c, ib + (size_t)n * 16, &tmp);
    if (rc) return rc;
    RefineItem *d_items = (RefineItem *)tmp;
    uint64_t *d_prefix = (uint64_t *)((char *)tmp + ib);
    uint64_t *d_out = d_prefix + n;
    if ((rc = fill_u32((uint32_t *)d_out, 2ull * n, 0, s))) return rc;
    COLO_HIP(hipMemcpyAsync(d_items, items, (size_t)n * sizeof(RefineItem), hipMemcpyHostToDevice, s));
    COLO_HIP(hipMemcpyAsync(d_prefix, prefix.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
    if (chunks) {
        hipLaunchKernelGGL(k_refine_tests, dim3((unsigned)chunks), dim3(kBlock), 0, s, d_items, d_prefix, (uint64_t)n,
                           (unsigned long long *)d_out);
        COLO_HIP(hipGetLastError());
    }
    COLO_HIP(hipMemcpyAsync(out, d_out, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    COLO_HIP(hipStreamSynchronize(s));
    return COLO_OK;
}

int colo::pairwork_batch(colo_ctx *c, const uint32_t *pairs, uint32_t P, uint64_t *out, hipStream_t s) {
    colo_ctx *ctx_for_err = c;
    if (!P) return COLO_OK;
    std::vector<FilterItem> items;
    std::vector<uint64_t> prefix;
    uint64_t chunks = 0;
    for (uint32_t i = 0; i < P; i++) {
        uint32_t f = pairs[2 * i];
        uint64_t rb = c->run_off[f], re = c->run_off[f + 1];
        if (re == rb) continue;
        items.push_back(FilterItem{i, 0, rb, re});
        prefix.push_back(chunks);
        chunks += (re - rb + kFilterRuns - 1) / kFilterRuns;
    }
    size_t bytes = items.size() * sizeof(FilterItem) + prefix.size() * 8 + (size_t)P * 8 + (size_t)P * 8 + 64;
    void *tmp;
    int rc = scratch(c, bytes, &tmp);
    if (rc) return rc;
    char *q = (char *)tmp;
    FilterItem *d_items = (FilterItem *)q;
    q += items.size() * sizeof(FilterItem);
    uint64_t *d_prefix = (uint64_t *)q;
    q += prefix.size() * 8;
    uint32_t *d_pairs = (uint32_t *)q;
    q += (size_t)P * 8;
    uint64_t *d_out = (uint64_t *)(((uintptr_t)q + 7) & ~(uintptr_t)7);
    if ((rc = fill_u32((uint32_t *)d_out, 2ull * P, 0, s))) return rc;
    COLO_HIP(hipMemcpyAsync(d_pairs, pairs, (size_t)P * 8, hipMemcpyHostToDevice, s));
    if (!items.empty()) {
        COLO_HIP(hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(FilterItem), hipMemcpyHostToDevice, s));
        COLO_HIP(hipMemcpyAsync(d_prefix, prefix.data(), prefix.size() * 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_pairwork, dim3((unsigned)chunks), dim3(kBlock), 0, s, d_items, d_prefix,
                           (uint64_t)items.size(), d_pairs, c->d_run_cell, c->d_dir, c->dev().g, c->d_mask3,
                           mask_words(c->F), (unsigned long long *)d_out);
        COLO_HIP(hipGetLastError());
    }
    COLO_HIP(hipMemcpyAsync(out, d_out, (size_t)P * 8, hipMemcpyDeviceToHost, s));
    COLO_HIP(hipStreamSynchronize(s));
    return COLO_OK;
}

// Level-2 count for one anchor at sorted position p of feature a
// (degree2TableSlotCounter_Kernel, K:12-55).  The anchor's slot and bit are
// indexed by its input rank (the reference's layout) or, with SORTED, by
// its sorted rank (the batched miner's layout).
template <bool SORTED = false>
__device__ inline uint32_t join2_count_anchor(const DevData &D, uint32_t a, uint32_t b, uint32_t p, uint32_t *slots,
                                              uint32_t *bits_a, uint32_t *bits_b) {
    const uint32_t sa = D.fstart[a], sb = D.fstart[b];
    const uint32_t t = SORTED ? p - sa : D.sidx[p] - sa;
    const float xa = D.sxy[p].x, ya = D.sxy[p].y;
    const uint32_t *dirb = D.dir + (uint64_t)b * (D.g.cells + 1);
    RowSpan rs;
    int m = row_spans(D.g, dirb, D.scell[p], rs);
    uint32_t cnt = 0;
    BitAcc acc(bits_b);
    for (int r = 0; r < m; r++) {
        uint32_t q = rs.lo[r];
        for (; q + kJoinUnroll <= rs.hi[r]; q += kJoinUnroll) {
            float qx[kJoinUnroll], qy[kJoinUnroll];
#pragma unroll
            for (int u = 0; u < kJoinUnroll; u++) {
                qx[u] = D.sxy[q + u].x;
                qy[u] = D.sxy[q + u].y;
            }
#pragma unroll
            for (int u = 0; u < kJoinUnroll; u++)
                if (sqdist(xa, ya, qx[u], qy[u]) <= D.g.r2) {
                    cnt++;
                    acc.set(q + u - sb);
                }
        }
        for (; q < rs.hi[r]; q++) {
            if (sqdist(xa, ya, D.sxy[q].x, D.sxy[q].y) <= D.g.r2) {
                cnt++;
                acc.set(q - sb);
            }
        }
    }
    acc.flush();
    slots[t] = cnt;
    if (cnt) set_bit(bits_a, t);
    return cnt;
}

// Level-k count for one parent row (generalInstanceTableSlotCounter_Kernel,
// K:255-316): members row[0..km1-1], anchor row[0], last feature fl.
__device__ inline void joink_count_row(const DevData &D, const uint32_t *row, uint32_t km1, const uint32_t *feat,
                                       uint32_t fl, uint32_t *slot, uint32_t *bits, const uint64_t *seg_off) {
    float mx[kMaxK], my[kMaxK];
    for (uint32_t m = 0; m < km1; m++) {
        mx[m] = D.x[row[m]];
        my[m] = D.y[row[m]];
    }
    const uint32_t sl = D.fstart[fl];
    const uint32_t *dirl = D.dir + (uint64_t)fl * (D.g.cells + 1);
    RowSpan rs;
    int nr = row_spans(D.g, dirl, D.cell_of[row[0]], rs);
    uint32_t cnt = 0;
    BitAcc acc(bits + seg_off[km1]);
    for (int r = 0; r < nr; r++) {
        for (uint32_t q = rs.lo[r]; q < rs.hi[r]; q++) {
            const float qx = D.sxy[q].x, qy = D.sxy[q].y;
            bool ok = true;
            for (uint32_t m = 0; m < km1; m++)
                if (sqdist(mx[m], my[m], qx, qy) > D.g.r2) { ok = false; break; }
            if (ok) {
                cnt++;
                acc.set(q - sl);
            }
        }
    }
    acc.flush();
    *slot = cnt;
    if (cnt)
        for (uint32_t m = 0; m < km1; m++) set_bit(bits + seg_off[m], row[m] - D.fstart[feat[m]]);
}

// Batched level-2 join of one anchor (sorted position p of feature a): its
// (up to three) row ranges of feature b are walked as one stream, range 0,
// then 1, then 2 -- the scan order -- FLAT entries per step (kFlat; the
// packed count takes 2: its lanes all have entries, and fewer loads in
// flight per lane leave room for more waves), so a short range does not
// cost a dependent round trip of its own.  f(q, hit) is
// called for every entry in stream order while keep() holds.  With MASK,
// an anchor whose cell's 3 x 3 mask lacks b has no entries and reads no
// directory (the write pass visits only anchors with rows).
constexpr int kFlat = 8;
template <bool MASK, int FLAT = kFlat, typename F, typename K>
__device__ inline void join2_stream(const DevData &D, uint32_t b, uint32_t p, F &&f, K &&keep) {
    const uint32_t cell = D.scell[p];
    if (MASK && D.mask3 && !((D.mask3[(uint64_t)cell * D.mask_w + (b >> 6)] >> (b & 63)) & 1)) return;
    const float2 pa = D.sxy[p];
    const uint32_t *dirb = D.dir + (uint64_t)b * (D.g.cells + 1);
    RowSpan rs;
    const int m = row_spans(D.g, dirb, cell, rs);
    const uint32_t lo0 = m > 0 ? rs.lo[0] : 0, lo1 = m > 1 ? rs.lo[1] : 0, lo2 = m > 2 ? rs.lo[2] : 0;
    const uint32_t p1 = m > 0 ? rs.hi[0] - lo0 : 0;
    const uint32_t p2 = p1 + (m > 1 ? rs.hi[1] - lo1 : 0);
    const uint32_t tot = p2 + (m > 2 ? rs.hi[2] - lo2 : 0);
    for (uint32_t s0 = 0; s0 < tot && keep(); s0 += FLAT) {
        uint32_t q[FLAT];
        float2 c[FLAT];
#pragma unroll
        for (int u = 0; u < FLAT; u++) {
            const uint32_t e = s0 + u;
            q[u] = e < p1 ? lo0 + e : e < p2 ? lo1 + (e - p1) : lo2 + (e - p2);
            if (e < tot) c[u] = D.sxy[q[u]];
        }
#pragma unroll
        for (int u = 0; u < FLAT; u++)
            if (s0 + u < tot) f(q[u], sqdist(pa.x, pa.y, c[u].x, c[u].y) <= D.g.r2);
    }
}

// Batched level-2 count of anchor rank t of feature a (join2_count_anchor
// with SORTED, entries walked by join2_stream; MASK: test the anchor cell's
// 3 x 3 mask first).  acc collects the bits of feature b; bits_a null: the
// anchor's own bit is left to the caller.
template <bool MASK = true, int FLAT = kFlat, typename Acc = BitAcc>
__device__ inline uint32_t join2_count_flat(const DevData &D, uint32_t a, uint32_t b, uint32_t t, uint32_t *slots,
                                            uint32_t *bits_a, Acc acc) {
    const uint32_t sb = D.fstart[b];
    uint32_t cnt = 0;
    join2_stream<MASK, FLAT>(
        D, b, D.fstart[a] + t,
        [&](uint32_t q, bool hit) {
            if (hit) {
                cnt++;
                acc.set(q - sb);
            }
        },
        [] { return true; });
    acc.flush();
    if (slots) slots[t] = cnt;
    if (cnt && bits_a) set_bit(bits_a, t);
    return cnt;
}

// Batched level-2 write of anchor rank t of feature a (join2_write_anchor
// with SORTED): its nrows rows (position, position) at row offset off, the
// second column also into nbr.
__device__ inline void join2_write_flat(const DevData &D, uint32_t a, uint32_t b, uint32_t t, uint32_t nrows,
                                        uint64_t off, uint32_t *table, uint32_t *nbr) {
    const uint32_t g = D.fstart[a] + t;
    gu32 *out = as_global_out(table + off * 2);
    gu32 *nb = nbr ? as_global_out(nbr + off) : nullptr;
    uint32_t w = 0;
    join2_stream<false>(
        D, b, g,
        [&](uint32_t q, bool hit) {
            if (hit) {
                if (w < nrows) {
                    out[2 * w] = g;
                    out[2 * w + 1] = q;
                    if (nb) nb[w] = q;
                }
                w++;
            }
        },
        [&] { return w < nrows; });
}

// Level-2 write for one anchor of feature a (degree2TableGenerator_Kernel,
// K:56-100): t is its input rank and the rows are (instance id, instance
// id) -- or, with SORTED, t is its sorted rank and the rows are (sorted
// position, sorted position).  nbr (optional) receives the second column as
// sorted positions.
template <bool SORTED = false>
__device__ inline void join2_write_anchor(const DevData &D, uint32_t a, uint32_t b, uint64_t t, uint32_t nrows,
                                          uint64_t off, uint32_t *table, uint32_t *nbr) {
    const uint32_t g = D.fstart[a] + (uint32_t)t;
    const float xa = SORTED ? D.sxy[g].x : D.x[g], ya = SORTED ? D.sxy[g].y : D.y[g];
    const uint32_t *dirb = D.dir + (uint64_t)b * (D.g.cells + 1);
    RowSpan rs;
    int m = row_spans(D.g, dirb, SORTED ? D.scell[g] : D.cell_of[g], rs);
    uint32_t *out = table + off * 2;
    uint32_t *nb = nbr ? nbr + off : nullptr;
    uint32_t w = 0;
    auto emit = [&](uint32_t q) {
        if (w < nrows) {
            out[2 * w] = g;
            out[2 * w + 1] = SORTED ? q : D.sidx[q];
            if (nb) nb[w] = q;
        }
        w++;
    };
    for (int r = 0; r < m && w < nrows; r++) {
        uint32_t q = rs.lo[r];
        for (; q + kJoinUnroll <= rs.hi[r] && w < nrows; q += kJoinUnroll) {
            float qx[kJoinUnroll], qy[kJoinUnroll];
#pragma unroll
            for (int u = 0; u < kJoinUnroll; u++) {
                qx[u] = D.sxy[q + u].x;
                qy[u] = D.sxy[q + u].y;
            }
#pragma unroll
            for (int u = 0; u < kJoinUnroll; u++)
                if (sqdist(xa, ya, qx[u], qy[u]) <= D.g.r2) emit(q + u);
        }
        for (; q < rs.hi[r] && w < nrows; q++)
            if (sqdist(xa, ya, D.sxy[q].x, D.sxy[q].y) <= D.g.r2) emit(q);
    }
}

// Level-k write for one parent row (generateInstanceTableGeneral_Kernel,
// K:318-385): rows row ++ [q].
__device__ inline void joink_write_row(const DevData &D, const uint32_t *row, uint32_t km1, uint32_t fl,
                                       uint32_t nrows, uint64_t off, uint32_t *table) {
    float mx[kMaxK], my[kMaxK];
    for (uint32_t m = 0; m < km1; m++) {
        mx[m] = D.x[row[m]];
        my[m] = D.y[row[m]];
    }
    const uint32_t *dirl = D.dir + (uint64_t)fl * (D.g.cells + 1);
    RowSpan rs;
    int nr = row_spans(D.g, dirl, D.cell_of[row[0]], rs);
    const uint32_t k = km1 + 1;
    uint32_t *out = table + off * k;
    uint32_t w = 0;
    for (int r = 0; r < nr && w < nrows; r++) {
        for (uint32_t q = rs.lo[r]; q < rs.hi[r]; q++) {
            const float qx = D.sxy[q].x, qy = D.sxy[q].y;
            bool ok = true;
            for (uint32_t m = 0; m < km1; m++)
                if (sqdist(mx[m], my[m], qx, qy) > D.g.r2) { ok = false; break; }
            if (ok) {
                for (uint32_t m = 0; m < km1; m++) out[(uint64_t)w * k + m] = row[m];
                out[(uint64_t)w * k + km1] = D.sidx[q];
                w++;
            }
        }
    }
}

// Level-k passes over the anchor's neighbour list in the last feature.
// The reference accepts q for a parent row when q lies in the anchor's 3x3
// cells and d2(row[m], q) <= r2 for every member (K:268-290).  For m = 0
// (the anchor) the first two conditions define exactly the rows of the
// anchor in the level-2 table of (feature of row[0], last feature), written
// in the same scan order (K:56-100); that pair is prevalent for every
// level-k candidate (all 2-subsets of a candidate are prevalent), so its
// table is kept.  Only members 1..k-2 remain to be tested.
__device__ inline void joink_count_row_nbr(const DevData &D, const uint32_t *row, uint32_t km1, const uint32_t *feat,
                                           uint32_t fl, uint32_t *slot, uint32_t *bits, const uint64_t *seg_off,
                                           const uint32_t *nbr, const uint64_t *nb_off, uint64_t nb_base) {
    float mx[kMaxK], my[kMaxK];
    for (uint32_t m = 1; m < km1; m++) {
        mx[m] = D.x[row[m]];
        my[m] = D.y[row[m]];
    }
    const uint32_t t = row[0] - D.fstart[feat[0]];
    uint64_t b0, b1;
    nb_bounds(nb_off, nb_base, t, b0, b1);
    const uint32_t sl = D.fstart[fl];
    uint32_t cnt = 0;
    BitAcc acc(bits + seg_off[km1]);
    for (uint64_t i = b0; i < b1; i++) {
        const uint32_t q = nbr[i];  // sorted position
        const float qx = D.sxy[q].x, qy = D.sxy[q].y;
        bool ok = true;
        for (uint32_t m = 1; m < km1; m++)
            if (sqdist(mx[m], my[m], qx, qy) > D.g.r2) { ok = false; break; }
        if (ok) {
            cnt++;
            acc.set(q - sl);
        }
    }
    acc.flush();
    *slot = cnt;
    if (cnt)
        for (uint32_t m = 0; m < km1; m++) set_bit(bits + seg_off[m], row[m] - D.fstart[feat[m]]);
}

__device__ inline void joink_write_row_nbr(const DevData &D, const uint32_t *row, uint32_t km1, const uint32_t *feat,
                                           uint32_t nrows, uint64_t off, uint32_t *table, const uint32_t *nbr,
                                           const uint64_t *nb_off, uint64_t nb_base) {
    float mx[kMaxK], my[kMaxK];
    for (uint32_t m = 1; m < km1; m++) {
        mx[m] = D.x[row[m]];
        my[m] = D.y[row[m]];
    }
    const uint32_t t = row[0] - D.fstart[feat[0]];
    uint64_t b0, b1;
    nb_bounds(nb_off, nb_base, t, b0, b1);
    const uint32_t k = km1 + 1;
    uint32_t *out = table + off * k;
    uint32_t w = 0;
    for (uint64_t i = b0; i < b1 && w < nrows; i++) {
        const uint32_t q = nbr[i];
        const float qx = D.sxy[q].x, qy = D.sxy[q].y;
        bool ok = true;
        for (uint32_t m = 1; m < km1; m++)
            if (sqdist(mx[m], my[m], qx, qy) > D.g.r2) { ok = false; break; }
        if (ok) {
            for (uint32_t m = 0; m < km1; m++) out[(uint64_t)w * k + m] = row[m];
            out[(uint64_t)w * k + km1] = D.sidx[q];
            w++;
        }
    }
}

// ---- per-candidate kernels (the plain schedule and the C ABI)

__global__ void k_join2_count(DevData D, uint32_t a, uint32_t b, uint32_t *slots, uint32_t *bits_a,
                              uint32_t *bits_b) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t Na = D.fcount[a];
    if (i >= Na) return;
    join2_count_anchor(D, a, b, D.fstart[a] + i, slots, bits_a, bits_b);
}

__global__ void k_join2_write(DevData D, uint32_t a, uint32_t b, const uint32_t *slots, const uint64_t *offs,
                              uint32_t *table, uint32_t *nbr) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= D.fcount[a]) return;
    uint32_t nr = slots[t];
    if (!nr) return;
    join2_write_anchor(D, a, b, t, nr, offs[t], table, nbr);
}

struct CandK {
    uint32_t km1, fl;
    uint32_t feat[kMaxK];
    uint64_t seg_off[kMaxK];
};

__global__ void k_joink_count(DevData D, const uint32_t *parent, uint64_t R, CandK ck, uint32_t *slots,
                              uint32_t *bits) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    joink_count_row(D, parent + i * ck.km1, ck.km1, ck.feat, ck.fl, slots + i, bits, ck.seg_off);
}

__global__ void k_joink_write(DevData D, const uint32_t *parent, uint64_t R, CandK ck, const uint32_t *slots,
                              const uint64_t *offs, uint32_t *table) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    uint32_t nr = slots[i];
    if (!nr) return;
    joink_write_row(D, parent + i * ck.km1, ck.km1, ck.fl, nr, offs[i], table);
}

struct NbrList {
    const uint32_t *table;
    const uint64_t *off;
    uint64_t base;
};

__global__ void k_joink_count_nbr(DevData D, const uint32_t *parent, uint64_t R, CandK ck, NbrList nb,
                                  uint32_t *slots, uint32_t *bits) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    joink_count_row_nbr(D, parent + i * ck.km1, ck.km1, ck.feat, ck.fl, slots + i, bits, ck.seg_off, nb.table,
                        nb.off, nb.base);
}

__global__ void k_joink_write_nbr(DevData D, const uint32_t *parent, uint64_t R, CandK ck, NbrList nb,
                                  const uint32_t *slots, const uint64_t *offs, uint32_t *table) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    uint32_t nr = slots[i];
    if (!nr) return;
    joink_write_row_nbr(D, parent + i * ck.km1, ck.km1, ck.feat, nr, offs[i], table, nb.table, nb.off, nb.base);
}

// ---- batched (persistent) kernels: a fixed grid drains a work queue of
// (candidate, anchor range) items in queue order.

// Work queue with one segment per XCD.  Workgroups are dispatched to the 8
// XCDs round-robin by workgroup id, so workgroup b drains segment b % 8
// first -- consecutive items (adjacent anchors of one candidate, which read
// the same neighbourhoods) then share that XCD's L2 -- and steals from the
// other segments when its own is empty.  seg holds 9 item boundaries; ctr
// holds 8 counters (zeroed before the launch).
constexpr int kXcds = 8;

__device__ inline bool next_item(uint32_t *ctr, const uint64_t *seg, uint64_t &item) {
    __shared__ uint64_t s_item;
    if (threadIdx.x == 0) {
        uint64_t got = ~0ull;
        const uint32_t home = blockIdx.x % kXcds;
        for (uint32_t d = 0; d < kXcds && got == ~0ull; d++) {
            const uint32_t x = (home + d) % kXcds;
            const uint64_t b = seg[x], e = seg[x + 1];
            if (b >= e || __hip_atomic_load(&ctr[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= e - b) continue;
            const uint32_t i = atomicAdd(&ctr[x], 1u);
            if (b + i < e) got = b + i;
        }
        s_item = got;
    }
    __syncthreads();
    item = s_item;
    __syncthreads();
    return item != ~0ull;
}

// Segmented OR over the lanes of a wave holding the same 32-bit word w, in
// runs of at most CAP lanes (CAP a power of two, log2(CAP) shuffle steps):
// the first lane of each run of equal words, and every CAP-th lane inside
// a run, gets the OR of the masks of the next CAP lanes of its word; every
// other lane 0.  Every lane's mask reaches a leader: leaders are at most CAP
// lanes apart inside a run.  Lanes of one word in separate runs each lead
// their own run (still correct, one atomic more).  All 64 lanes must take
// part; lanes without a bit pass msk 0.
template <uint32_t CAP>
__device__ inline uint32_t wave_word_or(uint32_t w, uint32_t msk) {
    static_assert(CAP >= 2 && CAP <= 64 && (CAP & (CAP - 1)) == 0, "CAP: power of two, 2..64");
    const uint32_t lane = __lane_id();
#pragma unroll
    for (uint32_t off = 1; off < CAP; off <<= 1) {
        const uint32_t ow = __shfl_down(w, off, 64);
        const uint32_t om = __shfl_down(msk, off, 64);
        if (lane + off < 64 && ow == w) msk |= om;
    }
    const uint32_t pw = __shfl_up(w, 1, 64);
    return ((lane & (CAP - 1)) == 0 || pw != w) ? msk : 0u;
}

// Participation bits of the members of the wave's rows for one candidate,
// combined over the wave: adjacent rows share their anchor and mostly their
// members' 32-bit words, so one check and one atomic per distinct word
// (per run of up to 8 lanes for the anchor, 2 for the other members; more
// shuffle steps cost more than the atomics they save) replace one per row.
// All 64 lanes must take part (hit false for lanes without a row or
// without an extension).
template <int kArr>
__device__ inline void wave_member_bits(uint32_t *bt, const uint64_t *seg_off, const uint32_t *rk, uint32_t km1,
                                        bool hit) {
    uint32_t *w[kArr];
    uint32_t lm[kArr], cur[kArr];
#pragma unroll
    for (uint32_t m = 0; m < (uint32_t)kArr; m++) {
        if (m >= km1) break;
        const uint32_t wd = hit ? rk[m] >> 5 : 0xffffffffu, b = hit ? 1u << (rk[m] & 31) : 0u;
        lm[m] = m == 0 ? wave_word_or<8>(wd, b) : wave_word_or<2>(wd, b);
        w[m] = bt + seg_off[m] + (rk[m] >> 5);
    }
#pragma unroll
    for (uint32_t m = 0; m < (uint32_t)kArr; m++) {
        if (m >= km1) break;
        cur[m] = lm[m] ? __hip_atomic_load(w[m], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : ~0u;
    }
#pragma unroll
    for (uint32_t m = 0; m < (uint32_t)kArr; m++) {
        if (m >= km1) break;
        if ((cur[m] & lm[m]) != lm[m]) atomicOr(w[m], lm[m]);
    }
}

// Sets the participation bits of a row's members m0 .. km1-1: every check
// load is issued before the first atomic, so the members cost one round
// trip.
template <int kArr>
__device__ inline void set_member_bits(uint32_t *bt, const uint64_t *seg_off, const uint32_t *rk, uint32_t km1,
                                       uint32_t m0 = 0) {
    uint32_t *w[kArr];
    uint32_t cur[kArr];
#pragma unroll
    for (uint32_t m = 0; m < (uint32_t)kArr; m++) {
        if (m >= km1) break;
        w[m] = bt + seg_off[m] + (rk[m] >> 5);
        cur[m] = m >= m0 ? __hip_atomic_load(w[m], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : ~0u;
    }
#pragma unroll
    for (uint32_t m = 0; m < (uint32_t)kArr; m++) {
        if (m >= km1) break;
        const uint32_t msk = 1u << (rk[m] & 31);
        if (!(cur[m] & msk)) atomicOr(w[m], msk);
    }
}

// Participation bits of a work item staged in LDS (count variants 10-13).
// Every member and every last-feature instance of a row lies within r of
// the row's anchor, hence in the anchor's 3x3 cells, whose cell ids lie
// within gx + 1 of the anchor's; anchors come in cell order.  So the ranks
// an item can touch in column c lie in a window: the item's anchor ranks
// for the anchor column, and for the others the instances of the column's
// feature in cells cell(first anchor) - gx - 1 .. cell(last anchor) + gx + 1
// (two directory reads).  Each (candidate, column) whose window fits gets
// words of an LDS bitmap; hits set bits there with LDS atomics (no return
// value, no wait) and the item ends with one check and one global atomic
// per non-zero word, instead of a check round trip per hit word.  Windows
// that do not fit, and any rank outside its window, take the global path.
constexpr uint32_t kLdsBitWords = 3072;  // 12 KiB of bits per workgroup
constexpr uint32_t kWinCols = 8;         // columns (k <= 8: the specialised kernels)
struct LdsWin {
    uint32_t base_n;  // LDS word base (low 16 bits) and words (high 16 bits); 0 words = global path
    uint32_t gword;   // segment word of LDS word 0
};
__device__ inline uint32_t win_base(const LdsWin &w) { return w.base_n & 0xffffu; }
__device__ inline uint32_t win_words(const LdsWin &w) { return w.base_n >> 16; }

// Sets bit `rank` of column c: in the LDS window when it holds the rank,
// else in global memory (check, then atomic).
__device__ inline void lds_or_bit(uint32_t *s_bits, const LdsWin &w, uint32_t *gseg, uint32_t rank) {
    const uint32_t wi = (rank >> 5) - w.gword;
    if (wi < win_words(w))
        atomicOr(&s_bits[win_base(w) + wi], 1u << (rank & 31));
    else
        set_bit(gseg, rank);
}

// Windows of the item's rows [begin, end) for its ncand candidates (k
// columns each), LDS bases assigned in candidate order while they fit;
// clears the words in use.  Workgroup-wide (all threads call).  Anchor
// column: the item's anchor ranks; other member columns: the exact rank
// range of the item's rows (one min/max pass over them, shared by the run's
// candidates); last column: from the first entry of the first anchor's
// neighbour list to the last entry of the last anchor's (lists follow the
// anchors' cell order), or the cell-range window above when either list is
// empty.  A rank outside its window takes the global path, so a window only
// has to be a good guess.
__device__ inline void lds_setup(const DevData &D, const CandDesc *descs, uint32_t ncand, uint32_t k, uint64_t begin,
                                 uint64_t end, uint32_t *s_bits, LdsWin (*s_win)[kWinCols], uint32_t *s_used,
                                 uint32_t *s_mlo, uint32_t *s_mhi) {
    const CandDesc &c0 = descs[0];
    const uint32_t km1 = k - 1;
    if (threadIdx.x < kWinCols) {
        s_mlo[threadIdx.x] = ~0u;
        s_mhi[threadIdx.x] = 0u;
    }
    __syncthreads();
    {
        uint32_t lo[kWinCols], hi[kWinCols];
#pragma unroll
        for (uint32_t m = 0; m < kWinCols; m++) {
            lo[m] = ~0u;
            hi[m] = 0u;
        }
        for (uint64_t i = begin + threadIdx.x; i < end; i += kBlock)
#pragma unroll
            for (uint32_t m = 1; m < kWinCols; m++)
                if (m < km1) {
                    const uint32_t v = c0.parent[i * km1 + m];
                    lo[m] = v < lo[m] ? v : lo[m];
                    hi[m] = v > hi[m] ? v : hi[m];
                }
#pragma unroll
        for (uint32_t m = 1; m < kWinCols; m++)
            if (m < km1) {
#pragma unroll
                for (int d = 32; d >= 1; d >>= 1) {
                    const uint32_t ol = __shfl_xor(lo[m], d, 64), oh = __shfl_xor(hi[m], d, 64);
                    lo[m] = ol < lo[m] ? ol : lo[m];
                    hi[m] = oh > hi[m] ? oh : hi[m];
                }
                if (__lane_id() == 0) {
                    atomicMin(&s_mlo[m], lo[m]);
                    atomicMax(&s_mhi[m], hi[m]);
                }
            }
    }
    __syncthreads();
    const uint32_t a_first = c0.parent[begin * km1], a_last = c0.parent[(end - 1) * km1];
    const uint32_t f0s = D.fstart[c0.f[0]], t_first = a_first - f0s, t_last = a_last - f0s;
    const uint32_t cf = D.scell[a_first], cl = D.scell[a_last];
    const uint32_t lo_cell = cf > D.g.gx + 1 ? cf - D.g.gx - 1 : 0;
    const uint32_t hi_cell = (cl + D.g.gx + 1 < D.g.cells ? cl + D.g.gx + 1 : D.g.cells - 1) + 1;
    for (uint32_t e = threadIdx.x; e < ncand * k; e += kBlock) {
        const uint32_t j = e / k, c = e - j * k;
        const CandDesc &cd = descs[j];
        const uint32_t f = cd.f[c], fs = D.fstart[f];
        uint32_t lo, hi;
        if (c == 0) {
            lo = t_first;
            hi = t_last + 1;
        } else if (c < km1) {
            lo = s_mlo[c] - fs;
            hi = s_mhi[c] - fs + 1;
        } else {
            uint64_t b0, b1;
            nb_bounds(cd.nb_off, cd.nb_base, t_first, b0, b1);
            uint64_t e0, e1;
            nb_bounds(cd.nb_off, cd.nb_base, t_last, e0, e1);
            if (b1 > b0 && e1 > e0) {
                lo = cd.nb_table[b0] - fs;
                hi = cd.nb_table[e1 - 1] - fs + 1;
            } else {
                const uint32_t *dirf = D.dir + (uint64_t)f * (D.g.cells + 1);
                lo = dirf[lo_cell] - fs;
                hi = dirf[hi_cell] - fs;
            }
        }
        const uint32_t lw = lo >> 5, hw = hi > lo ? (hi + 31) >> 5 : lw;
        const uint32_t n = hw - lw < 0xffffu ? hw - lw : 0u;  // a window that large stays global
        s_win[j][c] = LdsWin{n << 16, lw};
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t used = 0;
        for (uint32_t j = 0; j < ncand; j++) {
            uint32_t tot = 0;
            for (uint32_t c = 0; c < k; c++) tot += win_words(s_win[j][c]);
            const bool fits = used + tot <= kLdsBitWords;
            for (uint32_t c = 0; c < k; c++) {
                const uint32_t n = fits ? win_words(s_win[j][c]) : 0u;
                s_win[j][c].base_n = (n << 16) | (fits ? used : 0u);
                if (fits) used += n;
            }
        }
        *s_used = used;
    }
    __syncthreads();
    for (uint32_t w = threadIdx.x; w < *s_used; w += kBlock) s_bits[w] = 0;
    __syncthreads();
}

// Writes the item's LDS bits to the candidates' bitmaps: one check and one
// atomic per non-zero word.  Workgroup-wide, after a barrier.
__device__ inline void lds_flush(const CandDesc *descs, uint32_t ncand, uint32_t k, uint32_t *bits,
                                 const uint32_t *s_bits, const LdsWin (*s_win)[kWinCols]) {
    for (uint32_t j = 0; j < ncand; j++)
        for (uint32_t c = 0; c < k; c++) {
            const LdsWin w = s_win[j][c];
            const uint32_t n = win_words(w);
            if (!n) continue;
            uint32_t *g = bits + descs[j].bits_off + descs[j].seg_off[c] + w.gword;
            for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
                const uint32_t v = s_bits[win_base(w) + i];
                if (v && (__hip_atomic_load(g + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & v) != v)
                    atomicOr(g + i, v);
            }
        }
}

// LDS-staged participation bits of one candidate j for the rows of a wave
// (row_multi and rows2w with LDSB): the members of a row with an
// extension.  The anchor bit is set by the first lane of each run of lanes
// with the same anchor and an extension; all 64 lanes must take part.
template <int kArr>
__device__ inline void lds_member_bits(uint32_t *s_bits, const LdsWin *win, uint32_t *bt, const uint64_t *seg_off,
                                       const uint32_t *rk, uint32_t km1, bool hit) {
    const uint32_t key = hit ? rk[0] : ~0u;
    const uint32_t pk = __shfl_up(key, 1, 64);
    if (!hit) return;
    const bool lead = __lane_id() == 0 || pk != key;
#pragma unroll
    for (uint32_t m = 0; m < (uint32_t)kArr; m++) {
        if (m >= km1) break;
        if (m == 0 && !lead) continue;
        lds_or_bit(s_bits, win[m], bt + seg_off[m], rk[m]);
    }
}

// Level-k count over the rows of one parent table for ncand candidates that
// extend it: the row and its members' coordinates are loaded once, and the
// members' ranks too (every candidate extends the same parent features).
// Tables of the batched miner hold sorted positions; slots and participation
// bits are indexed by sorted rank.  The next candidate's list bounds are
// loaded before this candidate's stores, one step ahead.  WAVE: all 64
// lanes of the wave call together (act false for lanes without a row) and
// the members' bits are combined over the wave.
template <int KM1, bool WAVE = false, bool LDSB = false>
__device__ inline void joink_count_row_multi(const DevData &D, const CandDesc *descs, uint32_t ncand, uint64_t i,
                                             uint32_t *slots, uint32_t *bits, unsigned long long *s_rows,
                                             bool act = true, uint32_t *s_bits = nullptr,
                                             const LdsWin (*s_win)[kWinCols] = nullptr) {
    const CandDesc &c0 = descs[0];
    const uint32_t km1 = KM1 ? (uint32_t)KM1 : c0.k - 1;
    constexpr int kArr = KM1 ? KM1 : kMaxK;
    const uint32_t *row = c0.parent + i * km1;
    uint32_t rk[kArr];
    float mx[kArr], my[kArr];
#pragma unroll
    for (uint32_t m = 0; m < (uint32_t)kArr; m++) {
        if (m >= km1) break;
        rk[m] = 0;
        mx[m] = my[m] = 0.f;
        if (act) {
            const uint32_t id = row[m];
            rk[m] = id - D.fstart[c0.f[m]];
            if (m) {
                mx[m] = D.sxy[id].x;
                my[m] = D.sxy[id].y;
            }
        }
    }
    const uint32_t t = rk[0];
    uint64_t nb0 = 0, nb1 = 0;
    if (act) {
        nb_bounds(c0.nb_off, c0.nb_base, t, nb0, nb1);
    }
    for (uint32_t j = 0; j < ncand; j++) {
        const CandDesc &cd = descs[j];
        const uint64_t b0 = nb0, b1 = nb1;
        if (act && j + 1 < ncand) {
            const CandDesc &cn = descs[j + 1];
            nb_bounds(cn.nb_off, cn.nb_base, t, nb0, nb1);
        }
        // descriptor fields in registers: the loop's bit stores could alias
        // the descriptor, which would otherwise be re-read every entry
        const gcu32 *nbt = as_global(cd.nb_table);
        const uint32_t sl = D.fstart[cd.f[km1]];
        uint32_t *bt = bits + cd.bits_off;
        uint32_t cnt = 0;
        BitAcc acc(bt + cd.seg_off[km1]);
        const LdsWin wl = LDSB ? s_win[j][km1] : LdsWin{0, 0};
        auto test = [&](uint32_t q, float qx, float qy) {
            bool ok = true;
#pragma unroll
            for (uint32_t m = 1; m < (uint32_t)kArr; m++) {
                if (m >= km1) break;
                if (sqdist(mx[m], my[m], qx, qy) > D.g.r2) ok = false;
            }
            if (ok) {
                cnt++;
                if (LDSB && win_words(wl)) {
                    const uint32_t r = q - sl, w = (r >> 5) - wl.gword;
                    if (w < win_words(wl))
                        atomicOr(&s_bits[win_base(wl) + w], 1u << (r & 31));
                    else
                        acc.set(r);
                } else {
                    acc.set(q - sl);
                }
            }
        };
        uint64_t e = b0;
        // kJoinUnroll entries per step, software-pipelined: the positions of
        // the next step are loaded while this step's coordinates arrive
        if (e + kJoinUnroll <= b1) {
            uint32_t qn[kJoinUnroll];
#pragma unroll
            for (int u = 0; u < kJoinUnroll; u++) qn[u] = nbt[e + u];
            for (; e + kJoinUnroll <= b1; e += kJoinUnroll) {
                uint32_t q[kJoinUnroll];
                float qx[kJoinUnroll], qy[kJoinUnroll];
#pragma unroll
                for (int u = 0; u < kJoinUnroll; u++) {
                    q[u] = qn[u];
                    qx[u] = D.sxy[q[u]].x;
                    qy[u] = D.sxy[q[u]].y;
                }
                if (e + 2 * kJoinUnroll <= b1) {
#pragma unroll
                    for (int u = 0; u < kJoinUnroll; u++) qn[u] = nbt[e + kJoinUnroll + u];
                }
#pragma unroll
                for (int u = 0; u < kJoinUnroll; u++) test(q[u], qx[u], qy[u]);
            }
        }
        for (; e < b1; e++) {
            const uint32_t q = nbt[e];
            test(q, D.sxy[q].x, D.sxy[q].y);
        }
        acc.flush();
        if (act && cd.slots_off != kNoSlots) slots[cd.slots_off + i] = cnt;
        if (cnt) atomicAdd(&s_rows[j], (unsigned long long)cnt);
        if (LDSB) {
            if (__ballot(cnt != 0)) lds_member_bits<kArr>(s_bits, s_win[j], bt, cd.seg_off, rk, km1, act && cnt != 0);
        } else if (WAVE) {
            if (__ballot(cnt != 0)) wave_member_bits<kArr>(bt, cd.seg_off, rk, km1, cnt != 0);
        }
        else if (cnt)
            set_member_bits<kArr>(bt, cd.seg_off, rk, km1);
    }
}

// Level-k count of two adjacent parent rows i, i + 1 with the same anchor:
// every candidate's list is read once and each entry tested against both
// rows (results equal two calls of joink_count_row_multi).
template <int KM1>
__device__ inline void joink_count_rows2(const DevData &D, const CandDesc *descs, uint32_t ncand, uint64_t i,
                                         uint32_t *slots, uint32_t *bits, unsigned long long *s_rows) {
    const CandDesc &c0 = descs[0];
    const uint32_t km1 = KM1 ? (uint32_t)KM1 : c0.k - 1;
    constexpr int kArr = KM1 ? KM1 : kMaxK;
    const uint32_t *row0 = c0.parent + i * km1, *row1 = row0 + km1;
    uint32_t rk0[kArr], rk1[kArr];
    float mx0[kArr], my0[kArr], mx1[kArr], my1[kArr];
#pragma unroll
    for (uint32_t m = 0; m < (uint32_t)kArr; m++) {
        if (m >= km1) break;
        const uint32_t id0 = row0[m], id1 = row1[m], fs = D.fstart[c0.f[m]];
        rk0[m] = id0 - fs;
        rk1[m] = id1 - fs;
        if (m) {
            mx0[m] = D.sxy[id0].x;
            my0[m] = D.sxy[id0].y;
            mx1[m] = D.sxy[id1].x;
            my1[m] = D.sxy[id1].y;
        }
    }
    const uint32_t t = rk0[0];
    uint64_t nb0, nb1;
    nb_bounds(c0.nb_off, c0.nb_base, t, nb0, nb1);
    for (uint32_t j = 0; j < ncand; j++) {
        const CandDesc &cd = descs[j];
        const uint64_t b0 = nb0, b1 = nb1;
        if (j + 1 < ncand) {
            const CandDesc &cn = descs[j + 1];
            nb_bounds(cn.nb_off, cn.nb_base, t, nb0, nb1);
        }
        const gcu32 *nbt = as_global(cd.nb_table);
        const uint32_t sl = D.fstart[cd.f[km1]];
        uint32_t *bt = bits + cd.bits_off;
        uint32_t cnt0 = 0, cnt1 = 0;
        BitAcc acc(bt + cd.seg_off[km1]);
        auto test = [&](uint32_t q, float qx, float qy) {
            bool ok0 = true, ok1 = true;
#pragma unroll
            for (uint32_t m = 1; m < (uint32_t)kArr; m++) {
                if (m >= km1) break;
                if (sqdist(mx0[m], my0[m], qx, qy) > D.g.r2) ok0 = false;
                if (sqdist(mx1[m], my1[m], qx, qy) > D.g.r2) ok1 = false;
            }
            cnt0 += ok0;
            cnt1 += ok1;
            if (ok0 || ok1) acc.set(q - sl);
        };
        uint64_t e = b0;
        for (; e + kJoinUnroll <= b1; e += kJoinUnroll) {
            uint32_t q[kJoinUnroll];
            float qx[kJoinUnroll], qy[kJoinUnroll];
#pragma unroll
            for (int u = 0; u < kJoinUnroll; u++) q[u] = nbt[e + u];
#pragma unroll
            for (int u = 0; u < kJoinUnroll; u++) {
                qx[u] = D.sxy[q[u]].x;
                qy[u] = D.sxy[q[u]].y;
            }
#pragma unroll
            for (int u = 0; u < kJoinUnroll; u++) test(q[u], qx[u], qy[u]);
        }
        for (; e < b1; e++) {
            const uint32_t q = nbt[e];
            test(q, D.sxy[q].x, D.sxy[q].y);
        }
        acc.flush();
        if (cd.slots_off != kNoSlots) {
            slots[cd.slots_off + i] = cnt0;
            slots[cd.slots_off + i + 1] = cnt1;
        }
        if (cnt0 + cnt1) atomicAdd(&s_rows[j], (unsigned long long)(cnt0 + cnt1));
        // the two rows share the anchor: its bit is set once
        if (cnt0) set_member_bits<kArr>(bt, cd.seg_off, rk0, km1);
        if (cnt1) set_member_bits<kArr>(bt, cd.seg_off, rk1, km1, cnt0 ? 1u : 0u);
    }
}

// Level-k count of parent rows i and i + 1 (those below end) with the wave
// converged: all 64 lanes call together, each with its own pair.  When the
// two rows share the anchor one pass over the anchor's list tests both;
// otherwise each row's list is walked in turn.  The members' bits are
// combined over the wave as in joink_count_row_multi.  Results equal two
// calls of joink_count_row_multi.
template <int KM1, bool LDSB = false>
__device__ inline void joink_count_rows2w(const DevData &D, const CandDesc *descs, uint32_t ncand, uint64_t i,
                                          uint64_t end, uint32_t *slots, uint32_t *bits,
                                          unsigned long long *s_rows, uint32_t *s_bits = nullptr,
                                          const LdsWin (*s_win)[kWinCols] = nullptr) {
    const CandDesc &c0 = descs[0];
    const uint32_t km1 = KM1 ? (uint32_t)KM1 : c0.k - 1;
    constexpr int kArr = KM1 ? KM1 : kMaxK;
    const bool act0 = i < end, act1 = i + 1 < end;
    uint32_t rk0[kArr], rk1[kArr];
    float mx0[kArr], my0[kArr], mx1[kArr], my1[kArr];
#pragma unroll
    for (uint32_t m = 0; m < (uint32_t)kArr; m++) {
        if (m >= km1) break;
        rk0[m] = rk1[m] = 0;
        mx0[m] = my0[m] = mx1[m] = my1[m] = 0.f;
        const uint32_t fs = D.fstart[c0.f[m]];
        if (act0) {
            const uint32_t id = c0.parent[i * km1 + m];
            rk0[m] = id - fs;
            if (m) {
                mx0[m] = D.sxy[id].x;
                my0[m] = D.sxy[id].y;
            }
        }
        if (act1) {
            const uint32_t id = c0.parent[(i + 1) * km1 + m];
            rk1[m] = id - fs;
            if (m) {
                mx1[m] = D.sxy[id].x;
                my1[m] = D.sxy[id].y;
            }
        }
    }
    const bool shared = act1 && rk0[0] == rk1[0];
    const bool own1 = act1 && !shared;  // row i + 1 walks its own list
    const uint32_t t0 = rk0[0], t1 = rk1[0];
    uint64_t na0 = 0, na1 = 0, nb0 = 0, nb1 = 0;
    if (act0) {
        nb_bounds(c0.nb_off, c0.nb_base, t0, na0, na1);
    }
    if (own1) {
        nb_bounds(c0.nb_off, c0.nb_base, t1, nb0, nb1);
    }
    for (uint32_t j = 0; j < ncand; j++) {
        const CandDesc &cd = descs[j];
        const uint64_t a0 = na0, a1 = na1, e0 = nb0, e1 = nb1;
        if (j + 1 < ncand) {
            const CandDesc &cn = descs[j + 1];
            if (act0) {
                nb_bounds(cn.nb_off, cn.nb_base, t0, na0, na1);
            }
            if (own1) {
                nb_bounds(cn.nb_off, cn.nb_base, t1, nb0, nb1);
            }
        }
        const gcu32 *nbt = as_global(cd.nb_table);
        const uint32_t sl = D.fstart[cd.f[km1]];
        uint32_t *bt = bits + cd.bits_off;
        uint32_t cnt0 = 0, cnt1 = 0;
        BitAcc acc(bt + cd.seg_off[km1]);
        const LdsWin wl = LDSB ? s_win[j][km1] : LdsWin{0, 0};
        // pass 1: the anchor list of row i, tested against row i and, when
        // shared, row i + 1; pass 2: row i + 1's own list
        auto walk = [&](uint64_t b0, uint64_t b1, bool use0, bool use1) {
            auto test = [&](uint32_t q, float qx, float qy) {
                bool ok0 = use0, ok1 = use1;
#pragma unroll
                for (uint32_t m = 1; m < (uint32_t)kArr; m++) {
                    if (m >= km1) break;
                    if (sqdist(mx0[m], my0[m], qx, qy) > D.g.r2) ok0 = false;
                    if (sqdist(mx1[m], my1[m], qx, qy) > D.g.r2) ok1 = false;
                }
                cnt0 += ok0;
                cnt1 += ok1;
                if (ok0 || ok1) {
                    if (LDSB && win_words(wl)) {
                        const uint32_t r = q - sl, w = (r >> 5) - wl.gword;
                        if (w < win_words(wl))
                            atomicOr(&s_bits[win_base(wl) + w], 1u << (r & 31));
                        else
                            acc.set(r);
                    } else {
                        acc.set(q - sl);
                    }
                }
            };
            uint64_t e = b0;
            for (; e + kJoinUnroll <= b1; e += kJoinUnroll) {
                uint32_t q[kJoinUnroll];
                float qx[kJoinUnroll], qy[kJoinUnroll];
#pragma unroll
                for (int u = 0; u < kJoinUnroll; u++) q[u] = nbt[e + u];
#pragma unroll
                for (int u = 0; u < kJoinUnroll; u++) {
                    qx[u] = D.sxy[q[u]].x;
                    qy[u] = D.sxy[q[u]].y;
                }
#pragma unroll
                for (int u = 0; u < kJoinUnroll; u++) test(q[u], qx[u], qy[u]);
            }
            for (; e < b1; e++) {
                const uint32_t q = nbt[e];
                test(q, D.sxy[q].x, D.sxy[q].y);
            }
        };
        walk(a0, a1, act0, shared);
        walk(e0, e1, false, own1);
        acc.flush();
        if (cd.slots_off != kNoSlots) {
            if (act0) slots[cd.slots_off + i] = cnt0;
            if (act1) slots[cd.slots_off + i + 1] = cnt1;
        }
        if (cnt0 + cnt1) atomicAdd(&s_rows[j], (unsigned long long)(cnt0 + cnt1));
        if (LDSB) {
            if (__ballot(cnt0 + cnt1 != 0)) {
                lds_member_bits<kArr>(s_bits, s_win[j], bt, cd.seg_off, rk0, km1, act0 && cnt0 != 0);
                lds_member_bits<kArr>(s_bits, s_win[j], bt, cd.seg_off, rk1, km1, act1 && cnt1 != 0);
            }
        } else if (__ballot(cnt0 + cnt1 != 0)) {
            wave_member_bits<kArr>(bt, cd.seg_off, rk0, km1, cnt0 != 0);
            wave_member_bits<kArr>(bt, cd.seg_off, rk1, km1, cnt1 != 0);
        }
    }
}

// Level-k count of R adjacent parent rows i .. i + R - 1 with the same
// anchor: every candidate's list is read once and each entry tested against
// all R rows (results equal R calls of joink_count_row_multi).  Parent rows
// of deep levels come in long runs per anchor, so one list read serves
// several rows.
template <int KM1, int R>
__device__ inline void joink_count_rowsR(const DevData &D, const CandDesc *descs, uint32_t ncand, uint64_t i,
                                         uint32_t *slots, uint32_t *bits, unsigned long long *s_rows) {
    const CandDesc &c0 = descs[0];
    const uint32_t km1 = KM1 ? (uint32_t)KM1 : c0.k - 1;
    constexpr int kArr = KM1 ? KM1 : kMaxK;
    uint32_t rk[R][kArr];
    float mx[R][kArr], my[R][kArr];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t *row = c0.parent + (i + r) * km1;
#pragma unroll
        for (uint32_t m = 0; m < (uint32_t)kArr; m++) {
            if (m >= km1) break;
            const uint32_t id = row[m];
            rk[r][m] = id - D.fstart[c0.f[m]];
            if (m) {
                mx[r][m] = D.sxy[id].x;
                my[r][m] = D.sxy[id].y;
            }
        }
    }
    const uint32_t t = rk[0][0];
    uint64_t nb0, nb1;
    nb_bounds(c0.nb_off, c0.nb_base, t, nb0, nb1);
    for (uint32_t j = 0; j < ncand; j++) {
        const CandDesc &cd = descs[j];
        const uint64_t b0 = nb0, b1 = nb1;
        if (j + 1 < ncand) {
            const CandDesc &cn = descs[j + 1];
            nb_bounds(cn.nb_off, cn.nb_base, t, nb0, nb1);
        }
        const gcu32 *nbt = as_global(cd.nb_table);
        const uint32_t sl = D.fstart[cd.f[km1]];
        uint32_t *bt = bits + cd.bits_off;
        uint32_t cnt[R];
#pragma unroll
        for (int r = 0; r < R; r++) cnt[r] = 0;
        BitAcc acc(bt + cd.seg_off[km1]);
        auto test = [&](uint32_t q, float qx, float qy) {
            bool any = false;
#pragma unroll
            for (int r = 0; r < R; r++) {
                bool ok = true;
#pragma unroll
                for (uint32_t m = 1; m < (uint32_t)kArr; m++) {
                    if (m >= km1) break;
                    if (sqdist(mx[r][m], my[r][m], qx, qy) > D.g.r2) ok = false;
                }
                cnt[r] += ok;
                any |= ok;
            }
            if (any) acc.set(q - sl);
        };
        uint64_t e = b0;
        for (; e + kJoinUnroll <= b1; e += kJoinUnroll) {
            uint32_t q[kJoinUnroll];
            float qx[kJoinUnroll], qy[kJoinUnroll];
#pragma unroll
            for (int u = 0; u < kJoinUnroll; u++) q[u] = nbt[e + u];
#pragma unroll
            for (int u = 0; u < kJoinUnroll; u++) {
                qx[u] = D.sxy[q[u]].x;
                qy[u] = D.sxy[q[u]].y;
            }
#pragma unroll
            for (int u = 0; u < kJoinUnroll; u++) test(q[u], qx[u], qy[u]);
        }
        for (; e < b1; e++) {
            const uint32_t q = nbt[e];
            test(q, D.sxy[q].x, D.sxy[q].y);
        }
        acc.flush();
        uint32_t tot = 0;
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (cd.slots_off != kNoSlots) slots[cd.slots_off + i + r] = cnt[r];
            tot += cnt[r];
        }
        if (tot) atomicAdd(&s_rows[j], (unsigned long long)tot);
        // the R rows share the anchor: its bit is set with the first row
        // that has an extension, the other members per row
        bool first = true;
#pragma unroll
        for (int r = 0; r < R; r++)
            if (cnt[r]) {
                set_member_bits<kArr>(bt, cd.seg_off, rk[r], km1, first ? 0u : 1u);
                first = false;
            }
    }
}

// Sets bit `rank` of `base` for every lane with `hit`, one atomic per
// distinct 32-bit word.  Ranks must be ascending over the participating
// lanes (list entries are sorted positions), so lanes of one word are
// adjacent: a segmented OR over each run of equal words leaves the run's
// mask in its first lane.
__device__ inline void wave_set_bits(uint32_t *base, uint32_t rank, bool act, bool hit) {
    const uint32_t lane = __lane_id();
    const uint32_t w = act ? (rank >> 5) : 0xffffffffu;
    uint32_t msk = act && hit ? (1u << (rank & 31)) : 0u;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t ow = __shfl_down(w, off, 64);
        const uint32_t om = __shfl_down(msk, off, 64);
        if (lane + off < 64 && ow == w) msk |= om;
    }
    const uint32_t pw = __shfl_up(w, 1, 64);
    if (act && msk && (lane == 0 || pw != w)) {
        uint32_t *p = base + w;
        if ((__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & msk) != msk) atomicOr(p, msk);
    }
}

// Wave-cooperative level-k join over parent rows [r0, r0 + n), n <= 64, of
// the parent table shared by the ncand candidates descs[0..ncand).  Rows
// with the same anchor (adjacent in the parent table) form a group; the
// candidates for the last feature of every row of a group are the anchor's
// neighbour list of that feature.  The wave spreads the list over its lanes
// and steps through the group's rows, so each lane tests one (row, entry)
// pair per step against the row's other members; lists of at most 32
// entries are packed 64 / m rows per step.  Results equal the per-row
// passes (joink_count_row_multi / joink_write_row_multi): COUNT writes
// each row's slot and the participation bits, WRITE emits each row's
// extensions in list order at the row's offset.
constexpr uint32_t kRegChunks = 4;  // list chunks of 64 kept in registers

template <bool WRITE, int KM1>
__device__ void wave_rows_multi(const DevData &D, const CandDesc *descs, uint32_t ncand, uint64_t r0, uint32_t n,
                                uint32_t *slots, uint32_t *bits, unsigned long long *s_rows) {
    const uint32_t lane = __lane_id();
    const uint64_t lt = (1ull << lane) - 1;
    const CandDesc &c0 = descs[0];
    const uint32_t km1 = KM1 ? (uint32_t)KM1 : c0.k - 1, k = km1 + 1;
    constexpr int kArr = KM1 ? KM1 : kMaxK;
    const uint32_t *parent = c0.parent;
    const uint32_t fa0 = D.fstart[c0.f[0]];
    const float r2 = D.g.r2;
    const uint32_t anc = lane < n ? parent[(r0 + lane) * km1] : 0xffffffffu;
    const uint32_t prv = __shfl_up(anc, 1, 64);
    uint64_t starts = __ballot(lane < n && (lane == 0 || anc != prv));
    while (starts) {
        const uint32_t gs = (uint32_t)__builtin_ctzll(starts);
        starts &= starts - 1;
        const uint32_t ge = starts ? (uint32_t)__builtin_ctzll(starts) : n;
        const uint32_t nrows = ge - gs;
        const uint32_t t = __shfl(anc, gs, 64) - fa0;
        for (uint32_t j = 0; j < ncand; j++) {
            const CandDesc &cd = descs[j];
            uint64_t b0, b1;
            nb_bounds(cd.nb_off, cd.nb_base, t, b0, b1);
            const uint32_t m = (uint32_t)(b1 - b0);
            const uint32_t sl = D.fstart[cd.f[km1]];
            uint32_t *bt = WRITE ? nullptr : bits + cd.bits_off;
            const uint32_t *wsl = WRITE ? (cd.wslots ? cd.wslots : slots + cd.slots_off) : nullptr;
            if (m == 0) {
                if (!WRITE && lane < nrows && cd.slots_off != kNoSlots) slots[cd.slots_off + r0 + gs + lane] = 0;
                continue;
            }
            if (m <= 32) {
                // packed: lane = (row offset ro, list entry e)
                const uint32_t R = 64 / m;
                const uint32_t ro = lane / m, e = lane - ro * m;
                const bool act = ro < R;
                const uint64_t segm = ((1ull << m) - 1) << (act ? ro * m : 0);
                uint32_t q = 0;
                float qx = 0.f, qy = 0.f;
                if (act) {
                    q = cd.nb_table[b0 + e];
                    qx = D.sxy[q].x;
                    qy = D.sxy[q].y;
                }
                bool hit = false;
                for (uint32_t s0 = 0; s0 < nrows; s0 += R) {
                    const uint32_t rl = s0 + ro;
                    const bool valid = act && rl < nrows;
                    const uint64_t row = r0 + gs + rl;
                    const uint32_t *rp = parent + (valid ? row : r0) * km1;
                    bool ok = valid;
                    for (uint32_t mm = 1; mm < km1 && ok; mm++) {
                        const uint32_t id = rp[mm];
                        if (sqdist(D.sxy[id].x, D.sxy[id].y, qx, qy) > r2) ok = false;
                    }
                    const uint64_t bal = __ballot(ok);
                    if (!WRITE) {
                        const uint32_t cnt = (uint32_t)__popcll(bal & segm);
                        if (valid && e == 0) {
                            if (cd.slots_off != kNoSlots) slots[cd.slots_off + row] = cnt;
                            if (cnt) atomicAdd(&s_rows[j], (unsigned long long)cnt);
                            if (cnt)
                                for (uint32_t mm = 0; mm < km1; mm++)
                                    set_bit(bt + cd.seg_off[mm], rp[mm] - D.fstart[cd.f[mm]]);
                        }
                        hit |= ok;
                    } else if (ok) {
                        const uint32_t rank = (uint32_t)__popcll(bal & segm & lt);
                        if (rank < wsl[row]) {
                            uint32_t *out = cd.table + (cd.offsets[row] - cd.off_base + rank) * k;
                            for (uint32_t mm = 0; mm < km1; mm++) out[mm] = rp[mm];
                            out[km1] = q;
                        }
                    }
                }
                if (!WRITE) {
                    // fold the row offsets' hits onto entry e, then one
                    // participation bit per entry
                    const uint64_t hb = __ballot(hit);
                    uint64_t fold = 0;
                    for (uint32_t x = 0; x < R; x++) fold |= (hb >> (x * m)) & ((1ull << m) - 1);
                    wave_set_bits(bt + cd.seg_off[km1], q - sl, lane < m, (fold >> lane) & 1);
                }
                continue;
            }
            // chunked: the list in chunks of 64 lanes, the group's rows in turn
            const uint32_t nch = (m + 63) / 64;
            uint32_t qr[kRegChunks];
            float xr[kRegChunks], yr[kRegChunks];
            bool hr[kRegChunks];
#pragma unroll
            for (uint32_t c = 0; c < kRegChunks; c++) {
                const uint32_t e = c * 64 + lane;
                qr[c] = 0;
                xr[c] = yr[c] = 0.f;
                hr[c] = false;
                if (e < m) {
                    qr[c] = cd.nb_table[b0 + e];
                    xr[c] = D.sxy[qr[c]].x;
                    yr[c] = D.sxy[qr[c]].y;
                }
            }
            for (uint32_t rl = 0; rl < nrows; rl++) {
                const uint64_t row = r0 + gs + rl;
                const uint32_t *rp = parent + row * km1;
                uint32_t nr = 0;
                uint64_t base = 0;
                if (WRITE) {
                    nr = wsl[row];
                    if (!nr) continue;
                    base = cd.offsets[row] - cd.off_base;
                }
                float mx[kArr], my[kArr];
#pragma unroll
                for (uint32_t mm = 1; mm < (uint32_t)kArr; mm++) {
                    if (mm >= km1) break;
                    const uint32_t id = rp[mm];
                    mx[mm] = D.sxy[id].x;
                    my[mm] = D.sxy[id].y;
                }
                uint32_t cnt = 0;
                auto step = [&](uint32_t e, uint32_t q, float qx, float qy, bool &h) {
                    bool ok = e < m;
#pragma unroll
                    for (uint32_t mm = 1; mm < (uint32_t)kArr; mm++) {
                        if (mm >= km1) break;
                        if (sqdist(mx[mm], my[mm], qx, qy) > r2) ok = false;
                    }
                    const uint64_t bal = __ballot(ok);
                    if (!WRITE) {
                        h |= ok;
                    } else if (ok) {
                        const uint32_t rank = cnt + (uint32_t)__popcll(bal & lt);
                        if (rank < nr) {
                            uint32_t *out = cd.table + (base + rank) * k;
                            for (uint32_t mm = 0; mm < km1; mm++) out[mm] = rp[mm];
                            out[km1] = q;
                        }
                    }
                    cnt += (uint32_t)__popcll(bal);
                };
#pragma unroll
                for (uint32_t c = 0; c < kRegChunks; c++)
                    if (c < nch) step(c * 64 + lane, qr[c], xr[c], yr[c], hr[c]);
                for (uint32_t c = kRegChunks; c < nch; c++) {
                    const uint32_t e = c * 64 + lane;
                    uint32_t q = 0;
                    float qx = 0.f, qy = 0.f;
                    if (e < m) {
                        q = cd.nb_table[b0 + e];
                        qx = D.sxy[q].x;
                        qy = D.sxy[q].y;
                    }
                    bool h = false;
                    step(e, q, qx, qy, h);
                    if (!WRITE) wave_set_bits(bt + cd.seg_off[km1], q - sl, e < m, h);
                }
                if (!WRITE) {
                    if (lane == 0) {
                        if (cd.slots_off != kNoSlots) slots[cd.slots_off + row] = cnt;
                        if (cnt) atomicAdd(&s_rows[j], (unsigned long long)cnt);
                    }
                    if (cnt && lane < km1) set_bit(bt + cd.seg_off[lane], rp[lane] - D.fstart[cd.f[lane]]);
                }
            }
            if (!WRITE) {
#pragma unroll
                for (uint32_t c = 0; c < kRegChunks; c++)
                    if (c < nch) wave_set_bits(bt + cd.seg_off[km1], qr[c] - sl, c * 64 + lane < m, hr[c]);
            }
        }
    }
}

// Runs f(t, true) for the anchors t of [begin, end) that this wavefront
// owns (its 64-anchor chunks, one every kBlock anchors) and for which
// keep(t) holds, with those anchors packed onto the wave's lanes in
// ascending order: they are queued in LDS (q: 128 entries for this wave)
// and run 64 at a time, the last < 64 at the end, where the lanes without
// one call f(0, false) (f is always called by the whole wave); skip(t) is
// called for the others.  Every lane of the wave calls it with the same
// range.
template <typename P, typename S, typename F>
__device__ inline void wave_packed(uint64_t begin, uint64_t end, uint32_t *q, P &&keep, S &&skip, F &&f) {
    const uint32_t lane = __lane_id();
    const uint64_t lt = __lanemask_lt();
    uint32_t qn = 0;  // wave-uniform
    for (uint64_t i0 = begin + (threadIdx.x & ~63u); i0 < end; i0 += kBlock) {
        const uint64_t t = i0 + lane;
        const bool valid = t < end;
        const bool kp = valid && keep((uint32_t)t);
        if (valid && !kp) skip((uint32_t)t);
        const uint64_t bal = __ballot(kp);
        if (kp) q[qn + (uint32_t)__popcll(bal & lt)] = (uint32_t)t;
        qn += (uint32_t)__popcll(bal);
        __builtin_amdgcn_wave_barrier();
        if (qn >= 64) {
            const uint32_t mine = q[lane];
            const uint32_t rest = lane < qn - 64 ? q[64 + lane] : 0u;
            __builtin_amdgcn_wave_barrier();
            if (lane < qn - 64) q[lane] = rest;
            qn -= 64;
            __builtin_amdgcn_wave_barrier();
            f(mine, true);
        }
    }
    if (qn) f(lane < qn ? q[lane] : 0u, lane < qn);
}

// Level 2: the item's anchor range over the block's lanes; the item's rows
// are summed in the workgroup and added to the candidate's total.  VAR 0:
// one anchor per thread; VAR 1: the anchors whose cell's 3 x 3 mask holds b
// packed onto the wave's lanes (wave_packed), the others given zero slots.
template <int VAR>
__global__ __launch_bounds__(kBlock) void k_count2_batch(DevData D, const CandDesc *desc, const WorkItem *items,
                                                         const uint64_t *seg, uint32_t *slots, uint32_t *bits,
                                                         unsigned long long *rows, uint32_t *queue) {
    __shared__ uint32_t s_q[VAR == 1 ? kBlock / 64 : 1][128];
    __shared__ unsigned long long s_bc[VAR == 1 ? kBitCache : 1];
    if (VAR == 1)
        for (uint32_t i = threadIdx.x; i < kBitCache; i += kBlock) s_bc[i] = 0xffffffffull;
    __syncthreads();
    uint64_t it;
    while (next_item(queue, seg, it)) {
        const WorkItem wi = items[it];
        const CandDesc &cd = desc[wi.cand];
        uint32_t *sl = cd.slots_off != kNoSlots ? slots + cd.slots_off : nullptr;
        uint32_t *bt = bits + cd.bits_off;
        const uint32_t a = cd.f[0], b = cd.f[1];
        uint64_t sum = 0;
        if (VAR == 1) {
            const uint32_t sa = D.fstart[a];
            wave_packed(
                wi.begin, wi.end, s_q[VAR == 1 ? threadIdx.x >> 6 : 0],
                [&](uint32_t t) {
                    return !D.mask3 ||
                           ((D.mask3[(uint64_t)D.scell[sa + t] * D.mask_w + (b >> 6)] >> (b & 63)) & 1);
                },
                [&](uint32_t t) {
                    if (sl) sl[t] = 0;
                },
                [&](uint32_t t, bool valid) {
                    const uint32_t cnt =
                        valid ? join2_count_flat<false, 2>(D, a, b, t, sl, nullptr,
                                                        BitAccC(bt + cd.seg_off[1], bits, s_bc))
                              : 0u;
                    sum += cnt;
                    // the anchors' own bits, one check and atomic per word
                    // (the packed anchors ascend over the lanes)
                    wave_set_bits(bt + cd.seg_off[0], t, valid, cnt != 0);
                });
        } else {
            for (uint64_t i = wi.begin + threadIdx.x; i < wi.end; i += kBlock)
                sum += join2_count_flat(D, a, b, (uint32_t)i, sl, bt + cd.seg_off[0], BitAcc(bt + cd.seg_off[1]));
        }
        uint64_t ex;
        const uint64_t tot = block_excl_scan(sum, ex);
        if (threadIdx.x == 0 && tot) atomicAdd(&rows[wi.cand], (unsigned long long)tot);
    }
}

// Level k >= 3 (KM1 = k - 1; 0 for any k).  VAR 0: one parent row per
// lane, all of a run's candidates per row, waves converged and member bits
// combined over the wave; VAR 1: wave-cooperative (wave_rows_multi) over
// 64-row chunks taken by the block's waves in turn; VAR 2: one parent row
// per thread, the item's candidates one after another (one neighbour-list
// stream at a time); VAR 3: two adjacent rows per thread, their list read
// once when they share the anchor; VAR 6: four adjacent rows per thread,
// one list pass when all four share the anchor; VAR 8: two adjacent rows
// per lane, waves converged, member bits combined over the wave.  Rows per
// candidate are summed in LDS over the item and added to the candidates'
// totals.  VAR 4 and 5 run the code of VAR 0 and 3 with the registers
// capped at 64 per lane, for 8 waves per SIMD; VAR 7 and 9 the code of
// VAR 6 and 8 at 102 per lane, for 5 waves per SIMD (the compiler spills
// what does not fit).
constexpr int var_waves(int v) { return v == 4 || v == 5 || v == 10 ? 8 : v == 7 || v == 9 || v == 11 ? 5 : 1; }
// VAR 10-13: the code of VAR 4, 9, 0 and 8 with each item's participation
// bits staged in LDS (lds_setup / lds_flush), for k = 3..8.
constexpr bool var_lds(int v) { return v >= 10 && v <= 13; }

template <int KM1, int VAR>
__global__ __launch_bounds__(kBlock, var_waves(VAR)) void k_countk_batch(
    DevData D, const CandDesc *desc, const WorkItem *items, const uint64_t *seg, uint32_t *slots, uint32_t *bits,
    unsigned long long *rows, uint32_t *queue) {
    __shared__ unsigned long long s_rows[kMaxRun];
    constexpr bool LDS = var_lds(VAR) && KM1 >= 2 && KM1 + 1 <= (int)kWinCols;
    __shared__ uint32_t s_bits[LDS ? kLdsBitWords : 1];
    __shared__ LdsWin s_win[LDS ? kMaxRun : 1][kWinCols];
    __shared__ uint32_t s_used, s_mlo[kWinCols], s_mhi[kWinCols];
    uint64_t it;
    while (next_item(queue, seg, it)) {
        const WorkItem wi = items[it];
        if (threadIdx.x < kMaxRun) s_rows[threadIdx.x] = 0;
        __syncthreads();
        if (LDS)
            lds_setup(D, desc + wi.cand, wi.ncand, KM1 + 1, wi.begin, wi.end, s_bits, s_win, &s_used, s_mlo, s_mhi);
        if (VAR == 0 || VAR == 4 || VAR == 10 || VAR == 12) {
            // waves stay converged over the rows so the member bits can be
            // combined over the wave
            for (uint64_t i0 = wi.begin + (threadIdx.x & ~63u); i0 < wi.end; i0 += kBlock) {
                const uint64_t i = i0 + __lane_id();
                joink_count_row_multi<KM1, true, LDS>(D, desc + wi.cand, wi.ncand, i, slots, bits, s_rows,
                                                      i < wi.end, s_bits, s_win);
            }
        } else if (VAR == 2) {
            for (uint32_t j = 0; j < wi.ncand; j++)
                for (uint64_t i = wi.begin + threadIdx.x; i < wi.end; i += kBlock)
                    joink_count_row_multi<KM1>(D, desc + wi.cand + j, 1, i, slots, bits, s_rows + j);
        } else if (VAR == 8 || VAR == 9 || VAR == 11 || VAR == 13) {
            // two adjacent rows per lane, waves converged
            for (uint64_t i0 = wi.begin + 2 * (uint64_t)(threadIdx.x & ~63u); i0 < wi.end; i0 += 2 * kBlock)
                joink_count_rows2w<KM1, LDS>(D, desc + wi.cand, wi.ncand, i0 + 2 * __lane_id(), wi.end, slots, bits,
                                             s_rows, s_bits, s_win);
        } else if (VAR == 6 || VAR == 7) {
            // four adjacent rows per thread, one list pass when they share the
            // anchor, else pairs as in VAR 3
            const CandDesc &c0 = desc[wi.cand];
            const uint32_t km1 = c0.k - 1;
            for (uint64_t i = wi.begin + 4 * (uint64_t)threadIdx.x; i < wi.end; i += 4 * kBlock) {
                if (i + 3 < wi.end && c0.parent[i * km1] == c0.parent[(i + 3) * km1]) {
                    joink_count_rowsR<KM1, 4>(D, desc + wi.cand, wi.ncand, i, slots, bits, s_rows);
                } else {
                    for (uint64_t h = i; h < i + 4 && h < wi.end; h += 2) {
                        if (h + 1 < wi.end && h + 1 < i + 4 && c0.parent[h * km1] == c0.parent[(h + 1) * km1]) {
                            joink_count_rows2<KM1>(D, desc + wi.cand, wi.ncand, h, slots, bits, s_rows);
                        } else {
                            joink_count_row_multi<KM1>(D, desc + wi.cand, wi.ncand, h, slots, bits, s_rows);
                            if (h + 1 < wi.end && h + 1 < i + 4)
                                joink_count_row_multi<KM1>(D, desc + wi.cand, wi.ncand, h + 1, slots, bits, s_rows);
                        }
                    }
                }
            }
        } else if (VAR == 3 || VAR == 5) {
            // two adjacent rows per thread, one list pass when they share the anchor
            const CandDesc &c0 = desc[wi.cand];
            const uint32_t km1 = c0.k - 1;
            for (uint64_t i = wi.begin + 2 * (uint64_t)threadIdx.x; i < wi.end; i += 2 * kBlock) {
                if (i + 1 < wi.end && c0.parent[i * km1] == c0.parent[(i + 1) * km1]) {
                    joink_count_rows2<KM1>(D, desc + wi.cand, wi.ncand, i, slots, bits, s_rows);
                } else {
                    joink_count_row_multi<KM1>(D, desc + wi.cand, wi.ncand, i, slots, bits, s_rows);
                    if (i + 1 < wi.end)
                        joink_count_row_multi<KM1>(D, desc + wi.cand, wi.ncand, i + 1, slots, bits, s_rows);
                }
            }
        } else {
            for (uint64_t c0 = wi.begin + (threadIdx.x & ~63u); c0 < wi.end; c0 += kBlock) {
                const uint32_t n = (uint32_t)(wi.end - c0 < 64 ? wi.end - c0 : 64);
                wave_rows_multi<false, KM1>(D, desc + wi.cand, wi.ncand, c0, n, slots, bits, s_rows);
            }
        }
        __syncthreads();
        if (LDS) lds_flush(desc + wi.cand, wi.ncand, KM1 + 1, bits, s_bits, s_win);
        if (threadIdx.x < wi.ncand && s_rows[threadIdx.x]) atomicAdd(&rows[wi.cand + threadIdx.x], s_rows[threadIdx.x]);
    }
}

// Level-k write over the rows of one parent table for ncand tables.
// off1 (one candidate only): the row's table offset, already known (write
// variant 3); otherwise it is read from the candidate's row offsets.
template <int KM1>
__device__ inline void joink_write_row_multi(const DevData &D, const CandDesc *descs, uint32_t ncand, uint64_t i,
                                             const uint32_t *slots, uint64_t off1 = ~0ull) {
    const CandDesc &c0 = descs[0];
    const uint32_t km1 = KM1 ? (uint32_t)KM1 : c0.k - 1, k = km1 + 1;
    constexpr int kArr = KM1 ? KM1 : kMaxK;
    const gcu32 *row = as_global(c0.parent + i * km1);
    uint32_t ids[kArr];
    float mx[kArr], my[kArr];
    auto slot_of = [&](const CandDesc &cd) { return as_global(cd.wslots ? cd.wslots : slots + cd.slots_off)[i]; };
    // candidates with rows from this parent row (a run holds at most 64)
    uint64_t todo = 0;
    for (uint32_t j = 0; j < ncand; j++)
        if (slot_of(descs[j])) todo |= 1ull << j;
    if (!todo) return;
#pragma unroll
    for (uint32_t m = 0; m < (uint32_t)kArr; m++) {
        if (m >= km1) break;
        ids[m] = row[m];
        if (m) {
            mx[m] = D.sxy[ids[m]].x;
            my[m] = D.sxy[ids[m]].y;
        }
    }
    const uint32_t t = ids[0] - D.fstart[c0.f[0]];
    // list bounds loaded one candidate ahead (the candidate index stays
    // wave-uniform, so descriptor fields remain scalar loads)
    uint64_t nb0 = 0, nb1 = 0;
    if (todo & 1) {
        nb_bounds(c0.nb_off, c0.nb_base, t, nb0, nb1);
    }
    for (uint32_t j = 0; j < ncand; j++) {
        const uint64_t b0 = nb0, b1 = nb1;
        if (j + 1 < ncand && ((todo >> (j + 1)) & 1)) {
            const CandDesc &cn = descs[j + 1];
            nb_bounds(cn.nb_off, cn.nb_base, t, nb0, nb1);
        }
        if (!((todo >> j) & 1)) continue;
        const CandDesc &cd = descs[j];
        const uint32_t nrows = slot_of(cd);
        const gcu32 *nbt = as_global(cd.nb_table);
        gu32 *out = as_global_out(cd.table + (off1 != ~0ull ? off1 : cd.offsets[i] - cd.off_base) * k);
        uint32_t w = 0;
        auto test = [&](uint32_t q, float qx, float qy) {
            bool ok = true;
#pragma unroll
            for (uint32_t m = 1; m < (uint32_t)kArr; m++) {
                if (m >= km1) break;
                if (sqdist(mx[m], my[m], qx, qy) > D.g.r2) ok = false;
            }
            if (ok) {
                if (w < nrows) {
#pragma unroll
                    for (uint32_t m = 0; m < (uint32_t)kArr; m++) {
                        if (m >= km1) break;
                        out[(uint64_t)w * k + m] = ids[m];
                    }
                    out[(uint64_t)w * k + km1] = q;
                }
                w++;
            }
        };
        uint64_t e = b0;
        for (; e + kWriteUnroll <= b1 && w < nrows; e += kWriteUnroll) {
            uint32_t q[kWriteUnroll];
            float qx[kWriteUnroll], qy[kWriteUnroll];
#pragma unroll
            for (int u = 0; u < kWriteUnroll; u++) q[u] = nbt[e + u];
#pragma unroll
            for (int u = 0; u < kWriteUnroll; u++) {
                qx[u] = D.sxy[q[u]].x;
                qy[u] = D.sxy[q[u]].y;
            }
#pragma unroll
            for (int u = 0; u < kWriteUnroll; u++) test(q[u], qx[u], qy[u]);
        }
        for (; e < b1 && w < nrows; e++) {
            const uint32_t q = nbt[e];
            test(q, D.sxy[q].x, D.sxy[q].y);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_write2_batch(DevData D, const CandDesc *desc, const WorkItem *items,
                                                         const uint64_t *seg, const uint32_t *slots, uint32_t *queue) {
    uint64_t it;
    while (next_item(queue, seg, it)) {
        const WorkItem wi = items[it];
        const CandDesc &cd = desc[wi.cand];
        const gcu32 *sl = as_global(cd.wslots ? cd.wslots : slots + cd.slots_off);
        const gcu64 *offs = as_global(cd.offsets);
        for (uint64_t t = wi.begin + threadIdx.x; t < wi.end; t += kBlock) {
            uint32_t nr = sl[t];
            if (nr)
                join2_write_flat(D, cd.f[0], cd.f[1], (uint32_t)t, nr, offs[t] - cd.off_base, cd.table, cd.nbr_out);
        }
    }
}

template <int KM1, int VAR>
__global__ __launch_bounds__(kBlock) void k_writek_batch(DevData D, const CandDesc *desc, const WorkItem *items,
                                                         const uint64_t *seg, const uint32_t *slots, uint32_t *queue) {
    uint64_t it;
    while (next_item(queue, seg, it)) {
        const WorkItem wi = items[it];
        if (VAR == 0) {
            for (uint64_t i = wi.begin + threadIdx.x; i < wi.end; i += kBlock)
                joink_write_row_multi<KM1>(D, desc + wi.cand, wi.ncand, i, slots);
        } else if (VAR == 2) {
            for (uint32_t j = 0; j < wi.ncand; j++)
                for (uint64_t i = wi.begin + threadIdx.x; i < wi.end; i += kBlock)
                    joink_write_row_multi<KM1>(D, desc + wi.cand + j, 1, i, slots);
        } else if (VAR == 3) {
            // offsets kept per 64 rows (items start at multiples of 64, so
            // a wave's rows are one chunk): a row's offset is its chunk's
            // plus the wave's exclusive sum of the slots before it
            const uint32_t lane = threadIdx.x & 63;
            for (uint32_t j = 0; j < wi.ncand; j++) {
                const CandDesc &cd = desc[wi.cand + j];
                const gcu32 *sl = as_global(cd.wslots ? cd.wslots : slots + cd.slots_off);
                const gcu64 *offs = as_global(cd.offsets);
                for (uint64_t c0 = wi.begin + (threadIdx.x & ~63u); c0 < wi.end; c0 += kBlock) {
                    const uint64_t i = c0 + lane;
                    const uint64_t nr = i < wi.end ? sl[i] : 0;
                    const uint64_t off = offs[c0 >> 6] - cd.off_base + wave_incl_scan(nr) - nr;
                    if (nr) joink_write_row_multi<KM1>(D, &cd, 1, i, slots, off);
                }
            }
        } else {
            for (uint64_t c0 = wi.begin + (threadIdx.x & ~63u); c0 < wi.end; c0 += kBlock) {
                const uint32_t n = (uint32_t)(wi.end - c0 < 64 ? wi.end - c0 : 64);
                wave_rows_multi<true, KM1>(D, desc + wi.cand, wi.ncand, c0, n, const_cast<uint32_t *>(slots),
                                           nullptr, nullptr);
            }
        }
    }
}

static unsigned persistent_grid(const colo_ctx *c, uint64_t n_items) {
    uint64_t g = (uint64_t)num_cus() * c->tune.blocks_per_cu;
    if (g > n_items) g = (n_items + kXcds - 1) / kXcds * kXcds;
    return (unsigned)(g ? g : kXcds);
}

// Kernel of a level: k = 2, a specialisation for k = 3..8, any k above.
template <template <int> class Pick>
static auto pick_k(uint32_t k) -> decltype(Pick<0>::get()) {
    switch (k) {
        case 3: return Pick<2>::get();
        case 4: return Pick<3>::get();
        case 5: return Pick<4>::get();
        case 6: return Pick<5>::get();
        case 7: return Pick<6>::get();
        case 8: return Pick<7>::get();
        default: return Pick<0>::get();
    }
}
template <int KM1> struct PickCount0 { static auto get() { return &k_countk_batch<KM1, 0>; } };
template <int KM1> struct PickCount1 { static auto get() { return &k_countk_batch<KM1, 1>; } };
template <int KM1> struct PickCount2 { static auto get() { return &k_countk_batch<KM1, 2>; } };
template <int KM1> struct PickCount3 { static auto get() { return &k_countk_batch<KM1, 3>; } };
template <int KM1> struct PickCount4 { static auto get() { return &k_countk_batch<KM1, 4>; } };
template <int KM1> struct PickCount5 { static auto get() { return &k_countk_batch<KM1, 5>; } };
template <int KM1> struct PickCount6 { static auto get() { return &k_countk_batch<KM1, 6>; } };
template <int KM1> struct PickCount7 { static auto get() { return &k_countk_batch<KM1, 7>; } };
template <int KM1> struct PickCount8 { static auto get() { return &k_countk_batch<KM1, 8>; } };
template <int KM1> struct PickCount9 { static auto get() { return &k_countk_batch<KM1, 9>; } };
template <int KM1> struct PickCount10 { static auto get() { return &k_countk_batch<KM1, 10>; } };
template <int KM1> struct PickCount11 { static auto get() { return &k_countk_batch<KM1, 11>; } };
template <int KM1> struct PickCount12 { static auto get() { return &k_countk_batch<KM1, 12>; } };
template <int KM1> struct PickCount13 { static auto get() { return &k_countk_batch<KM1, 13>; } };
template <int KM1> struct PickWrite0 { static auto get() { return &k_writek_batch<KM1, 0>; } };
template <int KM1> struct PickWrite1 { static auto get() { return &k_writek_batch<KM1, 1>; } };
template <int KM1> struct PickWrite2 { static auto get() { return &k_writek_batch<KM1, 2>; } };
template <int KM1> struct PickWrite3 { static auto get() { return &k_writek_batch<KM1, 3>; } };

int colo::count_batch(colo_ctx *c, uint32_t k, const CandDesc *d_desc, const WorkItem *d_items, const uint64_t *d_seg,
                      uint64_t n_items, uint32_t *d_slots, uint32_t *d_bits, uint64_t *d_rows, uint32_t *d_queue,
                      hipStream_t s) {
    if (!n_items) return COLO_OK;
    int rc = fill_u32(d_queue, kXcds, 0, s);
    if (rc) return rc;
    if (k == 2)
        hipLaunchKernelGGL(c->tune.count_var_of(2) % 2 ? k_count2_batch<1> : k_count2_batch<0>,
                           dim3(persistent_grid(c, n_items)), dim3(kBlock), 0, s, c->dev(), d_desc, d_items, d_seg,
                           d_slots, d_bits, (unsigned long long *)d_rows, d_queue);
    else {
        const uint32_t var = c->tune.count_var_of(k);
        hipLaunchKernelGGL(var == 13   ? pick_k<PickCount13>(k)
                           : var == 12 ? pick_k<PickCount12>(k)
                           : var == 11 ? pick_k<PickCount11>(k)
                           : var == 10 ? pick_k<PickCount10>(k)
                           : var == 9 ? pick_k<PickCount9>(k)
                           : var == 8 ? pick_k<PickCount8>(k)
                           : var == 7 ? pick_k<PickCount7>(k)
                           : var == 6 ? pick_k<PickCount6>(k)
                           : var == 5 ? pick_k<PickCount5>(k)
                           : var == 4 ? pick_k<PickCount4>(k)
                           : var == 3 ? pick_k<PickCount3>(k)
                           : var == 2 ? pick_k<PickCount2>(k)
                           : var == 1 ? pick_k<PickCount1>(k)
                                      : pick_k<PickCount0>(k),
                           dim3(persistent_grid(c, n_items)), dim3(kBlock), 0, s, c->dev(), d_desc, d_items, d_seg,
                           d_slots, d_bits, (unsigned long long *)d_rows, d_queue);
    }
    return hipGetLastError() == hipSuccess ? COLO_OK : set_error(c, COLO_EHIP, "count_batch launch");
}

int colo::write_batch(colo_ctx *c, uint32_t k, const CandDesc *d_desc, const WorkItem *d_items, const uint64_t *d_seg,
                      uint64_t n_items, const uint32_t *d_slots, uint32_t *d_queue, hipStream_t s) {
    if (!n_items) return COLO_OK;
    int rc = fill_u32(d_queue, kXcds, 0, s);
    if (rc) return rc;
    if (k == 2)
        hipLaunchKernelGGL(k_write2_batch, dim3(persistent_grid(c, n_items)), dim3(kBlock), 0, s, c->dev(), d_desc,
                           d_items, d_seg, d_slots, d_queue);
    else
        hipLaunchKernelGGL(c->tune.write_var == 3   ? pick_k<PickWrite3>(k)
                           : c->tune.write_var == 2 ? pick_k<PickWrite2>(k)
                           : c->tune.write_var == 1 ? pick_k<PickWrite1>(k) : pick_k<PickWrite0>(k),
                           dim3(persistent_grid(c, n_items)), dim3(kBlock), 0, s, c->dev(), d_desc, d_items, d_seg,
                           d_slots, d_queue);
    return hipGetLastError() == hipSuccess ? COLO_OK : set_error(c, COLO_EHIP, "write_batch launch");
}


// -------------------------------------------------------------------- PI
// One workgroup per bitmap segment: popcount reduction plus the two
// boundary bits the reference's off-by-one windows need (cu:113-124).

__global__ __launch_bounds__(kBlock) void k_pi_abs(const uint32_t *bits, const SegDesc *segs, SegOut *out) {
    const SegDesc sd = segs[blockIdx.x];
    const uint32_t *w = bits + sd.off;
    uint64_t s = 0;
    for (uint64_t i = threadIdx.x; i < sd.words; i += kBlock) s += __popc(w[i]);
    uint64_t ex;
    uint64_t tot = block_excl_scan(s, ex);
    if (threadIdx.x == 0) {
        SegOut o;
        o.pop = tot;
        o.first = (w[sd.first_bit >> 5] >> (sd.first_bit & 31)) & 1u;
        o.last = (w[sd.last_bit >> 5] >> (sd.last_bit & 31)) & 1u;
        out[blockIdx.x] = o;
    }
}

// ------------------------------------------------------------ subset bound
// Workgroups take fixed chunks of kBoundWords words of one task (a bound
// task or a copy range), found by binary search over the tasks' first
// chunk numbers: a hot feature's windows (hundreds of thousands of words)
// spread over many workgroups instead of one.

constexpr uint64_t kBoundWords = 4096;

__device__ inline uint64_t chunk_owner(const uint64_t *chunk_prefix, uint64_t n, uint64_t b) {
    uint64_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) >> 1;
        if (chunk_prefix[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kBlock) void k_subset_bound(const colo::BoundTask *tasks, const uint64_t *chunk_prefix,
                                                         uint64_t n, unsigned long long *out) {
    const uint64_t b = blockIdx.x;
    const uint64_t t = chunk_owner(chunk_prefix, n, b);
    const colo::BoundTask *tp = tasks + t;
    const uint32_t nw = tp->nw, words = tp->words, last = tp->last_bit;
    const uint64_t w0 = (b - chunk_prefix[t]) * kBoundWords;
    const uint64_t w1 = w0 + kBoundWords < words ? w0 + kBoundWords : words;
    uint64_t sum = 0;
    for (uint64_t i = w0 + threadIdx.x; i < w1; i += kBlock) {
        uint32_t a = tp->w[0][i];
        for (uint32_t j = 1; j < nw; j++) a &= tp->w[j][i];
        sum += __popc(a);
        if (i == (last >> 5)) out[2 * t + 1] = (a >> (last & 31)) & 1u;
    }
    uint64_t ex;
    const uint64_t tot = block_excl_scan(sum, ex);
    if (threadIdx.x == 0 && tot) atomicAdd(&out[2 * t], (unsigned long long)tot);
}

__global__ __launch_bounds__(kBlock) void k_gather_ranges(const uint32_t *src, uint32_t *dst,
                                                          const colo::CopyRange *ranges, const uint64_t *chunk_prefix,
                                                          uint64_t n) {
    const uint64_t b = blockIdx.x;
    const uint64_t t = chunk_owner(chunk_prefix, n, b);
    const colo::CopyRange r = ranges[t];
    const uint64_t w0 = (b - chunk_prefix[t]) * kBoundWords;
    const uint64_t w1 = w0 + kBoundWords < r.words ? w0 + kBoundWords : r.words;
    for (uint64_t i = w0 + threadIdx.x; i < w1; i += kBlock) dst[r.dst + i] = src[r.src + i];
}

__global__ __launch_bounds__(kBlock) void k_rank_to_sorted(const uint32_t *src, uint32_t N, const uint32_t *spos_f,
                                                           uint32_t base, uint32_t *dst) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    if ((src[i >> 5] >> (i & 31)) & 1u) {
        const uint32_t r = spos_f[i] - base;
        atomicOr(&dst[r >> 5], 1u << (r & 31));
    }
}

// Chunk numbers of n tasks of the given word counts (prefix[i] = first
// chunk of task i); returns the total.
template <typename T, typename W>
static uint64_t chunk_prefix_of(const T *v, uint64_t n, W words_of, std::vector<uint64_t> &prefix) {
    prefix.resize(n);
    uint64_t chunks = 0;
    for (uint64_t i = 0; i < n; i++) {
        prefix[i] = chunks;
        chunks += (words_of(v[i]) + kBoundWords - 1) / kBoundWords;
    }
    return chunks;
}

int colo::subset_bound_batch(colo_ctx *c, const BoundTask *tasks, uint64_t n, uint64_t *out, hipStream_t s) {
    colo_ctx *ctx_for_err = c;
    if (!n) return COLO_OK;
    for (uint64_t i = 0; i < n; i++)
        if (tasks[i].nw < 1 || tasks[i].nw > kMaxK - 1 || (tasks[i].words && tasks[i].last_bit >= 32ull * tasks[i].words))
            return set_error(c, COLO_EARG, "subset bound: malformed task");
    std::vector<uint64_t> prefix;
    const uint64_t chunks = chunk_prefix_of(tasks, n, [](const BoundTask &t) { return (uint64_t)t.words; }, prefix);
    if (chunks >= (1ull << 31)) return set_error(c, COLO_EUNSUPPORTED, "subset bound: too many chunks");
    const size_t tb = ((size_t)n * sizeof(BoundTask) + 15) & ~(size_t)15;
    void *tmp;
    int rc = scratch(c, tb + (size_t)n * 24, &tmp);
    if (rc) return rc;
    BoundTask *d_tasks = (BoundTask *)tmp;
    uint64_t *d_prefix = (uint64_t *)((char *)tmp + tb);
    uint64_t *d_out = d_prefix + n;
    if ((rc = fill_u32((uint32_t *)d_out, 4ull * n, 0, s))) return rc;
    COLO_HIP(hipMemcpyAsync(d_tasks, tasks, (size_t)n * sizeof(BoundTask), hipMemcpyHostToDevice, s));
    COLO_HIP(hipMemcpyAsync(d_prefix, prefix.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
    if (chunks) {
        hipLaunchKernelGGL(k_subset_bound, dim3((unsigned)chunks), dim3(kBlock), 0, s, d_tasks, d_prefix, n,
                           (unsigned long long *)d_out);
        COLO_HIP(hipGetLastError());
    }
    COLO_HIP(hipMemcpyAsync(out, d_out, (size_t)n * 16, hipMemcpyDeviceToHost, s));
    COLO_HIP(hipStreamSynchronize(s));
    return COLO_OK;
}

int colo::gather_ranges(colo_ctx *c, const uint32_t *d_src, uint32_t *d_dst, const CopyRange *ranges, uint64_t n,
                        hipStream_t s) {
    colo_ctx *ctx_for_err = c;
    if (!n) return COLO_OK;
    std::vector<uint64_t> prefix;
    const uint64_t chunks = chunk_prefix_of(ranges, n, [](const CopyRange &r) { return r.words; }, prefix);
    if (chunks >= (1ull << 31)) return set_error(c, COLO_EUNSUPPORTED, "gather: too many chunks");
    const size_t rb = ((size_t)n * sizeof(CopyRange) + 15) & ~(size_t)15;
    void *tmp;
    int rc = scratch(c, rb + (size_t)n * 8, &tmp);
    if (rc) return rc;
    CopyRange *d_ranges = (CopyRange *)tmp;
    uint64_t *d_prefix = (uint64_t *)((char *)tmp + rb);
    COLO_HIP(hipMemcpyAsync(d_ranges, ranges, (size_t)n * sizeof(CopyRange), hipMemcpyHostToDevice, s));
    COLO_HIP(hipMemcpyAsync(d_prefix, prefix.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
    if (chunks) {
        hipLaunchKernelGGL(k_gather_ranges, dim3((unsigned)chunks), dim3(kBlock), 0, s, d_src, d_dst, d_ranges,
                           d_prefix, n);
        COLO_HIP(hipGetLastError());
    }
    COLO_HIP(hipStreamSynchronize(s));
    return COLO_OK;
}

int colo::rank_to_sorted(colo_ctx *c, const uint32_t *d_src, uint32_t f, uint32_t *d_dst, hipStream_t s) {
    colo_ctx *ctx_for_err = c;
    if (f >= c->F) return set_error(c, COLO_EARG, "rank_to_sorted: feature out of range");
    const uint32_t N = c->fcount[f];
    if (!N) return COLO_OK;
    hipLaunchKernelGGL(k_rank_to_sorted, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0, s, d_src, N,
                       c->d_spos + c->fstart[f], c->fstart[f], d_dst);
    COLO_HIP(hipGetLastError());
    return COLO_OK;
}

// The miner's PI segments in chunks of kBoundWords words: popcounts summed
// with one atomic per workgroup; the chunk holding a boundary bit reports it.
__global__ __launch_bounds__(kBlock) void k_pi_chunks(const uint32_t *bits, const SegDesc *segs,
                                                      const uint64_t *chunk_prefix, uint64_t n, SegOut *out) {
    const uint64_t b = blockIdx.x;
    const uint64_t t = chunk_owner(chunk_prefix, n, b);
    const SegDesc sd = segs[t];
    const uint64_t w0 = (b - chunk_prefix[t]) * kBoundWords;
    const uint64_t w1 = w0 + kBoundWords < sd.words ? w0 + kBoundWords : sd.words;
    const uint32_t *w = bits + sd.off;
    uint64_t sum = 0;
    for (uint64_t i = w0 + threadIdx.x; i < w1; i += kBlock) sum += __popc(w[i]);
    uint64_t ex;
    const uint64_t tot = block_excl_scan(sum, ex);
    if (threadIdx.x == 0) {
        if (tot) atomicAdd((unsigned long long *)&out[t].pop, (unsigned long long)tot);
        const uint64_t fw = sd.first_bit >> 5, lw = sd.last_bit >> 5;
        if (fw >= w0 && fw < w1) out[t].first = (w[fw] >> (sd.first_bit & 31)) & 1u;
        if (lw >= w0 && lw < w1) out[t].last = (w[lw] >> (sd.last_bit & 31)) & 1u;
    }
}

void colo::seg_layout(const colo_ctx *c, uint32_t k, const uint32_t *cand, uint64_t bits_off, uint64_t *seg_off,
                      uint64_t *words_total) {
    uint64_t o = bits_off;
    for (uint32_t j = 0; j < k; j++) {
        seg_off[j] = o;
        o += (c->fcount[cand[j]] + 31) / 32;
    }
    if (words_total) *words_total = o - bits_off;
}

void colo::seg_descs(const colo_ctx *c, uint32_t k, const uint32_t *cand, uint64_t bits_off, SegDesc *out,
                     bool all_sorted) {
    uint64_t so[kMaxK];
    seg_layout(c, k, cand, bits_off, so, nullptr);
    for (uint32_t j = 0; j < k; j++) {
        uint32_t f = cand[j], N = c->fcount[f];
        out[j].off = so[j];
        out[j].words = (N + 31) / 32;
        if (j + 1 < k && !all_sorted) {  // input-rank indexed
            out[j].first_bit = 0;
            out[j].last_bit = N - 1;
        } else {          // last feature (every feature with all_sorted): sorted-rank indexed
            out[j].first_bit = c->first_sorted[f];
            out[j].last_bit = c->last_sorted[f];
        }
    }
}

// Numerators and PI (cu:108-132).  Reference windows over its 0/1 words:
// j = 0: [0, N0-1); j >= 1: [to-1, to-1+N_j) with to advanced per feature.
// With the reference's layouts (degree 2: b's slot 0 shares a word with a's
// slot N_a-1, K:45-47; degree k: features back to back, K:298-310) these
// windows sum to
//   num_0 = pop(P_0) - P_0[N_0-1]
//   k == 2: num_1 = (P_1[0] | P_0[N_0-1]) + pop(P_1) - P_1[0]
//   k >= 3: num_j = P_{j-1}[N_{j-1}-1] + pop(P_j) - P_j[N_j-1]
// pr_j = fl32(num_j) / fl32(N_j), PI = min(1, min_j pr_j).
void colo::pi_from_segs(const colo_ctx *c, uint32_t k, const uint32_t *cand, const SegOut *so, uint64_t *num_compat,
                        uint64_t *num_exact, float *pi_compat, float *pi_exact) {
    float pc = 1.0f, pe = 1.0f;
    for (uint32_t j = 0; j < k; j++) {
        uint64_t ex = so[j].pop, cm;
        if (j == 0) cm = so[0].pop - so[0].last;
        else if (k == 2) cm = (uint64_t)(so[1].first | so[0].last) + so[1].pop - so[1].first;
        else cm = (uint64_t)so[j - 1].last + so[j].pop - so[j].last;
        float N = (float)c->fcount[cand[j]];
        volatile float prc = (float)cm / N;
        volatile float pre = (float)ex / N;
        if (prc < pc) pc = prc;
        if (pre < pe) pe = pre;
        if (num_compat) num_compat[j] = cm;
        if (num_exact) num_exact[j] = ex;
    }
    if (pi_compat) *pi_compat = pc;
    if (pi_exact) *pi_exact = pe;
}

// participation expansion (parity hook)
__global__ void k_participation(const uint32_t *bits, SegDesc sd, uint32_t N, uint32_t f, bool sorted_rank,
                                const uint32_t *spos, const uint32_t *fstart, uint8_t *flags) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    uint64_t b = sorted_rank ? (uint64_t)(spos[fstart[f] + i] - fstart[f]) : i;
    flags[i] = (bits[sd.off + (b >> 5)] >> (b & 31)) & 1u;
}

// ------------------------------------------------------------ C ABI

static int alloc_dev(colo_ctx *c, void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 4);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(c, COLO_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                             hipGetErrorName(e));
    }
    return COLO_OK;
}

static uint32_t bits_for(uint64_t v) {
    uint32_t b = 0;
    while (b < 64 && (v >> b)) b++;
    return b;
}

static void free_bin_bufs(colo_ctx *c) {
    if (c->bin_exec) (void)hipGraphExecDestroy(c->bin_exec);
    if (c->bin_graph) (void)hipGraphDestroy(c->bin_graph);
    c->bin_exec = nullptr;
    c->bin_graph = nullptr;
    void *ps[] = {c->b_keys, c->b_keys2, c->b_vals2, c->b_hist, c->b_runidx, c->b_stmp, c->b_small};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    c->b_keys = c->b_keys2 = c->b_vals2 = c->b_hist = c->b_runidx = nullptr;
    c->b_stmp = c->b_small = nullptr;
}

static uint64_t dir_tiles(const colo_ctx *c) { return (c->dir_len + kDirTile - 1) / kDirTile; }

// binning workspace (u64 words): radix and run scans, the directory's tile
// starts (dir_tiles + 1 u32), the results block
static uint64_t bin_tmp_words(const colo_ctx *c) {
    const uint64_t n = c->n;
    const uint64_t ntiles = (n + kRadixTile - 1) / kRadixTile;
    return std::max<uint64_t>(std::max(scan_tmp_words(ntiles * 256), scan_tmp_words(n)),
                              std::max<uint64_t>(dir_tiles(c) / 2 + 1, (uint64_t)c->F + 3));
}

// Work buffers of the binning for the current grid.
static int alloc_bin_bufs(colo_ctx *c, uint32_t keybytes) {
    int rc;
    const uint64_t n = c->n;
    const uint64_t ntiles = (n + kRadixTile - 1) / kRadixTile;
    if ((rc = alloc_dev(c, &c->b_keys, keybytes * n))) return rc;
    if ((rc = alloc_dev(c, &c->b_keys2, keybytes * n))) return rc;
    if ((rc = alloc_dev(c, &c->b_vals2, 4 * n))) return rc;
    if ((rc = alloc_dev(c, &c->b_hist, 4 * 256 * (ntiles ? ntiles : 1)))) return rc;
    if (keybytes < 8 && (rc = alloc_dev(c, &c->b_runidx, 8 * n + 8))) return rc;
    if ((rc = alloc_dev(c, (void **)&c->b_stmp, 8 * bin_tmp_words(c)))) return rc;
    if ((rc = alloc_dev(c, (void **)&c->b_small, 8 * (3 + 3ull * c->F)))) return rc;
    if (!c->h_bin && hipHostMalloc((void **)&c->h_bin, 8 * (3 + 3ull * c->F)) != hipSuccess) {
        c->h_bin = nullptr;
        return set_error(c, COLO_ENOMEM, "pinned binning results");
    }
    c->b_keybytes = keybytes;
    return COLO_OK;
}

// The binning as launches and copies only (no allocation, no host wait),
// so that it can be captured: cell keys, stable radix sort by (feature,
// cell), sorted gathers, the dense (feature, cell) directory, the runs of
// distinct (feature, cell) pairs, cell and quadrant masks, and the results
// block copied to pinned host memory.
template <typename K>
static int bin_launch(colo_ctx *c, hipStream_t s) {
    colo_ctx *ctx_for_err = c;
    const uint32_t n = c->n;
    K *keys = (K *)c->b_keys, *keys2 = (K *)c->b_keys2;
    uint32_t *vals2 = (uint32_t *)c->b_vals2, *hist = (uint32_t *)c->b_hist;
    uint64_t *stmp = c->b_stmp, *small = c->b_small;
    int rc;
    DevGrid g = c->dev().g;
    if ((rc = fill_u32(c->d_flag, 4, 0, s))) return rc;
    if ((rc = fill_u32((uint32_t *)small, 2 * (3 + 3ull * c->F), 0, s))) return rc;
    K *sk = keys;
    if (n) {
        hipLaunchKernelGGL((k_cell_keys<K>), dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, c->d_x, c->d_y, n,
                           c->d_fstart, c->F, g, c->cellbits, c->d_cell_of, keys, c->d_sidx, c->d_flag);
        uint32_t *sv = c->d_sidx;
        int bits = (int)(c->cellbits + c->fidbits);
        if ((rc = radix_sort<K>(c, keys, c->d_sidx, keys2, vals2, n, bits, hist, stmp, &sk, &sv, s))) return rc;
        if (sv != c->d_sidx) COLO_HIP(hipMemcpyAsync(c->d_sidx, sv, 4ull * n, hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL((k_gather_sorted<K>), dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, sk,
                           c->d_sidx, n, c->cellbits, c->d_x, c->d_y, c->d_sxy, c->d_spos, c->d_scell);
    }
    // directory, run starts and cell masks (W 64-bit words per cell, when
    // the filter's masks fit): each tile's first element (in stmp), then
    // the tiles
    const uint32_t W = mask_words(c->F);
    const bool masks = W <= kMaxMaskWords;
    if (masks && (rc = fill_u32((uint32_t *)c->d_cellmask, 2ull * c->cells * W, 0, s))) return rc;
    uint32_t *run_flag = vals2;  // free once the sorted indices are in d_sidx
    if (n) {
        const uint64_t ntiles = dir_tiles(c);
        uint32_t *tile_lo = (uint32_t *)stmp;
        hipLaunchKernelGGL((k_dir_tiles<K>), dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, sk, n,
                           c->cellbits, c->cells, ntiles, run_flag, tile_lo);
        hipLaunchKernelGGL((k_dir_fill<K>), dim3((unsigned)std::min<uint64_t>(ntiles, 1u << 20)), dim3(kBlock), 0, s,
                           sk, c->cellbits, c->cells, c->dir_len, ntiles, (const uint32_t *)tile_lo, c->d_dir,
                           masks ? (unsigned long long *)c->d_cellmask : nullptr, W);
    } else if ((rc = fill_u32(c->d_dir, c->dir_len, 0, s))) {
        return rc;
    }
    // runs (distinct (feature, cell) pairs) for the filter; H lands in small[1]
    uint64_t *run_idx = sizeof(K) < 8 ? (uint64_t *)c->b_runidx : (uint64_t *)keys2;
    if (n) {
        scan_rec<uint32_t, uint64_t>(run_flag, n, run_idx, small + 1, stmp, s);
        hipLaunchKernelGGL(k_runs, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, run_flag, run_idx,
                           c->d_scell, n, c->d_run_cell);
    }
    if (c->F && n)
        hipLaunchKernelGGL(k_bin_tail, dim3((c->F + kBlock - 1) / kBlock), dim3(kBlock), 0, s, c->d_flag, run_idx,
                           c->d_fstart, c->d_fcount, c->F, c->d_spos, small);
    // quadrant and 3x3 masks from the cell masks
    if (masks) {
        uint64_t nq = (c->gx >= 2 && c->gy >= 2) ? (uint64_t)(c->gx - 1) * (c->gy - 1) : 0;
        if (nq) hipLaunchKernelGGL(k_qmask, dim3(grid_for(nq * W, 1u << 30)), dim3(kBlock), 0, s, c->d_cellmask, c->gx,
                                   c->gy, W, c->d_qmask);
        hipLaunchKernelGGL(k_mask3, dim3(grid_for((uint64_t)c->cells * W, 1u << 30)), dim3(kBlock), 0, s,
                           c->d_cellmask, c->gx, c->gy, W, c->d_mask3);
    }
    COLO_HIP(hipMemcpyAsync(c->h_bin, small, 8 * (3 + 3ull * c->F), hipMemcpyDeviceToHost, s));
    COLO_HIP(hipGetLastError());
    return COLO_OK;
}

// Runs the binning for the current grid -- replaying its captured graph
// when graphs are on -- and reads the results block.
template <typename K>
static int bin_impl(colo_ctx *c, hipStream_t s) {
    colo_ctx *ctx_for_err = c;
    int rc;
    if (!c->b_keys || c->b_keybytes != sizeof(K)) {
        free_bin_bufs(c);
        if ((rc = alloc_bin_bufs(c, sizeof(K)))) return rc;
    }
    bool done = false;
    if (c->use_graphs && s) {
        if (!c->bin_exec) {
            if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                int lrc = bin_launch<K>(c, s);
                hipGraph_t graph = nullptr;
                hipError_t e = hipStreamEndCapture(s, &graph);
                if (lrc == COLO_OK && e == hipSuccess && graph &&
                    hipGraphInstantiate(&c->bin_exec, graph, nullptr, nullptr, 0) == hipSuccess) {
                    c->bin_graph = graph;
                    c->graph_captures++;
                } else {
                    if (graph) (void)hipGraphDestroy(graph);
                    c->bin_exec = nullptr;
                    c->use_graphs = false;  // capture unavailable here: eager from now on
                }
            } else {
                c->use_graphs = false;
            }
            (void)hipGetLastError();
        }
        if (c->bin_exec) {
            COLO_HIP(hipGraphLaunch(c->bin_exec, s));
            c->graph_replays++;
            done = true;
        }
    }
    if (!done && (rc = bin_launch<K>(c, s))) return rc;
    COLO_HIP(hipStreamSynchronize(s));
    const uint64_t *h = c->h_bin;
    if (h[0]) return set_error(c, COLO_EBOUNDS, "a point lies outside the grid (cell id >= gx*gy)");
    c->H = c->n ? h[1] : 0;
    c->run_off.assign(c->F + 1, c->H);
    c->first_sorted.assign(c->F, 0);
    c->last_sorted.assign(c->F, 0);
    if (c->n)
        for (uint32_t f = 0; f < c->F; f++) {
            c->run_off[f] = h[2 + f];
            c->first_sorted[f] = (uint32_t)h[3 + c->F + f];
            c->last_sorted[f] = (uint32_t)h[3 + 2 * c->F + f];
        }
    return COLO_OK;
}

// setGrid (cpp:102-115) for distance r, binary32 throughout; (re)allocates
// the r-dependent buffers.
static int set_grid(colo_ctx *c, float r) {
    free_bin_bufs(c);
    void *old[] = {c->d_dir, c->d_cellmask, c->d_qmask, c->d_mask3, c->d_run_cell};
    for (void *p : old)
        if (p) (void)hipFree(p);
    c->d_dir = nullptr; c->d_cellmask = nullptr; c->d_qmask = nullptr; c->d_mask3 = nullptr; c->d_run_cell = nullptr;
    c->r = r;
    volatile float r2 = r * r;
    c->r2 = r2;
    volatile float maxBy = c->maxY + r;
    volatile float gsx = (c->maxX - c->minX) / r;
    volatile float gsy = (maxBy - c->minY) / r;
    float cx = ceilf(gsx), cy = ceilf(gsy);
    if (!(r > 0.f) || !(cx >= 1.f) || !(cy >= 1.f) || cx > 4.0e9f || cy > 4.0e9f)
        return set_error(c, COLO_EARG, "degenerate grid");
    c->gx = (uint32_t)cx;
    c->gy = (uint32_t)cy;
    uint64_t cells = (uint64_t)c->gx * c->gy;
    if (cells >= (1ull << 31)) return set_error(c, COLO_EUNSUPPORTED, "more than 2^31 cells");
    c->cells = (uint32_t)cells;
    c->cellbits = std::max<uint32_t>(1, bits_for(c->cells));
    c->fidbits = bits_for(c->F ? c->F - 1 : 0);
    c->dir_len = (uint64_t)c->F * (c->cells + 1ull);
    int rc = alloc_dev(c, (void **)&c->d_dir, 4 * c->dir_len);
    if (!rc) rc = alloc_dev(c, (void **)&c->d_run_cell, 4 * ((uint64_t)c->n + 1));  // runs <= n
    if (!rc && mask_words(c->F) <= kMaxMaskWords) {
        const uint64_t W = mask_words(c->F);
        uint64_t nq = (c->gx >= 2 && c->gy >= 2) ? (uint64_t)(c->gx - 1) * (c->gy - 1) : 0;
        rc = alloc_dev(c, (void **)&c->d_cellmask, 8ull * c->cells * W);
        if (!rc) rc = alloc_dev(c, (void **)&c->d_qmask, 8 * (nq ? nq : 1) * W);
        if (!rc) rc = alloc_dev(c, (void **)&c->d_mask3, 8ull * c->cells * W);
    }
    return rc;
}

extern "C" {

const char *colo_last_error(const colo_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int colo_bin(colo_ctx *c, float r, void *stream) {
    if (!c) return COLO_EARG;
    hipStream_t s = pick_stream(c, stream);
    int rc = COLO_OK;
    if (r != c->r || !c->d_dir) {  // new grid: new buffers and a new capture
        c->pairwork.clear();
        rc = set_grid(c, r);
    }
    if (!rc) rc = (c->cellbits + c->fidbits <= 32) ? bin_impl<uint32_t>(c, s) : bin_impl<uint64_t>(c, s);
    return rc;
}

int colo_ctx_create(const colo_dataset *ds, float r, void *stream, colo_ctx **out) {
    if (!ds || !out) return COLO_EARG;
    *out = nullptr;
    colo_ctx *c = new colo_ctx();
    int rc = COLO_OK;
    c->n = ds->n;
    c->F = ds->F;
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return COLO_EHIP;
        }
        c->own_stream = true;
    }
    hipStream_t s = c->stream;
    c->fstart.resize(c->F);
    c->fcount.resize(c->F);
    uint64_t tot = 0;
    for (uint32_t f = 0; f < c->F; f++) {
        c->fstart[f] = (uint32_t)tot;
        c->fcount[f] = ds->fcount[f];
        if (!ds->fcount[f]) rc = set_error(c, COLO_EARG, "feature with no instances");
        tot += ds->fcount[f];
        c->labels.push_back(ds->labels ? std::string(ds->labels[f]) : std::to_string(f));
    }
    if (!rc && tot != c->n) rc = set_error(c, COLO_EARG, "sum of fcount != n");
    c->minX = ds->minX; c->minY = ds->minY; c->maxX = ds->maxX; c->maxY = ds->maxY;
    using Clock = std::chrono::steady_clock;
    auto t_up = Clock::now();
    if (!rc) {
        const uint64_t n = c->n;
        if (!rc) rc = alloc_dev(c, (void **)&c->d_x, 4 * n);
        if (!rc) rc = alloc_dev(c, (void **)&c->d_y, 4 * n);
        if (!rc) rc = alloc_dev(c, (void **)&c->d_sxy, 8 * n);
        if (!rc) rc = alloc_dev(c, (void **)&c->d_cell_of, 4 * n);
        if (!rc) rc = alloc_dev(c, (void **)&c->d_sidx, 4 * n);
        if (!rc) rc = alloc_dev(c, (void **)&c->d_spos, 4 * n);
        if (!rc) rc = alloc_dev(c, (void **)&c->d_scell, 4 * n);
        if (!rc) rc = alloc_dev(c, (void **)&c->d_fstart, 4ull * c->F);
        if (!rc) rc = alloc_dev(c, (void **)&c->d_fcount, 4ull * c->F);
        if (!rc) rc = alloc_dev(c, (void **)&c->d_flag, 16);
        if (!rc && hipHostMalloc((void **)&c->h_pinned, 4096, hipHostMallocDefault) != hipSuccess)
            rc = set_error(c, COLO_ENOMEM, "pinned staging");
        if (!rc) {
            hipError_t e = hipSuccess;
            if (n) {
                e = hipMemcpyAsync(c->d_x, ds->x, 4 * n, hipMemcpyHostToDevice, s);
                if (e == hipSuccess) e = hipMemcpyAsync(c->d_y, ds->y, 4 * n, hipMemcpyHostToDevice, s);
            }
            if (e == hipSuccess && c->F) e = hipMemcpyAsync(c->d_fstart, c->fstart.data(), 4ull * c->F, hipMemcpyHostToDevice, s);
            if (e == hipSuccess && c->F) e = hipMemcpyAsync(c->d_fcount, c->fcount.data(), 4ull * c->F, hipMemcpyHostToDevice, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) rc = set_error(c, COLO_EHIP, hipGetErrorString(e));
        }
        c->upload_ms = std::chrono::duration<double, std::milli>(Clock::now() - t_up).count();
        auto t_bin = Clock::now();
        if (!rc) rc = colo_bin(c, r, nullptr);
        if (!rc) rc = hipStreamSynchronize(s) == hipSuccess ? COLO_OK : set_error(c, COLO_EHIP, "binning");
        c->bin_ms = std::chrono::duration<double, std::milli>(Clock::now() - t_bin).count();
    }
    *out = c;  // returned even on error so colo_last_error can be read
    return rc;
}

int colo_ctx_destroy(colo_ctx *c) {
    if (!c) return COLO_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void *ptrs[] = {c->d_x, c->d_y, c->d_sxy, c->d_cell_of, c->d_sidx, c->d_spos, c->d_scell, c->d_dir,
                    c->d_fstart, c->d_fcount, c->d_cellmask, c->d_qmask, c->d_mask3, c->d_run_cell, c->d_flag,
                    c->d_scratch};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (c->h_pinned) (void)hipHostFree(c->h_pinned);
    free_bin_bufs(c);
    if (c->h_bin) (void)hipHostFree(c->h_bin);
    c->pool.release();
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return COLO_OK;
}

int colo_set_verbose(colo_ctx *c, int level) {
    if (!c) return COLO_EARG;
    c->verbose = level;
    return COLO_OK;
}

int colo_set_graphs(colo_ctx *c, int on) {
    if (!c) return COLO_EARG;
    c->use_graphs = on != 0;
    if (on) c->level_graphs_off = false;  // try level captures again
    return COLO_OK;
}

int colo_graph_info(const colo_ctx *c, uint32_t *captures, uint32_t *replays) {
    if (!c) return COLO_EARG;
    if (captures) *captures = c->graph_captures;
    if (replays) *replays = c->graph_replays;
    return COLO_OK;
}

int colo_get_tune(const colo_ctx *c, colo_tune *o) {
    if (!c || !o) return COLO_EARG;
    o->count_var = c->tune.count_var;
    for (int k = 0; k < 16; k++) {
        o->count_var_k[k] = c->tune.count_var_k[k];
        o->item_rows_k[k] = c->tune.item_rows_k[k];
    }
    o->write_var = c->tune.write_var;
    o->blocks_per_cu = c->tune.blocks_per_cu;
    o->item_rows = c->tune.item_rows;
    o->slots_budget_kb = c->tune.slots_budget_kb;
    for (int k = 0; k < 16; k++) o->write_rows_k[k] = c->tune.write_rows_k[k];
    return COLO_OK;
}

int colo_set_tune(colo_ctx *c, const colo_tune *t) {
    if (!c || !t) return COLO_EARG;
    for (int k = 0; k < 16; k++)
        if ((t->count_var_k[k] != COLO_TUNE_UNSET && t->count_var_k[k] > 13) ||
            (t->item_rows_k[k] != COLO_TUNE_UNSET && (t->item_rows_k[k] < 64 || t->item_rows_k[k] > 65536)) ||
            (t->write_rows_k[k] != COLO_TUNE_UNSET && (t->write_rows_k[k] < 64 || t->write_rows_k[k] > 65536)))
            return set_error(c, COLO_EARG, "launch shape out of range");
    if (t->count_var > 13 || t->write_var > 3 || t->blocks_per_cu < 1 || t->blocks_per_cu > 32 || t->item_rows < 64 ||
        t->item_rows > 65536)
        return set_error(c, COLO_EARG, "launch shape out of range");
    c->tune.count_var = t->count_var;
    for (int k = 0; k < 16; k++) {
        c->tune.count_var_k[k] = t->count_var_k[k];
        c->tune.item_rows_k[k] = t->item_rows_k[k];
    }
    c->tune.write_var = t->write_var;
    c->tune.blocks_per_cu = t->blocks_per_cu;
    c->tune.item_rows = t->item_rows;
    c->tune.slots_budget_kb = t->slots_budget_kb;
    for (int k = 0; k < 16; k++) c->tune.write_rows_k[k] = t->write_rows_k[k];
    return COLO_OK;
}

int colo_trim(colo_ctx *c) {
    if (!c) return COLO_EARG;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->pool.release();
    return COLO_OK;
}

int colo_get_grid(const colo_ctx *c, colo_grid *o) {
    if (!c || !o) return COLO_EARG;
    o->minX = c->minX; o->minY = c->minY; o->cellWidth = c->r; o->r2 = c->r2;
    o->gx = c->gx; o->gy = c->gy; o->cells = c->cells; o->n = c->n; o->F = c->F;
    return COLO_OK;
}

int colo_binning_export(colo_ctx *c, uint32_t *d_cell_of, uint32_t *d_sorted_idx, uint32_t *d_dir,
                        uint64_t *dir_len, void *stream) {
    if (!c) return COLO_EARG;
    colo_ctx *ctx_for_err = c;
    hipStream_t s = pick_stream(c, stream);
    if (d_cell_of && c->n) COLO_HIP(hipMemcpyAsync(d_cell_of, c->d_cell_of, 4ull * c->n, hipMemcpyDeviceToDevice, s));
    if (d_sorted_idx && c->n) COLO_HIP(hipMemcpyAsync(d_sorted_idx, c->d_sidx, 4ull * c->n, hipMemcpyDeviceToDevice, s));
    if (d_dir) COLO_HIP(hipMemcpyAsync(d_dir, c->d_dir, 4ull * c->dir_len, hipMemcpyDeviceToDevice, s));
    if (dir_len) *dir_len = c->dir_len;
    return COLO_OK;
}

int colo_filter(colo_ctx *c, const uint32_t *d_cands, uint32_t C, uint32_t k, float theta, uint32_t *d_counts,
                uint8_t *d_pass, void *stream) {
    if (!c || k < 2 || k > kMaxK) return COLO_EARG;
    return filter_batch(c, d_cands, C, k, theta, d_counts, d_pass, pick_stream(c, stream));
}

uint64_t colo_bits_words(const colo_ctx *c, uint32_t k, const uint32_t *cand) {
    uint64_t so[kMaxK], w = 0;
    if (!c || k < 2 || k > kMaxK) return 0;
    seg_layout(c, k, cand, 0, so, &w);
    return w;
}

static int check_cand(colo_ctx *c, uint32_t k, const uint32_t *cand) {
    if (k < 2 || k > kMaxK) return set_error(c, COLO_EARG, "pattern size out of range");
    for (uint32_t j = 0; j < k; j++)
        if (cand[j] >= c->F) return set_error(c, COLO_EARG, "feature id out of range");
    return COLO_OK;
}

static int sum_slots(colo_ctx *c, const uint32_t *d_slots, uint64_t n, uint64_t *rows, hipStream_t s) {
    colo_ctx *ctx_for_err = c;
    if (!rows) return COLO_OK;
    if (!n) { *rows = 0; return COLO_OK; }
    void *tmp;
    int rc = scratch(c, 8 * n + 8 * scan_tmp_words(n) + 16, &tmp);
    if (rc) return rc;
    uint64_t *offs = (uint64_t *)tmp;
    uint64_t *tot = offs + n;
    scan_rec<uint32_t, uint64_t>(d_slots, n, offs, tot, tot + 1, s);
    COLO_HIP(hipMemcpyAsync(c->h_pinned, tot, 8, hipMemcpyDeviceToHost, s));
    COLO_HIP(hipStreamSynchronize(s));
    *rows = c->h_pinned[0];
    return COLO_OK;
}

int colo_join2_count(colo_ctx *c, uint32_t a, uint32_t b, uint32_t *d_slots, uint32_t *d_bits, uint64_t *rows,
                     void *stream) {
    if (!c) return COLO_EARG;
    colo_ctx *ctx_for_err = c;
    uint32_t cand[2] = {a, b};
    int rc = check_cand(c, 2, cand);
    if (rc) return rc;
    hipStream_t s = pick_stream(c, stream);
    uint64_t so[2], words;
    seg_layout(c, 2, cand, 0, so, &words);
    if ((rc = fill_u32(d_bits, words, 0, s))) return rc;
    uint32_t Na = c->fcount[a];
    hipLaunchKernelGGL(k_join2_count, dim3((Na + kBlock - 1) / kBlock), dim3(kBlock), 0, s, c->dev(), a, b, d_slots,
                       d_bits + so[0], d_bits + so[1]);
    COLO_HIP(hipGetLastError());
    return sum_slots(c, d_slots, Na, rows, s);
}

int colo_scan_slots(colo_ctx *c, const uint32_t *d_slots, uint64_t n, uint64_t *d_offsets, uint64_t *total,
                    void *stream) {
    if (!c) return COLO_EARG;
    colo_ctx *ctx_for_err = c;
    hipStream_t s = pick_stream(c, stream);
    void *tmp;
    int rc = scratch(c, 8 * scan_tmp_words(n) + 16, &tmp);
    if (rc) return rc;
    uint64_t *tot = d_offsets + n;  // offsets has n + 1 entries: the sum lands at [n]
    if (!n) {
        if ((rc = fill_u32((uint32_t *)tot, 2, 0, s))) return rc;
    } else {
        scan_rec<uint32_t, uint64_t>(d_slots, n, d_offsets, tot, (uint64_t *)tmp, s);
    }
    COLO_HIP(hipGetLastError());
    if (total) {
        COLO_HIP(hipMemcpyAsync(c->h_pinned, tot, 8, hipMemcpyDeviceToHost, s));
        COLO_HIP(hipStreamSynchronize(s));
        *total = c->h_pinned[0];
    }
    return COLO_OK;
}

int colo_join2_write(colo_ctx *c, uint32_t a, uint32_t b, const uint32_t *d_slots, const uint64_t *d_offsets,
                     uint32_t *d_table, uint32_t *d_nbr, void *stream) {
    if (!c) return COLO_EARG;
    colo_ctx *ctx_for_err = c;
    uint32_t cand[2] = {a, b};
    int rc = check_cand(c, 2, cand);
    if (rc) return rc;
    hipStream_t s = pick_stream(c, stream);
    uint32_t Na = c->fcount[a];
    hipLaunchKernelGGL(k_join2_write, dim3((Na + kBlock - 1) / kBlock), dim3(kBlock), 0, s, c->dev(), a, b, d_slots,
                       d_offsets, d_table, d_nbr);
    COLO_HIP(hipGetLastError());
    return COLO_OK;
}

static CandK make_candk(const colo_ctx *c, uint32_t k, const uint32_t *cand) {
    CandK ck{};
    ck.km1 = k - 1;
    ck.fl = cand[k - 1];
    for (uint32_t j = 0; j < k; j++) ck.feat[j] = cand[j];
    seg_layout(c, k, cand, 0, ck.seg_off, nullptr);
    return ck;
}

int colo_joink_count(colo_ctx *c, const uint32_t *d_parent, uint64_t R, uint32_t k, const uint32_t *cand,
                     uint32_t *d_slots, uint32_t *d_bits, uint64_t *rows, void *stream) {
    if (!c) return COLO_EARG;
    colo_ctx *ctx_for_err = c;
    int rc = check_cand(c, k, cand);
    if (rc) return rc;
    if (k < 3) return set_error(c, COLO_EARG, "colo_joink_count needs k >= 3");
    hipStream_t s = pick_stream(c, stream);
    CandK ck = make_candk(c, k, cand);
    uint64_t words;
    uint64_t so[kMaxK];
    seg_layout(c, k, cand, 0, so, &words);
    if ((rc = fill_u32(d_bits, words, 0, s))) return rc;
    if (R) {
        hipLaunchKernelGGL(k_joink_count, dim3((unsigned)((R + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, c->dev(),
                           d_parent, R, ck, d_slots, d_bits);
        COLO_HIP(hipGetLastError());
    }
    return sum_slots(c, d_slots, R, rows, s);
}

int colo_joink_write(colo_ctx *c, const uint32_t *d_parent, uint64_t R, uint32_t k, const uint32_t *cand,
                     const uint32_t *d_slots, const uint64_t *d_offsets, uint32_t *d_table, void *stream) {
    if (!c) return COLO_EARG;
    colo_ctx *ctx_for_err = c;
    int rc = check_cand(c, k, cand);
    if (rc) return rc;
    if (k < 3) return set_error(c, COLO_EARG, "colo_joink_write needs k >= 3");
    hipStream_t s = pick_stream(c, stream);
    CandK ck = make_candk(c, k, cand);
    if (R) {
        hipLaunchKernelGGL(k_joink_write, dim3((unsigned)((R + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, c->dev(),
                           d_parent, R, ck, d_slots, d_offsets, d_table);
        COLO_HIP(hipGetLastError());
    }
    return COLO_OK;
}

int colo_joink_count_nbr(colo_ctx *c, const uint32_t *d_parent, uint64_t R, uint32_t k, const uint32_t *cand,
                         const uint32_t *d_nbr2, const uint64_t *d_offsets2, uint32_t *d_slots, uint32_t *d_bits,
                         uint64_t *rows, void *stream) {
    if (!c) return COLO_EARG;
    colo_ctx *ctx_for_err = c;
    int rc = check_cand(c, k, cand);
    if (rc) return rc;
    if (k < 3) return set_error(c, COLO_EARG, "colo_joink_count_nbr needs k >= 3");
    hipStream_t s = pick_stream(c, stream);
    CandK ck = make_candk(c, k, cand);
    uint64_t words;
    uint64_t so[kMaxK];
    seg_layout(c, k, cand, 0, so, &words);
    if ((rc = fill_u32(d_bits, words, 0, s))) return rc;
    if (R) {
        hipLaunchKernelGGL(k_joink_count_nbr, dim3((unsigned)((R + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                           c->dev(), d_parent, R, ck, NbrList{d_nbr2, d_offsets2, 0}, d_slots, d_bits);
        COLO_HIP(hipGetLastError());
    }
    return sum_slots(c, d_slots, R, rows, s);
}

int colo_joink_write_nbr(colo_ctx *c, const uint32_t *d_parent, uint64_t R, uint32_t k, const uint32_t *cand,
                         const uint32_t *d_nbr2, const uint64_t *d_offsets2, const uint32_t *d_slots,
                         const uint64_t *d_offsets, uint32_t *d_table, void *stream) {
    if (!c) return COLO_EARG;
    colo_ctx *ctx_for_err = c;
    int rc = check_cand(c, k, cand);
    if (rc) return rc;
    if (k < 3) return set_error(c, COLO_EARG, "colo_joink_write_nbr needs k >= 3");
    hipStream_t s = pick_stream(c, stream);
    CandK ck = make_candk(c, k, cand);
    if (R) {
        hipLaunchKernelGGL(k_joink_write_nbr, dim3((unsigned)((R + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                           c->dev(), d_parent, R, ck, NbrList{d_nbr2, d_offsets2, 0}, d_slots, d_offsets, d_table);
        COLO_HIP(hipGetLastError());
    }
    return COLO_OK;
}

int colo_pi(colo_ctx *c, const uint32_t *d_bits, uint32_t k, const uint32_t *cand, uint64_t *num_compat,
            uint64_t *num_exact, float *pi_compat, float *pi_exact, void *stream) {
    if (!c) return COLO_EARG;
    colo_ctx *ctx_for_err = c;
    int rc = check_cand(c, k, cand);
    if (rc) return rc;
    hipStream_t s = pick_stream(c, stream);
    SegDesc sd[kMaxK];
    seg_descs(c, k, cand, 0, sd);
    void *tmp;
    if ((rc = scratch(c, sizeof(sd) + sizeof(SegOut) * kMaxK, &tmp))) return rc;
    SegDesc *d_sd = (SegDesc *)tmp;
    SegOut *d_so = (SegOut *)((char *)tmp + sizeof(sd));
    COLO_HIP(hipMemcpyAsync(d_sd, sd, sizeof(SegDesc) * k, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_pi_abs, dim3(k), dim3(kBlock), 0, s, d_bits, d_sd, d_so);
    COLO_HIP(hipGetLastError());
    SegOut so[kMaxK];
    COLO_HIP(hipMemcpyAsync(so, d_so, sizeof(SegOut) * k, hipMemcpyDeviceToHost, s));
    COLO_HIP(hipStreamSynchronize(s));
    pi_from_segs(c, k, cand, so, num_compat, num_exact, pi_compat, pi_exact);
    return COLO_OK;
}

int colo_participation(colo_ctx *c, const uint32_t *d_bits, uint32_t k, const uint32_t *cand, uint8_t *d_flags,
                       void *stream) {
    if (!c) return COLO_EARG;
    colo_ctx *ctx_for_err = c;
    int rc = check_cand(c, k, cand);
    if (rc) return rc;
    hipStream_t s = pick_stream(c, stream);
    SegDesc sd[kMaxK];
    seg_descs(c, k, cand, 0, sd);
    uint64_t o = 0;
    for (uint32_t j = 0; j < k; j++) {
        uint32_t N = c->fcount[cand[j]];
        hipLaunchKernelGGL(k_participation, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0, s, d_bits, sd[j], N,
                           cand[j], j + 1 == k, c->d_spos, c->d_fstart, d_flags + o);
        o += N;
    }
    COLO_HIP(hipGetLastError());
    return COLO_OK;
}

}  // extern "C"

// batched PI over absolute segment offsets (the miner's arena)
namespace colo {
uint64_t pi_chunk_prefix(const SegDesc *segs, uint64_t n_segs, std::vector<uint64_t> &prefix) {
    return chunk_prefix_of(segs, n_segs, [](const SegDesc &d) { return d.words; }, prefix);
}
int pi_batch_launch(colo_ctx *c, const uint32_t *d_bits, const SegDesc *d_segs, const uint64_t *d_prefix,
                    uint64_t n_segs, uint64_t chunks, SegOut *d_out, hipStream_t s) {
    if (!n_segs) return COLO_OK;
    if (chunks >= (1ull << 31)) return set_error(c, COLO_EUNSUPPORTED, "pi: too many chunks");
    static_assert(sizeof(SegOut) % 4 == 0, "SegOut is whole words");
    int rc = fill_u32((uint32_t *)d_out, n_segs * (sizeof(SegOut) / 4), 0, s);
    if (rc) return rc;
    if (chunks)
        hipLaunchKernelGGL(k_pi_chunks, dim3((unsigned)chunks), dim3(kBlock), 0, s, d_bits, d_segs, d_prefix,
                           n_segs, d_out);
    return hipGetLastError() == hipSuccess ? COLO_OK : set_error(c, COLO_EHIP, "pi launch");
}
int scan_prepare(colo_ctx *c, const ScanSeg *segs, uint64_t nseg, ScanPlan &plan, hipStream_t s) {
    colo_ctx *ctx_for_err = c;
    plan = ScanPlan{};
    plan.nseg = nseg;
    if (!nseg) return COLO_OK;
    std::vector<ScanSegDev> sd(nseg);
    std::vector<ScanTile> tiles;
    for (uint64_t i = 0; i < nseg; i++) {
        sd[i] = ScanSegDev{segs[i].in, segs[i].out, segs[i].n, tiles.size(), 0, segs[i].chunked ? 1u : 0u, 0};
        for (uint64_t b = 0; b < segs[i].n; b += kScanTile) tiles.push_back(ScanTile{(uint32_t)i, 0, b});
        sd[i].tile1 = tiles.size();
        (segs[i].chunked ? plan.any_chunks : plan.any_rows) = true;
    }
    const uint64_t nt = tiles.size();
    plan.nt = nt;
    const size_t sd_b = nseg * sizeof(ScanSegDev), t_b = nt * sizeof(ScanTile);
    void *tmp;
    int rc = scratch(c, sd_b + t_b + 8 * nt + 64, &tmp);
    if (rc) return rc;
    char *q = (char *)tmp;
    plan.d_sd = q;
    q += sd_b;
    plan.d_tiles = q;
    q += t_b;
    plan.d_part = (uint64_t *)(((uintptr_t)q + 7) & ~(uintptr_t)7);
    // one host block holding both arrays, kept with the plan until the
    // stream has consumed the upload
    plan.host.resize(sd_b + t_b);
    memcpy(plan.host.data(), sd.data(), sd_b);
    if (t_b) memcpy(plan.host.data() + sd_b, tiles.data(), t_b);
    COLO_HIP(hipMemcpyAsync(plan.d_sd, plan.host.data(), sd_b + t_b, hipMemcpyHostToDevice, s));
    return COLO_OK;
}
int scan_launch(colo_ctx *c, const ScanPlan &plan, hipStream_t s) {
    if (!plan.nseg) return COLO_OK;
    const ScanSegDev *d_sd = (const ScanSegDev *)plan.d_sd;
    const ScanTile *d_tiles = (const ScanTile *)plan.d_tiles;
    const uint64_t nt = plan.nt;
    if (nt) hipLaunchKernelGGL(k_segscan_reduce, dim3((unsigned)nt), dim3(kBlock), 0, s, d_sd, d_tiles, plan.d_part);
    hipLaunchKernelGGL(k_segscan_carry, dim3((unsigned)plan.nseg), dim3(kBlock), 0, s, d_sd, plan.d_part);
    // per-row segments: tile scans; chunked ones: prefixes of the chunk sums
    if (nt && plan.any_rows)
        hipLaunchKernelGGL(k_segscan_apply, dim3((unsigned)nt), dim3(kBlock), 0, s, d_sd, d_tiles, plan.d_part);
    if (nt && plan.any_chunks)
        hipLaunchKernelGGL(k_segscan_apply_chunks, dim3((unsigned)((nt + kBlock / 32 - 1) / (kBlock / 32))),
                           dim3(kBlock), 0, s, d_sd, d_tiles, nt, plan.d_part);
    return hipGetLastError() == hipSuccess ? COLO_OK : set_error(c, COLO_EHIP, "scan launch");
}
int scan_segments(colo_ctx *c, const ScanSeg *segs, uint64_t nseg, hipStream_t s) {
    colo_ctx *ctx_for_err = c;
    ScanPlan plan;
    int rc = scan_prepare(c, segs, nseg, plan, s);
    if (!rc) rc = scan_launch(c, plan, s);
    if (rc) return rc;
    // the descriptors live in scratch: finish before the next scratch use
    COLO_HIP(hipStreamSynchronize(s));
    return COLO_OK;
}
}  // namespace colo

// Table checksum independent of row order: the 64-bit sum over rows of a
// hash of the row's instance ids in column order (positions are mapped to
// instance ids through map when given).
__device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlock) void k_table_hash(const uint32_t *t, uint64_t rows, uint32_t k,
                                                       const uint32_t *map, unsigned long long *out) {
    uint64_t s = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < rows; i += (uint64_t)gridDim.x * kBlock) {
        uint64_t h = 0x2545F4914F6CDD1Dull ^ k;
        for (uint32_t m = 0; m < k; m++) {
            const uint32_t v = map ? map[t[i * k + m]] : t[i * k + m];
            h = mix64(h ^ (((uint64_t)m << 32) | v));
        }
        s += h;
    }
    uint64_t ex;
    uint64_t tot = block_excl_scan(s, ex);
    if (threadIdx.x == 0) atomicAdd(out, (unsigned long long)tot);
}


namespace colo {
int table_hash(colo_ctx *c, const uint32_t *d_t, uint64_t rows, uint32_t k, const uint32_t *d_map, uint64_t *h,
               hipStream_t s) {
    colo_ctx *ctx_for_err = c;
    void *tmp;
    int rc = scratch(c, 16, &tmp);
    if (rc) return rc;
    if ((rc = fill_u32((uint32_t *)tmp, 2, 0, s))) return rc;
    if (rows) hipLaunchKernelGGL(k_table_hash, dim3(grid_for(rows, 2048)), dim3(kBlock), 0, s, d_t, rows, k, d_map,
                                 (unsigned long long *)tmp);
    COLO_HIP(hipGetLastError());
    COLO_HIP(hipMemcpyAsync(c->h_pinned, tmp, 8, hipMemcpyDeviceToHost, s));
    COLO_HIP(hipStreamSynchronize(s));
    *h = c->h_pinned[0];
    return COLO_OK;
}
}  // namespace colo
