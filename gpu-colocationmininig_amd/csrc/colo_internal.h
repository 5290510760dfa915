// colo_internal.h -- shared definitions of the device layer and the host
// miner.  Not part of the C ABI (include/colo.h is).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <chrono>
#include <iterator>
#include <map>
#include <string>
#include <vector>

#include "../../include/colo.h"

namespace colo {

constexpr int kMaxK = 16;          // largest pattern size the kernels accept
constexpr int kBlock = 256;        // threads per workgroup (4 wavefronts)
constexpr int kMaxRun = 64;        // candidates sharing a parent table in one work item
constexpr uint32_t kMaxMaskWords = 4;  // feature masks of the filter: 64-bit words, so at most 256 features
constexpr uint32_t mask_words(uint32_t F) { return F ? (F + 63) / 64 : 1; }

// Grid geometry as the kernels read it (gridBox, Codes/DataTypes.h:77-83).
struct DevGrid {
    float minX, minY, r, r2;
    uint32_t gx, gy, cells, F;
};

// Device-resident dataset after binning.  Feature-major sorted order: the
// instances of feature f occupy sorted positions [fstart[f], fstart[f]+N_f),
// ordered by (cell, input index).
struct DevData {
    DevGrid g;
    const float *x, *y;            // input order
    const uint32_t *cell_of;       // input order -> cell
    const uint32_t *sidx;          // sorted position -> input index
    const uint32_t *spos;          // input index -> sorted position
    const float2 *sxy;             // coordinates in sorted order, interleaved
    const uint32_t *scell;         // cell in sorted order
    const uint32_t *dir;           // F x (cells+1): first sorted position of (f, c)
    const uint32_t *fstart;        // F
    const uint32_t *fcount;        // F
    // presence of each feature in the clamped 3 x 3 block of cells around
    // each cell (mask_w words per cell, bit f of word f / 64), or null
    const uint64_t *mask3;
    uint32_t mask_w;
};

// slots_off of a candidate whose count pass keeps no per-row slots (its
// rows and participation bits are still counted; a table built for it later
// is recounted with slots).
constexpr uint64_t kNoSlots = ~0ull;

// One candidate of a batched (persistent) level launch.
struct CandDesc {
    uint32_t k;
    uint32_t f[kMaxK];
    uint64_t n_anchor;             // N_{p0} at level 2, parent rows at level k
    const uint32_t *parent;        // level k: parent table (n_anchor x (k-1))
    uint64_t slots_off;            // offset of this candidate's slots in the arena, or kNoSlots
    uint64_t bits_off;             // word offset of the bitmap in the bits arena
    uint64_t seg_off[kMaxK];       // word offset of segment j inside the bitmap
    const uint64_t *offsets;       // write pass: row offsets (n_anchor), global
    uint64_t off_base;             // write pass: global offset of row 0
    uint32_t *table;               // write pass: output table
    // level k: neighbour lists of the anchors in the last feature (sorted
    // positions, scan order), i.e. the second column of the level-2 instance
    // table of (f[0], f[k-1]), with its row offsets per anchor sorted rank
    // (nb_off[t] - nb_base .. nb_off[t+1] - nb_base)
    const uint32_t *nb_table;
    const uint64_t *nb_off;
    uint64_t nb_base;
    uint32_t *nbr_out;             // level-2 write pass: neighbour-list output
    const uint32_t *wslots;        // write pass: this table's slot counts (n_anchor)
};

// One work item of a batched launch: anchors / parent rows [begin, end) of
// candidates cand .. cand+ncand-1, which share their parent table (level k:
// one pass over a parent row serves every candidate that extends it).
struct WorkItem {
    uint32_t cand;
    uint32_t ncand;
    uint64_t begin, end;
};

// PI segment descriptor: popcount of words [off, off+words) plus the bits of
// the instances with input rank 0 and N-1.
struct SegDesc {
    uint64_t off, words;
    uint64_t first_bit, last_bit;
};
struct SegOut {
    uint64_t pop;
    uint32_t first, last;
};

// Caching device allocator of one context: freed blocks are kept by size and
// handed out again (a mining run allocates and frees multi-GB table arenas
// per level; hipMalloc/hipFree of such sizes cost milliseconds).  Blocks are
// rounded up to 2 MiB (or a power of two below that) and reused when at most
// 1.5x the request.  Up to 3/4 of the device's memory stays cached; when
// hipMalloc fails, cached blocks are returned largest first until it
// succeeds.  Everything is returned to HIP on release() (colo_trim) and when
// the context is destroyed.
struct DevPool {
    std::multimap<size_t, void *> free_blocks;
    size_t cached = 0;
    size_t in_use = 0, peak = 0;  // bytes handed out and not yet returned; their maximum
    double sys_ms = 0, sys_bytes = 0;  // time in hipMalloc/hipFree, bytes allocated
    struct Timer {
        double &acc;
        std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
        explicit Timer(double &a) : acc(a) {}
        ~Timer() { acc += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
    };
    size_t cap = 0, total = 0;  // bytes kept cached at most (3/4 of the device's memory); device memory
    size_t cap_bytes() {
        if (!cap) cap = total_bytes() / 4 * 3;
        return cap;
    }
    size_t total_bytes() {
        if (!total) {
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) != hipSuccess || !tot) {
                (void)hipGetLastError();
                tot = 64ull << 30;
            }
            total = tot;
        }
        return total;
    }
    static size_t round(size_t b) {
        if (b >= (2u << 20)) return (b + (2u << 20) - 1) & ~(size_t)((2u << 20) - 1);
        size_t r = 256;
        while (r < b) r <<= 1;
        return r;
    }
    void *alloc(size_t bytes, size_t *got) {
        size_t want = round(bytes ? bytes : 1);
        auto it = free_blocks.lower_bound(want);
        if (it != free_blocks.end() && it->first <= want + want / 2) {
            void *p = it->second;
            *got = it->first;
            cached -= it->first;
            free_blocks.erase(it);
            note_out(*got);
            return p;
        }
        void *p = nullptr;
        Timer tm(sys_ms);
        bool synced = false;
        while (hipMalloc(&p, want) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            if (free_blocks.empty()) return nullptr;
            // hand cached blocks back, largest first, until the request fits
            if (!synced) {
                (void)hipDeviceSynchronize();
                synced = true;
            }
            auto last = std::prev(free_blocks.end());
            (void)hipFree(last->second);
            cached -= last->first;
            free_blocks.erase(last);
        }
        sys_bytes += (double)want;
        *got = want;
        note_out(want);
        return p;
    }
    void note_out(size_t b) {
        in_use += b;
        if (in_use > peak) peak = in_use;
    }
    void free(void *p, size_t bytes) {
        if (!p) return;
        in_use -= bytes < in_use ? bytes : in_use;
        if (cached + bytes > cap_bytes()) {
            Timer tm(sys_ms);
            (void)hipFree(p);
            return;
        }
        free_blocks.emplace(bytes, p);
        cached += bytes;
    }
    void release() {
        Timer tm(sys_ms);
        (void)hipDeviceSynchronize();
        for (auto &kv : free_blocks) (void)hipFree(kv.second);
        free_blocks.clear();
        cached = 0;
    }
};

}  // namespace colo

// Launch shapes of the batched passes (chosen by the autotuner).
struct TuneParams {
    uint32_t count_var = 0;       // level-k count: 0 row per thread, 1 wave-cooperative,
                                  // 2 row per thread, candidates in turn, 3 two rows per thread
    uint32_t count_var_k[16] = {COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET,
                                COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET,
                                COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET,
                                COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET};
    uint32_t item_rows_k[16] = {COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET,
                                COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET,
                                COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET,
                                COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET};
    uint32_t count_var_of(uint32_t k) const {
        return (k < 16 && count_var_k[k] != COLO_TUNE_UNSET) ? count_var_k[k] : count_var;
    }
    uint32_t item_rows_of(uint32_t k) const {
        return (k < 16 && item_rows_k[k] != COLO_TUNE_UNSET) ? item_rows_k[k] : item_rows;
    }
    uint32_t write_rows_k[16] = {COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET,
                                 COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET,
                                 COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET,
                                 COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET, COLO_TUNE_UNSET};
    uint32_t write_rows_of(uint32_t k) const {
        return (k < 16 && write_rows_k[k] != COLO_TUNE_UNSET) ? write_rows_k[k] : item_rows_of(k);
    }
    uint32_t write_var = 0;       // level-k write: same choice
    uint32_t blocks_per_cu = 8;   // persistent grid: workgroups per CU
    uint32_t item_rows = 1024;    // rows (anchors) per work item without the cost model
    uint32_t slots_budget_kb = 0; // slots arena limit per level pass, KiB (0 = 1/16 of device memory)
};

struct colo_ctx {
    hipStream_t stream = nullptr;
    TuneParams tune;
    bool own_stream = false;
    uint32_t n = 0, F = 0;
    std::vector<uint32_t> fstart, fcount;
    std::vector<uint32_t> first_sorted, last_sorted;  // sorted rank of input rank 0 / N-1
    std::vector<std::string> labels;
    float minX = 0, minY = 0, maxX = 0, maxY = 0, r = 0, r2 = 0;
    uint32_t gx = 0, gy = 0, cells = 0;
    uint32_t cellbits = 0, fidbits = 0;
    // device buffers
    float *d_x = nullptr, *d_y = nullptr;
    float2 *d_sxy = nullptr;
    uint32_t *d_cell_of = nullptr, *d_sidx = nullptr, *d_spos = nullptr, *d_scell = nullptr;
    uint32_t *d_dir = nullptr;
    uint64_t dir_len = 0;
    uint32_t *d_fstart = nullptr, *d_fcount = nullptr;
    // filter support
    uint64_t *d_cellmask = nullptr, *d_qmask = nullptr, *d_mask3 = nullptr;
    uint32_t *d_run_cell = nullptr;
    uint64_t H = 0;
    std::vector<uint64_t> run_off;  // F+1
    uint32_t *d_flag = nullptr;     // device error flags
    // reusable scratch
    void *d_scratch = nullptr;
    size_t scratch_bytes = 0;
    uint64_t *h_pinned = nullptr;   // small pinned host staging
    // binning work buffers, kept while the grid is unchanged so that the
    // binning launches can be captured once per grid and replayed
    void *b_keys = nullptr, *b_keys2 = nullptr, *b_vals2 = nullptr, *b_hist = nullptr, *b_runidx = nullptr;
    uint64_t *b_stmp = nullptr, *b_small = nullptr;  // scan temporaries; results block (see bin_launch)
    uint64_t *h_bin = nullptr;      // pinned copy of the results block
    uint32_t b_keybytes = 0;
    hipGraph_t bin_graph = nullptr;
    hipGraphExec_t bin_exec = nullptr;
    bool use_graphs = true;
    bool level_graphs_off = false;  // set when a level's capture failed: eager from then on
    uint32_t graph_captures = 0, graph_replays = 0;
    std::vector<uint64_t> pairwork; // F x F 3x3 join work per (first, last), filled on first use
    colo::DevPool pool;             // miner buffers (tables, arenas, slots, offsets)
    double upload_ms = 0, bin_ms = 0; // colo_ctx_create: host-to-device copy; first binning
    int verbose = 0;                // progress lines on stderr (colo_set_verbose)
    std::string err;
    colo::DevData dev() const;
};

namespace colo {
// Device-layer helpers used by the miner (colo_device.hip).
int set_error(colo_ctx *c, int code, const std::string &msg);
hipStream_t pick_stream(colo_ctx *c, void *s);
int scratch(colo_ctx *c, size_t bytes, void **p);
int scan_u32_u64(colo_ctx *c, const uint32_t *in, uint64_t n, uint64_t *out, uint64_t *total_dev,
                 hipStream_t s);
int fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t s);
// Exclusive scan of n u32 slot counts into out[0..n) (u64), the total at
// out[n], with a caller-provided workspace of scan_ws_words(n) u64: launches
// only, no allocation.
uint64_t scan_ws_words(uint64_t n);
int scan_slots_ws(const uint32_t *in, uint64_t n, uint64_t *out, uint64_t *ws, hipStream_t s);
int filter_batch(colo_ctx *c, const uint32_t *d_cands, uint32_t C, uint32_t k, float theta,
                 uint32_t *d_counts, uint8_t *d_pass, hipStream_t s);
// The filter's pass flags (threshold_Kernel, K:545-573) from counts already
// known (a threshold sweep keeps the theta-independent counts); with
// d_pass NULL each flag replaces d_counts[c*k] (the reference's layout).
int threshold_batch(colo_ctx *c, const uint32_t *d_cands, uint32_t C, uint32_t k, float theta,
                    uint32_t *d_counts, uint8_t *d_pass, hipStream_t s);
// Batched passes over many candidates (one launch each).
// d_seg: 9 item boundaries, one segment per XCD (see xcd_segments);
// d_queue: 8 counters.
// d_rows: one u64 per descriptor, zeroed by the caller; receives the rows
// (sum of the slots) of every candidate.
int count_batch(colo_ctx *c, uint32_t k, const CandDesc *d_desc, const WorkItem *d_items, const uint64_t *d_seg,
                uint64_t n_items, uint32_t *d_slots, uint32_t *d_bits, uint64_t *d_rows, uint32_t *d_queue,
                hipStream_t s);
int write_batch(colo_ctx *c, uint32_t k, const CandDesc *d_desc, const WorkItem *d_items, const uint64_t *d_seg,
                uint64_t n_items, const uint32_t *d_slots, uint32_t *d_queue, hipStream_t s);
// PI segments of a level in one launch of fixed chunks of words (a hot
// feature's segment spreads over many workgroups).  pi_chunk_prefix (host):
// each segment's first chunk; returns the number of chunks.  pi_batch_launch
// clears d_out and launches (launches only: capturable in a graph); d_prefix
// holds the prefix.
uint64_t pi_chunk_prefix(const SegDesc *segs, uint64_t n_segs, std::vector<uint64_t> &prefix);
int pi_batch_launch(colo_ctx *c, const uint32_t *d_bits, const SegDesc *d_segs, const uint64_t *d_prefix,
                    uint64_t n_segs, uint64_t chunks, SegOut *d_out, hipStream_t s);
// Row-order independent checksum of a rows x k table; d_map (optional) maps
// stored sorted positions to instance ids.
int table_hash(colo_ctx *c, const uint32_t *d_t, uint64_t rows, uint32_t k, const uint32_t *d_map, uint64_t *h,
               hipStream_t s);
// Exclusive scans of many slot arrays in one pass: out has n + 1 entries,
// the total last.  With chunked, out holds one entry per 64 rows instead,
// out[c] = sum of in[0 .. 64c), ceil(n/64) + 1 entries, the total last.
// Waits for the stream before returning.
struct ScanSeg {
    const uint32_t *in;
    uint64_t *out;
    uint64_t n;
    bool chunked = false;
};
int scan_segments(colo_ctx *c, const ScanSeg *segs, uint64_t nseg, hipStream_t s);
// The same in two steps: scan_prepare builds the tiles and uploads the
// descriptors into the context's scratch (plan keeps the host copies);
// scan_launch launches the kernels (launches only: capturable in a graph).
// Nothing else may use the scratch until the launches have run.
struct ScanPlan {
    std::vector<unsigned char> host;  // descriptors and tiles as uploaded
    void *d_sd = nullptr, *d_tiles = nullptr;
    uint64_t *d_part = nullptr;
    uint64_t nseg = 0, nt = 0;
    bool any_rows = false, any_chunks = false;
};
int scan_prepare(colo_ctx *c, const ScanSeg *segs, uint64_t nseg, ScanPlan &plan, hipStream_t s);
int scan_launch(colo_ctx *c, const ScanPlan &plan, hipStream_t s);

// Segment descriptors of one candidate's bitmap (bits_off = base word).
void seg_layout(const colo_ctx *c, uint32_t k, const uint32_t *cand, uint64_t bits_off,
                uint64_t *seg_off, uint64_t *words_total);
// all_sorted: every segment indexed by sorted rank (batched miner); else
// segments 0..k-2 by input rank and the last by sorted rank.
void seg_descs(const colo_ctx *c, uint32_t k, const uint32_t *cand, uint64_t bits_off, SegDesc *out,
               bool all_sorted = false);
void pi_from_segs(const colo_ctx *c, uint32_t k, const uint32_t *cand, const SegOut *so,
                  uint64_t *num_compat, uint64_t *num_exact, float *pi_compat, float *pi_exact);
int num_cus();
// 3x3 join work of (first, last) pairs (P pairs, host in/out).
int pairwork_batch(colo_ctx *c, const uint32_t *pairs, uint32_t P, uint64_t *out, hipStream_t s);
// One size-k candidate of refine_tests_batch: its parent table (rows x km1
// instance ids), the neighbour-list offsets of its (first, last) features
// indexed by the first feature's sorted rank, and that feature's fstart.
struct RefineItem {
    const uint32_t *parent;
    const uint64_t *nb_off;
    uint64_t rows;
    uint32_t km1, fstart0;
};
// Refinement tests of n size-k candidates (host in/out): the sum over each
// one's parent rows of the neighbour-list length of the row's first
// instance, i.e. the last-feature instances its count pass tests.
int refine_tests_batch(colo_ctx *c, const RefineItem *items, uint32_t n, uint64_t *out, hipStream_t s);
int mlp_predict_host(colo_mlp *m, const float *X, uint32_t n, float *y);

// Subset bound of the miner: for one (candidate, member) the AND of the
// member's participation windows in the candidate's (k-1)-subsets (nw
// windows of words u32 each, bit = sorted rank inside the feature), its
// popcount and its bit last_bit (the sorted rank of input rank N-1).
struct BoundTask {
    const uint32_t *w[kMaxK - 1];
    uint32_t nw, words, last_bit, pad;
};
// n tasks (host in/out): out[2i] = popcount, out[2i+1] = last bit.  Waits
// for the stream before returning.
int subset_bound_batch(colo_ctx *c, const BoundTask *tasks, uint64_t n, uint64_t *out, hipStream_t s);
// Word ranges copied in one launch (ranges host-side): d_dst[r.dst + i] =
// d_src[r.src + i], i < r.words.  Waits for the stream before returning.
struct CopyRange {
    uint64_t src, dst, words;
};
int gather_ranges(colo_ctx *c, const uint32_t *d_src, uint32_t *d_dst, const CopyRange *ranges, uint64_t n,
                  hipStream_t s);
// Participation bits of feature f from input-rank to sorted-rank order:
// d_dst (ceil(N/32) words, zeroed by the caller) gets the bit of the sorted
// rank of every input rank set in d_src.
int rank_to_sorted(colo_ctx *c, const uint32_t *d_src, uint32_t f, uint32_t *d_dst, hipStream_t s);
}  // namespace colo

#define COLO_HIP(call)                                                                  \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return colo::set_error(ctx_for_err, COLO_EHIP,                              \
                                   std::string(#call) + ": " + hipGetErrorString(e_));  \
    } while (0)
