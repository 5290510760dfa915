/*
 * colo.h -- C ABI of the MI355X colocation-mining hot path.
 *
 * Drop-in boundary for the device layer of spatialdatasciencegroup/
 * GPU-ColocationMininig (Codes/DeviceProcess.h:4-36, implemented in
 * Codes/DeviceProcess.cu:56-374).  Each entry point below names the
 * reference function it replaces.  Conventions:
 *
 *  - Plain C types only: pointers + sizes, POD structs, int return codes
 *    (COLO_OK = 0).  Nothing calls exit().
 *  - The reference keeps the dataset in __constant__ globals set by
 *    initializeDeviceMemConst (Codes/Constant.h:5-22); here a context handle
 *    (colo_ctx) owns the device-resident dataset, grid and binning.  Calls on
 *    one context are single-threaded; separate contexts are independent.
 *  - Pointers named d_* are device pointers (hipMalloc'd by the caller);
 *    the others are host pointers.  "stream" is a hipStream_t (NULL = the
 *    context's own stream).  Entry points that return host values
 *    synchronise that stream; the others only enqueue work.
 *  - Arithmetic follows the reference types (Codes/DataTypes.h:5-11):
 *    coordinates are binary32, ids are uint32; offsets are 64-bit (the
 *    reference's int32 offsets overflow past 2^31 rows).
 *
 * Participation bitmaps ("bits"): for a candidate P = (p_0..p_{k-1}) the
 * buffer holds k segments, segment j = colo_bits_words() layout, one bit per
 * instance of p_j.  The segment of the last feature is indexed by the
 * instance's rank in the cell-sorted order of its feature, the others by the
 * instance's rank in input order; colo_participation() expands a bitmap into
 * per-instance flags in input order.
 */
#ifndef COLO_H
#define COLO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes (the reference's StatusType, Codes/DataTypes.h:85-94, widened). */
#define COLO_OK 0
#define COLO_EFILE 1       /* input file missing or unreadable            */
#define COLO_EPARSE 2      /* malformed config / parameter / CSV input    */
#define COLO_EBOUNDS 3     /* a point lies outside the grid               */
#define COLO_ENOMEM 4      /* device or host allocation failed            */
#define COLO_EHIP 5        /* HIP runtime error                           */
#define COLO_EARG 6        /* invalid argument                            */
#define COLO_EUNSUPPORTED 7/* outside the supported envelope (e.g. F>256) */

typedef struct colo_ctx colo_ctx;

/* Dataset as colocationFinder::populateData leaves it (cpp:139-254):
 * n instances grouped by feature, fcount[f] instances of feature f, the
 * bounds line of the CSV. */
typedef struct colo_dataset {
    uint32_t n;
    uint32_t F;
    const float *x;        /* host, n */
    const float *y;        /* host, n */
    const uint32_t *fcount;/* host, F */
    float minX, minY, maxX, maxY;
    const char *const *labels; /* host, F feature labels for the pattern files; NULL = "0".."F-1" */
} colo_dataset;

/* Grid geometry (gridBox, Codes/DataTypes.h:77-83, set by setGrid cpp:102-115). */
typedef struct colo_grid {
    float minX, minY, cellWidth, r2;
    uint32_t gx, gy, cells;
    uint32_t n, F;
} colo_grid;

/* Creates a context: copies the dataset to the device, sets the grid for
 * neighbour distance r and bins the instances (on-device stable radix sort
 * by (feature, cell), cell offsets by prefix sum).
 * Replaces: initializeDeviceMemConst (DeviceProcess.h:11, cu:166-184) and
 * CountCellFeatures (DeviceProcess.h:33-34, cu:312-369), called from
 * colocationFinder::setGrid (cpp:102-126). */
int colo_ctx_create(const colo_dataset *ds, float r, void *stream, colo_ctx **out);
int colo_ctx_destroy(colo_ctx *ctx);
/* Returns the device memory the context keeps cached between calls (table
 * arenas and scratch of earlier mining runs) to the HIP runtime.  No
 * counterpart in the reference, which allocates and frees per call. */
int colo_trim(colo_ctx *ctx);
/* Re-grids and re-bins the resident dataset for distance r (the same work
 * colo_ctx_create does after its upload).
 * Replaces: colocationFinder::setGrid (cpp:102-126) -> CountCellFeatures
 * (DeviceProcess.h:33-34, cu:312-369). */
int colo_bin(colo_ctx *ctx, float r, void *stream);
int colo_get_grid(const colo_ctx *ctx, colo_grid *out);
/* Human-readable message for the last error on this context. */
const char *colo_last_error(const colo_ctx *ctx);
/* level > 0: progress lines on stderr (one per autotuner measurement).  No
 * counterpart in the reference. */
int colo_set_verbose(colo_ctx *ctx, int level);

/* Copies the binning out (device -> device, input order for cell_of,
 * sorted order for sorted_idx; dir = F x (cells+1) first sorted position of
 * every (feature, cell)).  Any pointer may be NULL.  dir_len receives
 * F*(cells+1).  Parity hook for the binning kernels. */
int colo_binning_export(colo_ctx *ctx, uint32_t *d_cell_of, uint32_t *d_sorted_idx,
                        uint32_t *d_dir, uint64_t *dir_len, void *stream);

/* Cell-aggregate upper-bound filter for C candidates of size k (d_cands:
 * C*k feature ids, device).  d_counts (C*k u32) receives the per-feature
 * counts summed over the cells of qualifying 2x2 quadrants; d_pass (C
 * bytes) receives min_t fl32(cnt_t / N_t) >= theta.  With d_pass NULL the
 * pass flag of candidate c replaces d_counts[c*k], where the reference's
 * threshold_Kernel leaves it (DeviceProcess_Kernel.h:570).
 * Replaces: filter2 (DeviceProcess.h:18-19, cu:219-235) and filter2_2
 * (DeviceProcess.h:20-21, cu:237-249; applied with the same rule). */
int colo_filter(colo_ctx *ctx, const uint32_t *d_cands, uint32_t C, uint32_t k, float theta,
                uint32_t *d_counts, uint8_t *d_pass, void *stream);

/* Number of 32-bit words of the participation bitmap of a candidate. */
uint64_t colo_bits_words(const colo_ctx *ctx, uint32_t k, const uint32_t *cand);

/* Level-2 count pass for candidate (a, b): d_slots[t] = number of
 * b-instances within r of the t-th a-instance; d_bits (zeroed by the call)
 * receives the participation bits.  *rows receives sum of slots if non-NULL
 * (synchronises).
 * Replaces: degree2InstanceTableSlotCounter (DeviceProcess.h:5, cu:58-68). */
int colo_join2_count(colo_ctx *ctx, uint32_t a, uint32_t b, uint32_t *d_slots, uint32_t *d_bits,
                     uint64_t *rows, void *stream);

/* Exclusive scan of n slot counts into 64-bit row offsets; d_offsets has
 * n + 1 entries, the sum lands in d_offsets[n]; *total receives the sum if
 * non-NULL (synchronises).
 * Replaces: the host thrust::exclusive_scan (cpp:391, 887), getTotalInstances
 * (DeviceProcess.h:13, cu:154-159) and scalebyConstant (DeviceProcess.h:15,
 * cu:187-196; the write passes scale by the degree themselves). */
int colo_scan_slots(colo_ctx *ctx, const uint32_t *d_slots, uint64_t n, uint64_t *d_offsets,
                    uint64_t *total, void *stream);

/* colo_scan_slots without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * scan's workspace is a static device array of about 1 MiB per GPU, held
 * until the process ends, so n is limited to COLO_SCAN_NOALLOC_MAX_N
 * (COLO_EARG above it).  Calls on one GPU share that array: calls that
 * overlap on one GPU race.  Results equal colo_scan_slots's. */
#define COLO_SCAN_NOALLOC_MAX_N (1ull << 28)
int colo_scan_slots_noalloc(colo_ctx *ctx, const uint32_t *d_slots, uint64_t n, uint64_t *d_offsets,
                            uint64_t *total, void *stream);

/* Level-2 write pass: rows (g, q) in a-instance order, neighbours in scan
 * order, into d_table (2 * rows u32).  d_nbr (rows u32, may be NULL)
 * receives the same neighbours as positions in the cell-sorted order of
 * feature b: the anchors' neighbour lists of the level-k passes.
 * Replaces: degree2InstanceTableGenerator (DeviceProcess.h:6, cu:69-79). */
int colo_join2_write(colo_ctx *ctx, uint32_t a, uint32_t b, const uint32_t *d_slots,
                     const uint64_t *d_offsets, uint32_t *d_table, uint32_t *d_nbr, void *stream);

/* Level-k count pass (k >= 3) over the R rows of the parent table d_parent
 * (R*(k-1) u32, instance table of cand[0..k-2]); cand is host, k ids.
 * Replaces: generalInstanceTableSlotCounter (DeviceProcess.h:8, cu:82-92). */
int colo_joink_count(colo_ctx *ctx, const uint32_t *d_parent, uint64_t R, uint32_t k,
                     const uint32_t *cand, uint32_t *d_slots, uint32_t *d_bits, uint64_t *rows,
                     void *stream);

/* Level-k write pass: rows parent_row ++ [q] into d_table (k * rows u32).
 * Replaces: calcInstanceTableGeneral (DeviceProcess.h:9, cu:94-105). */
int colo_joink_write(colo_ctx *ctx, const uint32_t *d_parent, uint64_t R, uint32_t k,
                     const uint32_t *cand, const uint32_t *d_slots, const uint64_t *d_offsets,
                     uint32_t *d_table, void *stream);

/* Level-k passes that take the candidates of each parent row from the
 * anchor's neighbour list: d_nbr2 / d_offsets2 are the neighbour lists
 * written by colo_join2_write for (cand[0], cand[k-1]) and their n+1 row
 * offsets per cand[0] instance (colo_scan_slots of that pair's slots).  Same results as
 * colo_joink_count / colo_joink_write: the list holds exactly the
 * last-feature instances in the anchor's 3x3 cells within r of the anchor,
 * in scan order, so only the other members are tested.
 * Replaces: generalInstanceTableSlotCounter / calcInstanceTableGeneral
 * (DeviceProcess.h:8-9, cu:82-105). */
int colo_joink_count_nbr(colo_ctx *ctx, const uint32_t *d_parent, uint64_t R, uint32_t k,
                         const uint32_t *cand, const uint32_t *d_nbr2, const uint64_t *d_offsets2,
                         uint32_t *d_slots, uint32_t *d_bits, uint64_t *rows, void *stream);
int colo_joink_write_nbr(colo_ctx *ctx, const uint32_t *d_parent, uint64_t R, uint32_t k,
                         const uint32_t *cand, const uint32_t *d_nbr2, const uint64_t *d_offsets2,
                         const uint32_t *d_slots, const uint64_t *d_offsets, uint32_t *d_table,
                         void *stream);

/* Participation-index numerators and PI of one candidate from its bitmap
 * (synchronises).  num_compat: the reference's window sums (cu:113-124,
 * off by one); num_exact: participating instances per feature; pi_*:
 * min(1, min_j fl32(num_j) / fl32(N_j)).  Any output may be NULL.
 * Replaces: getParticipationIndex (DeviceProcess.h:12, cu:108-132). */
int colo_pi(colo_ctx *ctx, const uint32_t *d_bits, uint32_t k, const uint32_t *cand,
            uint64_t *num_compat, uint64_t *num_exact, float *pi_compat, float *pi_exact,
            void *stream);

/* Expands a bitmap into d_flags (sum_j N_{p_j} bytes, feature segments in
 * candidate order, instances in input order).  Parity hook. */
int colo_participation(colo_ctx *ctx, const uint32_t *d_bits, uint32_t k, const uint32_t *cand,
                       uint8_t *d_flags, void *stream);

/* ------------------------------------------------------------- miner */

/* Schedules of the per-level pipeline. */
#define COLO_SCHED_PLAIN 0      /* one count launch + one PI sync per candidate */
#define COLO_SCHED_PERSISTENT 1 /* one persistent launch per level phase       */
#define COLO_SCHED_LEARNED 2    /* persistent, queue ordered/pre-sized by MLP  */

typedef struct colo_mlp colo_mlp;

typedef struct colo_mine_opts {
    float theta;            /* prevalence threshold (parameter.txt line 2)   */
    uint32_t max_k;         /* largest pattern size, 0 = no cap               */
    int32_t pi_mode;        /* 0 = reference numerators, 1 = exact            */
    int32_t filter_on;      /* 1 = cell-aggregate filter (reference default)  */
    int32_t sched;          /* COLO_SCHED_*                                   */
    int32_t hash_tables;    /* 1 = report a position-sensitive table checksum */
    const char *out_prefix; /* ColocationDegree_<k>.txt prefix, NULL = none   */
    colo_mlp *model;        /* cost model of COLO_SCHED_LEARNED              */
    const char *log_path;   /* append (features, rows, refinement tests)
                               records of every counted candidate (training
                               data of the cost model), or NULL             */
    int32_t subset_bound;   /* 1 = from size 3 on, skip the candidates whose
                               subset bound is below theta: the bound caps
                               both PIs by the AND of each member's
                               participation in the candidate's (k-1)-subsets
                               (no counterpart in the reference; the pattern
                               files are unchanged, the report shows the
                               bound after "q" and omits the skipped
                               candidates' counts)                         */
} colo_mine_opts;

/* Per-run measurements (milliseconds unless stated).  The *_kernel_ms
 * fields are HIP-event times of the batched launches on the context's
 * stream (persistent and learned schedules); *_bytes are the algorithmic
 * bytes those launches must move (each input and output datum once). */
typedef struct colo_mine_stats {
    double filter_ms, count_ms, pi_ms, scan_ms, write_ms, host_ms, total_ms;
    double l2_join_ms;      /* level-2 count + scan + write                 */
    uint64_t l2_rows;       /* rows emitted by the level-2 tables            */
    uint64_t l2_pairs;      /* neighbour pairs counted at level 2            */
    uint64_t levels, patterns, candidates, launches;
    double count_kernel_ms, write_kernel_ms, filter_kernel_ms, pi_kernel_ms;
    double count_bytes, write_bytes;
    uint64_t count_launches, write_launches;
    double l2_count_kernel_ms, l2_count_bytes;
    double l2_write_kernel_ms, l2_write_bytes;
    /* learned schedule: predictions vs counted rows (|log1p| error summed
     * over candidates), table arenas that the prediction under-sized (the
     * exact allocation was taken instead), and arena bytes reserved */
    double model_abs_err, model_n, arena_bytes;
    uint64_t overflows, arenas;
    double model_ms;        /* features + prediction (learned schedule, logging) */
    double alloc_ms;        /* host time in hipMalloc / hipFree (pool misses) */
    double alloc_bytes;     /* bytes obtained from hipMalloc during the run   */
    double level_ms[16];    /* wall time of level k (index k)                 */
    double level_count_kernel_ms[16]; /* count-pass kernel time of level k    */
    double pool_peak_bytes; /* most device bytes the run held at once (tables,
                               arenas, slots, reservations)                */
    uint64_t slots_deferred;/* candidates counted without per-row slots
                               (slots arena above its budget)              */
    uint64_t recounts;      /* candidates recounted for their table's slots */
    uint64_t pruned;        /* candidates skipped by the subset bound        */
    double bound_ms;        /* subset bounds: wall time                      */
    double scan_kernel_ms;  /* table row-offset scans (batched schedules)    */
    uint64_t graphs;        /* level passes launched as captured HIP graphs  */
    uint64_t tables_skipped;/* tables of the reference's set not built: every
                               candidate extending them fails the next
                               level's filter or subset bound (only without
                               table checksums)                            */
} colo_mine_stats;

/* HIP graphs: the binning (colo_bin, colo_ctx_create) is captured once per
 * grid (distance r) as a graph and replayed by later colo_bin calls with the
 * same r.  Each pass of a level of the batched schedules (count, PI
 * popcounts, offsets scan, table write) is launched as a captured graph
 * (colo_mine_stats.graphs counts them).  on = 0 runs everything
 * eagerly.  Capture falls back to eager launches where the stream cannot be
 * captured.  colo_graph_info reports the binning's captures and replays so
 * far.  No counterpart in the reference. */
int colo_set_graphs(colo_ctx *ctx, int on);
int colo_graph_info(const colo_ctx *ctx, uint32_t *captures, uint32_t *replays);

/* ------------------------------------------------------- launch shapes */

/* Launch shapes of the batched (persistent) passes.  No counterpart in the
 * reference, whose kernels have fixed 1-D launches.
 *   count_var, write_var  level-k count / write pass: 0 = one parent row
 *                         per thread, all candidates of a work item per
 *                         row; 1 = wave-cooperative over 64-row chunks (a
 *                         row group's neighbour list spread over the
 *                         lanes); 2 = one parent row per thread, the work
 *                         item's candidates one after another; 3 (write)
 *                         = 2 with one row offset kept per 64 rows (the
 *                         wave rebuilds its rows' offsets from the slots);
 *                         3 (count) = two adjacent rows per thread, their list
 *                         read once when they share the anchor; 6 (count
 *                         only) = four adjacent rows per thread, one list
 *                         pass when all four share the anchor; 8 (count
 *                         only) = two adjacent rows per lane with the wave
 *                         converged, one list pass when they share the
 *                         anchor, member bits combined over the wave; 4
 *                         and 5 (count only) = 0 and 3 with registers
 *                         capped for 8 waves per SIMD, 7 and 9 = 6 and 8
 *                         capped for 5; 10-13 (count only) = 4, 9, 0 and 8
 *                         with each work item's participation bits staged
 *                         in LDS (k = 3..8); the autotuner tries 0 and 3-13.
 *                         The level-2 count takes the variant's parity:
 *                         even = one anchor per thread, odd = the anchors
 *                         whose cell's 3x3 block holds the second feature
 *                         packed onto the wavefront's lanes (the others
 *                         have no rows); the autotuner tries both
 *   blocks_per_cu         persistent workgroups per compute unit (1..32)
 *   item_rows             rows per work-queue item; with the cost model,
 *                         items are cut at equal predicted cost between
 *                         item_rows / 4 and 8 x item_rows (64..65536) */
typedef struct colo_tune {
    uint32_t count_var, write_var, blocks_per_cu, item_rows;
    /* count variant of level k (index k, 2..15); COLO_TUNE_UNSET = count_var */
    uint32_t count_var_k[16];
    /* item rows of level k (index k, 2..15); COLO_TUNE_UNSET = item_rows */
    uint32_t item_rows_k[16];
    /* largest slots arena of a level's count pass in KiB; 0 = 1/16 of the
     * device's memory.  Candidates above it are counted without per-row
     * slots and recounted when their table is built (results unchanged). */
    uint32_t slots_budget_kb;
    /* parent rows per work item of level k's batched table write (index
     * 2..15); COLO_TUNE_UNSET = the level's item rows */
    uint32_t write_rows_k[16];
} colo_tune;
#define COLO_TUNE_UNSET 0xffffffffu

int colo_get_tune(const colo_ctx *ctx, colo_tune *out);
int colo_set_tune(colo_ctx *ctx, const colo_tune *t);
/* Autotuner: mines the context's data with opts and keeps the launch
 * shapes with the least HIP-event kernel time.  Per level, the count
 * variant and item size are chosen inside one mining run: each pair is
 * timed on a sample of that level's work items (one block of 16 items in
 * 8), and the write item size on that level's whole table write.  The
 * write variant and the workgroups per CU are then chosen one at a time by
 * full runs (coordinate descent).  The
 * context then uses the result.  With cache_path, results are stored under
 * a key of device and workload (instances, features, r, theta, max_k,
 * schedule, subset bound) and a cached
 * entry is taken without measuring.  Every shape gives the same results;
 * only time differs. */
int colo_autotune(colo_ctx *ctx, const colo_mine_opts *opts, const char *cache_path, colo_tune *out);

/* ------------------------------------------------------- cost model */

/* Candidate features the miner logs and the cost model reads (16 floats per
 * candidate, see COLO_FEATURE_NAMES in colo.py) and the model's shape:
 * 16 -> 64 -> 64 -> 64 -> 2, ReLU, FP32.  Output 0 predicts log1p(table
 * rows), output 1 log1p(refinement tests): the instances of the last
 * feature the count pass tests against the candidate's parent rows (at
 * size 2 the instances of the second feature in the 3x3 cells of each
 * anchor, at size k > 2 the neighbours of each parent row's first instance
 * in the last feature).  A log record is COLO_MLP_IN features, then
 * log1p(rows), log1p(tests) and k.
 * No counterpart in the reference: these serve the learned schedule. */
#define COLO_MLP_IN 16
#define COLO_MLP_HIDDEN 64
#define COLO_MLP_OUT 2
#define COLO_MLP_NPARAMS (64 * 16 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 2 * 64 + 2)
#define COLO_MLP_RECORD (COLO_MLP_IN + COLO_MLP_OUT + 1)

int colo_mlp_create(uint32_t seed, colo_mlp **out);
int colo_mlp_destroy(colo_mlp *m);
int colo_mlp_set_params(colo_mlp *m, const float *params, size_t n);
int colo_mlp_get_params(colo_mlp *m, float *params, size_t n);
int colo_mlp_set_norm(colo_mlp *m, const float *mean, const float *stdv);
/* y[i*COLO_MLP_OUT + o] = output o of model(X[i]) for n rows of
 * COLO_MLP_IN host floats. */
int colo_mlp_predict(colo_mlp *m, const float *X, uint32_t n, float *y);
/* Gradient of the loss (1/n) sum_i sum_o (y[i][o] - T[i][o])^2 over the n
 * rows (HIP forward and backward kernels); T has n * COLO_MLP_OUT floats,
 * grad COLO_MLP_NPARAMS. */
int colo_mlp_grad(colo_mlp *m, const float *X, const float *T, uint32_t n, float *grad, float *loss);
/* Adam training on the GPU: steps minibatches of `batch` rows (T as in
 * colo_mlp_grad). */
int colo_mlp_train(colo_mlp *m, const float *X, const float *T, uint32_t n, uint32_t steps, uint32_t batch,
                   float lr, uint32_t seed, float *final_loss);
int colo_mlp_save(colo_mlp *m, const char *path);
int colo_mlp_load(const char *path, colo_mlp **out);

/* Runs the Apriori level loop (colocationFinder::Begin cpp:1104-1137 with
 * degree2Processing cpp:256-425, generatePrevalentPatternsGeneral
 * cpp:736-854, generateInstanceTableGeneral cpp:856-909) on the binned
 * context; writes <out_prefix>ColocationDegree_<k>.txt (savetoFile2 /
 * savetoFileGen format, cpp:960-997) and returns a report string
 * (free with colo_free_str). */
int colo_mine(colo_ctx *ctx, const colo_mine_opts *opts, char **report, colo_mine_stats *stats);
void colo_free_str(char *s);

/* Threshold sweep in one call (the c5 workload's theta in {0.1, 0.2, 0.3,
 * 0.5}): runs colo_mine for each theta in ascending order, reusing what
 * does not depend on theta -- every counted candidate's rows, numerators
 * and PIs, and the level-2 tables -- across the runs.  Each report equals
 * the one colo_mine would give for that theta.  With n > 1 the pattern
 * files go to <out_prefix>theta_<theta>_ColocationDegree_<k>.txt.
 * No counterpart in the reference (which mines one theta per run). */
int colo_mine_sweep(colo_ctx *ctx, const colo_mine_opts *opts, const float *thetas, uint32_t n, char **reports,
                    colo_mine_stats *stats);

/* End-to-end drop-in for colocationFinder::Begin(argc, argv): reads
 * config.txt (cpp:1077-1093), parameter.txt (cpp:57-74) and the dataset CSV
 * (cpp:139-254), mines, and writes the pattern files under the config's
 * location line.  opts may be NULL (reference defaults); its theta is
 * ignored (parameter.txt holds it).  timing_line receives the stdout line
 * of cpp:1144 if non-NULL (free with colo_free_str). */
int colo_run_config(const char *config_path, const colo_mine_opts *opts, char **timing_line);

/* Next-level candidates from the prevalent list L (P patterns of size km1,
 * in list order) into out (cap patterns of size km1+1); returns the number
 * of candidates (call with cap = 0 to size out).  Host only.
 * Replaces: colocationFinder::candidateColocationGeneral (cpp:517-557) with
 * generateandCheckSubsets / subsetGen / checkSubset (cpp:598-701),
 * including its stop at the first candidate with a missing subset. */
uint64_t colo_candidates(const uint32_t *L, uint64_t P, uint32_t km1, uint32_t *out, uint64_t cap);

/* Parses a dataset CSV (cpp:139-254) into host arrays owned by the
 * returned handle. */
typedef struct colo_host_data colo_host_data;
int colo_load_csv(const char *path, colo_host_data **out);
/* The same parse on `threads` host threads (0 = one per MiB of input, up
 * to the hardware's), with an optional binary cache: when cache_path names
 * a cache written for this file (same size and modification time) it is
 * read instead of the CSV; otherwise the CSV is parsed and the cache
 * written.  Results equal colo_load_csv's.  No counterpart in the
 * reference, which parses line by line (cpp:139-254). */
int colo_load_csv_opts(const char *path, const char *cache_path, uint32_t threads, colo_host_data **out);
int colo_host_dataset(const colo_host_data *h, colo_dataset *ds);
const char *colo_host_label(const colo_host_data *h, uint32_t f);
void colo_host_free(colo_host_data *h);

#ifdef __cplusplus
}
#endif
#endif /* COLO_H */
