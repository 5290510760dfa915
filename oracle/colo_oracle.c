/*
 * colo_oracle.c -- CPU restatement of the reference colocation miner.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is linked
 * into tests, the smoke check and the CPU-baseline leg of the benchmark, and
 * never into the product library.  It restates, in plain C with OpenMP, the
 * algorithm of spatialdatasciencegroup/GPU-ColocationMininig as it is written
 * in Codes/colocationFinder.cpp (host driver), Codes/DeviceProcess.cu and
 * Codes/DeviceProcess_Kernel.h (device code), keeping the reference's own
 * dense layouts (cells x features count/start arrays, 0/1 bitmap words with
 * the overlapping degree-2 layout) so that each step can be read against the
 * cited lines.
 *
 * Arithmetic types follow Codes/DataTypes.h:5-11: Integer = uint32_t,
 * SInteger = int32_t, Real = float.  All coordinate arithmetic is IEEE
 * binary32; build with -ffp-contract=off and without -ffast-math.
 *
 * Pinning (each function is checked by tests against an independent fact):
 *   orc_load_csv      : hand-written CSV fixtures with known parse results.
 *   grid / binning    : closed-form FP32 cell ids; stable order checked
 *                       against numpy's stable argsort.
 *   nbr cells         : literal Utility_Inline.h:35-53; checked on every cell
 *                       of grids with gx = 1, 2, 3, 7 against the clipping
 *                       rule derived from it.
 *   join count/write  : brute-force O(N^2) all-pairs / all-cliques on small
 *                       inputs; hand-placed points at r-ulp, r, r+ulp.
 *   pi (compat/exact) : hand-computed bitmaps, including N_f = 1 and the
 *                       degree-2 overlapping word.
 *   filter            : brute-force quadrant enumeration on small grids and
 *                       the soundness invariant ub_t >= exact numerator.
 *   candgen           : hand-made prevalent lists that exercise the
 *                       first-failure break.
 *   miner             : the pattern set with the filter on equals the set
 *                       with the filter off; anti-monotonicity of exact PI.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <errno.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_EFILE 1
#define ORC_EPARSE 2
#define ORC_EBOUNDS 3
#define ORC_EMEM 4
#define ORC_EARG 5

typedef struct orc_ctx {
    uint32_t n, F;
    float *x, *y;               /* instanceLocationX/Y, cpp:185-186 */
    uint32_t *fid;              /* instanceList, cpp:184 */
    uint32_t *fstart, *fend, *fcount; /* featureInstanceStart/End/Count, cpp:218-229 */
    char **labels;              /* featureTypes, cpp:170,181 */
    float minX, minY, maxX, maxY; /* bounds line, cpp:191-197 */
    float r, r2;                /* param, cpp:66,73 */
    /* gridBox, cpp:102-115 */
    float gMinX, gMinY, gMaxX, gMaxY, cellWidth;
    uint32_t gx, gy, cells;
    uint32_t *cellId;           /* d_instanceCellIDs */
    uint32_t *sortedIdx;        /* d_cellBasedSortedIndex */
    int32_t *cnt, *start;       /* d_cellEventInstanceCount / Start, cells*F */
    int binned;
} orc_ctx;

/* ---------------------------------------------------------------- parsing */

/* stoi(): leading blanks, optional sign, decimal digits (libstdc++ uses strtol). */
static int parse_int(const char *s, long *out) {
    char *end;
    errno = 0;
    long v = strtol(s, &end, 10);
    if (end == s || errno) return -1;
    *out = v;
    return 0;
}
/* stof(): strtof, decimal -> binary32 correctly rounded. */
static int parse_float(const char *s, float *out) {
    char *end;
    errno = 0;
    float v = strtof(s, &end);
    if (end == s) return -1;
    *out = v;
    return 0;
}

/* Split one line on ',' the way boost::escaped_list_separator does for
 * unquoted numeric fields (cpp:156,166); surrounding double quotes are
 * dropped.  Returns the number of fields (at most maxf). */
static int split_fields(char *line, char **f, int maxf) {
    int k = 0;
    char *p = line;
    while (k < maxf) {
        char *q = p;
        if (*q == '"') {
            q++;
            char *e = strchr(q, '"');
            if (!e) return -1;
            *e = 0;
            f[k++] = q;
            p = e + 1;
            if (*p == ',') { p++; continue; }
            break;
        }
        f[k++] = p;
        char *c = strchr(p, ',');
        if (!c) break;
        *c = 0;
        p = c + 1;
    }
    return k;
}

static char *dupstr(const char *s) {
    size_t l = strlen(s);
    char *d = (char *)malloc(l + 1);
    memcpy(d, s, l + 1);
    return d;
}

void orc_free(orc_ctx *c);

static orc_ctx *ctx_alloc(uint32_t n) {
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    c->n = n;
    c->x = (float *)malloc(sizeof(float) * (n ? n : 1));
    c->y = (float *)malloc(sizeof(float) * (n ? n : 1));
    c->fid = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    return c;
}

/* Reads the reference dataset format (colocationFinder::populateData,
 * cpp:139-254): line 1 = N; rows "label,x,y" until the first empty line;
 * a new feature starts whenever stoi(label) differs from the previous row's;
 * then the bounds line "minX,minY,maxX,maxY".  Row count must equal N. */
orc_ctx *orc_load_csv(const char *path, int *err) {
    FILE *fp = fopen(path, "rb");
    *err = ORC_OK;
    if (!fp) { *err = ORC_EFILE; return NULL; }
    size_t cap = 0;
    char *line = NULL;
    ssize_t len;
    len = getline(&line, &cap, fp);
    long nl;
    if (len <= 0) { *err = ORC_EPARSE; free(line); fclose(fp); return NULL; }
    if (line[len - 1] == '\n') line[--len] = 0;
    if (parse_int(line, &nl) || nl < 0 || nl > 0xFFFFFFFFL) { *err = ORC_EPARSE; free(line); fclose(fp); return NULL; }
    orc_ctx *c = ctx_alloc((uint32_t)nl);
    uint32_t fcap = 16, F = 0;
    c->fstart = (uint32_t *)malloc(sizeof(uint32_t) * fcap);
    c->fcount = (uint32_t *)malloc(sizeof(uint32_t) * fcap);
    c->labels = (char **)malloc(sizeof(char *) * fcap);
    long last = 0;
    int have_last = 0;
    uint64_t row = 0;
    for (;;) {
        len = getline(&line, &cap, fp);
        if (len < 0) { len = 0; line[0] = 0; }
        if (len > 0 && line[len - 1] == '\n') line[--len] = 0;
        if (len == 0) break;                     /* cpp:164 */
        char *fld[3];
        int nf = split_fields(line, fld, 3);
        long lab;
        if (nf < 3 || parse_int(fld[0], &lab)) { *err = ORC_EPARSE; goto fail; }
        if (row >= c->n) { *err = ORC_EPARSE; goto fail; }
        if (!have_last || lab != last) {          /* cpp:168-183 */
            if (F == fcap) {
                fcap *= 2;
                c->fstart = (uint32_t *)realloc(c->fstart, sizeof(uint32_t) * fcap);
                c->fcount = (uint32_t *)realloc(c->fcount, sizeof(uint32_t) * fcap);
                c->labels = (char **)realloc(c->labels, sizeof(char *) * fcap);
            }
            c->fstart[F] = (uint32_t)row;
            c->fcount[F] = 0;
            c->labels[F] = dupstr(fld[0]);
            F++;
            last = lab;
            have_last = 1;
        }
        c->fid[row] = F - 1;
        if (parse_float(fld[1], &c->x[row]) || parse_float(fld[2], &c->y[row])) { *err = ORC_EPARSE; goto fail; }
        c->fcount[F - 1]++;
        row++;
    }
    if (row != c->n) { *err = ORC_EPARSE; goto fail; }
    len = getline(&line, &cap, fp);
    {
        if (len <= 0) { *err = ORC_EPARSE; goto fail; }
        if (line[len - 1] == '\n') line[--len] = 0;
        char *fld[4];
        int nf = split_fields(line, fld, 4);
        if (nf < 4 || parse_float(fld[0], &c->minX) || parse_float(fld[1], &c->minY) ||
            parse_float(fld[2], &c->maxX) || parse_float(fld[3], &c->maxY)) { *err = ORC_EPARSE; goto fail; }
    }
    c->F = F;
    c->fend = (uint32_t *)malloc(sizeof(uint32_t) * (F ? F : 1));
    for (uint32_t f = 0; f < F; f++) c->fend[f] = c->fstart[f] + c->fcount[f] - 1;
    free(line);
    fclose(fp);
    return c;
fail:
    c->F = F;
    free(line);
    fclose(fp);
    orc_free(c);
    return NULL;
}

/* Builds a context from arrays already in memory: instances grouped by
 * feature, fcount[f] instances of feature f in order. */
orc_ctx *orc_from_arrays(uint32_t n, uint32_t F, const float *x, const float *y,
                         const uint32_t *fcount, const float *bounds) {
    orc_ctx *c = ctx_alloc(n);
    memcpy(c->x, x, sizeof(float) * n);
    memcpy(c->y, y, sizeof(float) * n);
    c->F = F;
    c->fstart = (uint32_t *)malloc(sizeof(uint32_t) * (F ? F : 1));
    c->fend = (uint32_t *)malloc(sizeof(uint32_t) * (F ? F : 1));
    c->fcount = (uint32_t *)malloc(sizeof(uint32_t) * (F ? F : 1));
    c->labels = (char **)malloc(sizeof(char *) * (F ? F : 1));
    uint32_t s = 0;
    for (uint32_t f = 0; f < F; f++) {
        char buf[32];
        c->fstart[f] = s;
        c->fcount[f] = fcount[f];
        c->fend[f] = s + fcount[f] - 1;
        snprintf(buf, sizeof buf, "%u", f);
        c->labels[f] = dupstr(buf);
        for (uint32_t i = 0; i < fcount[f]; i++) c->fid[s + i] = f;
        s += fcount[f];
    }
    c->minX = bounds[0]; c->minY = bounds[1]; c->maxX = bounds[2]; c->maxY = bounds[3];
    return c;
}

void orc_free(orc_ctx *c) {
    if (!c) return;
    free(c->x); free(c->y); free(c->fid);
    free(c->fstart); free(c->fend); free(c->fcount);
    if (c->labels) for (uint32_t f = 0; f < c->F; f++) free(c->labels[f]);
    free(c->labels);
    free(c->cellId); free(c->sortedIdx); free(c->cnt); free(c->start);
    free(c);
}

uint32_t orc_n(const orc_ctx *c) { return c->n; }
uint32_t orc_F(const orc_ctx *c) { return c->F; }
void orc_features(const orc_ctx *c, uint32_t *fstart, uint32_t *fcount) {
    memcpy(fstart, c->fstart, sizeof(uint32_t) * c->F);
    memcpy(fcount, c->fcount, sizeof(uint32_t) * c->F);
}
void orc_points(const orc_ctx *c, float *x, float *y) {
    memcpy(x, c->x, sizeof(float) * c->n);
    memcpy(y, c->y, sizeof(float) * c->n);
}
void orc_bounds(const orc_ctx *c, float *b) { b[0] = c->minX; b[1] = c->minY; b[2] = c->maxX; b[3] = c->maxY; }
const char *orc_label(const orc_ctx *c, uint32_t f) { return f < c->F ? c->labels[f] : ""; }

/* parameter.txt (LoadParameter, cpp:57-74): r, theta, filter flag read with
 * operator>> (correctly rounded float); r2 = fl32(r*r). */
int orc_load_params(const char *path, float *r, float *theta, uint32_t *flag) {
    FILE *fp = fopen(path, "rb");
    if (!fp) return ORC_EFILE;
    char a[128], b[128], d[128];
    int k = fscanf(fp, "%127s %127s %127s", a, b, d);
    fclose(fp);
    if (k < 2) return ORC_EPARSE;
    if (parse_float(a, r) || parse_float(b, theta)) return ORC_EPARSE;
    long fl = 0;
    if (k == 3 && parse_int(d, &fl)) return ORC_EPARSE;
    *flag = (uint32_t)fl;
    return ORC_OK;
}

/* ------------------------------------------------------------ grid + bins */

/* setGrid (cpp:102-115): all FP32. */
static void make_grid(orc_ctx *c, float r) {
    c->r = r;
    volatile float rr = r * r;           /* cpp:73 */
    c->r2 = rr;
    c->gMinX = c->minX;
    c->gMinY = c->minY;
    c->gMaxX = c->maxX;
    c->gMaxY = c->maxY + r;              /* cpp:108 */
    c->cellWidth = r;
    volatile float gsx = (c->gMaxX - c->gMinX) / c->cellWidth;  /* cpp:111 */
    volatile float gsy = (c->gMaxY - c->gMinY) / c->cellWidth;  /* cpp:112 */
    c->gx = (uint32_t)ceilf(gsx);
    c->gy = (uint32_t)ceilf(gsy);
    c->cells = c->gx * c->gy;             /* cpp:115, u32 */
}

/* getCellID (K:615-625). */
static uint32_t cell_of(const orc_ctx *c, float x, float y) {
    volatile float rx = x - c->gMinX;
    volatile float ry = y - c->gMinY;
    rx = rx / c->cellWidth;
    ry = ry / c->cellWidth;
    uint32_t i = (uint32_t)rx;
    uint32_t j = (uint32_t)ry;
    return c->gx * j + i;
}

/* CountCellFeatures (cu:312-369): cell id per point, per-(cell,feature)
 * counts, stable sort of point ids by cell, exclusive scan of the counts in
 * cell-major order.  Points whose cell lies outside [0,cells) are rejected
 * (the reference would write outside its count array). */
int orc_bin(orc_ctx *c, float r) {
    make_grid(c, r);
    uint64_t cf = (uint64_t)c->cells * c->F;
    free(c->cellId); free(c->sortedIdx); free(c->cnt); free(c->start);
    c->cellId = (uint32_t *)malloc(sizeof(uint32_t) * (c->n ? c->n : 1));
    c->sortedIdx = (uint32_t *)malloc(sizeof(uint32_t) * (c->n ? c->n : 1));
    c->cnt = (int32_t *)calloc(cf ? cf : 1, sizeof(int32_t));
    c->start = (int32_t *)calloc(cf ? cf : 1, sizeof(int32_t));
    if (!c->cellId || !c->sortedIdx || !c->cnt || !c->start) return ORC_EMEM;
    for (uint32_t p = 0; p < c->n; p++) {
        uint32_t cid = cell_of(c, c->x[p], c->y[p]);
        if (cid >= c->cells) return ORC_EBOUNDS;
        c->cellId[p] = cid;
        c->cnt[(uint64_t)cid * c->F + c->fid[p]]++;
    }
    /* stable counting sort by cell id (thrust::sort_by_key, cu:342) */
    uint32_t *cc = (uint32_t *)calloc((uint64_t)c->cells + 1, sizeof(uint32_t));
    for (uint32_t p = 0; p < c->n; p++) cc[c->cellId[p] + 1]++;
    for (uint64_t i = 0; i < c->cells; i++) cc[i + 1] += cc[i];
    for (uint32_t p = 0; p < c->n; p++) c->sortedIdx[cc[c->cellId[p]]++] = p;
    free(cc);
    /* exclusive scan (cu:344) */
    int64_t run = 0;
    for (uint64_t i = 0; i < cf; i++) { c->start[i] = (int32_t)run; run += c->cnt[i]; }
    c->binned = 1;
    return ORC_OK;
}

void orc_grid(const orc_ctx *c, uint32_t *gx, uint32_t *gy, uint32_t *cells, float *r2) {
    *gx = c->gx; *gy = c->gy; *cells = c->cells; *r2 = c->r2;
}
void orc_binning(const orc_ctx *c, uint32_t *cellId, uint32_t *sortedIdx) {
    memcpy(cellId, c->cellId, sizeof(uint32_t) * c->n);
    memcpy(sortedIdx, c->sortedIdx, sizeof(uint32_t) * c->n);
}
void orc_cell_counts(const orc_ctx *c, int32_t *cnt, int32_t *start) {
    uint64_t cf = (uint64_t)c->cells * c->F;
    memcpy(cnt, c->cnt, sizeof(int32_t) * cf);
    memcpy(start, c->start, sizeof(int32_t) * cf);
}

/* getNeighborCellID (U:35-53) followed by the caller's range check
 * (K:28-29), in the reference's own unsigned/signed arithmetic.  Returns the
 * neighbour cell, or -1 when it is skipped. */
int32_t orc_nbr_cell(uint32_t gx, uint32_t cells, uint32_t cur, int32_t i, int32_t j) {
    uint32_t cy = cur / gx;
    uint32_t cx = cur - gx * cy;
    uint32_t x = cx + (uint32_t)i, y = cy + (uint32_t)j;
    int32_t ncid = (int32_t)(y * gx + x);
    if (cur % gx == 0) {
        if ((uint32_t)(ncid + 1) % gx == 0) ncid = -1;
    } else if ((cur + 1) % gx == 0) {
        if ((uint32_t)ncid % gx == 0) ncid = -1;
    }
    if (ncid < 0 || (uint32_t)ncid >= cells) return -1;
    return ncid;
}

/* device_distanceinMeters (U:25-33): squared distance in binary32. */
static inline float d2_of(const orc_ctx *c, uint32_t p, uint32_t q) {
    volatile float dx = c->x[q] - c->x[p];
    volatile float dy = c->y[q] - c->y[p];
    volatile float sx = dx * dx;
    volatile float sy = dy * dy;
    return sx + sy;
}

/* The one rounding question the reference's source leaves open (U:31): how
 * nvcc evaluated pow(x2-x1, 2) + pow(y2-y1, 2) (a float or a double pow,
 * and whether the sum was contracted into a fused multiply-add).  Over
 * every distance test of the level-2 scan (K:24-41: each instance against
 * every instance of a later feature in its 3x3 cells), counts the tests
 * where another evaluation decides d2 <= r2 differently from the canonical
 * binary32 form fl(fl(dx*dx) + fl(dy*dy)):
 *   out[0] tests, out[1] neighbours under the canonical form,
 *   out[2] fmaf(dx, dx, fl(dy*dy)) decides differently,
 *   out[3] fmaf(dy, dy, fl(dx*dx)) decides differently,
 *   out[4] fl32(double(dx)*dx + double(dy)*dy) decides differently,
 *   out[5] at least one of the three does.
 * CUDA's pow is not reproduced: each form is an IEEE operation sequence. */
void orc_boundary_pairs(const orc_ctx *c, uint64_t *out) {
    uint64_t t = 0, nb = 0, fa = 0, fb = 0, db = 0, any = 0;
    const uint32_t F = c->F;
    #pragma omp parallel for schedule(dynamic, 1024) reduction(+:t,nb,fa,fb,db,any)
    for (int64_t pi = 0; pi < (int64_t)c->n; pi++) {
        const uint32_t p = (uint32_t)pi, fp = c->fid[p];
        if (fp + 1 >= F) continue;
        const uint32_t cur = c->cellId[p];
        for (int32_t j = -1; j <= 1; j++)
            for (int32_t i = -1; i <= 1; i++) {
                int32_t ncl = orc_nbr_cell(c->gx, c->cells, cur, i, j);
                if (ncl < 0) continue;
                const uint64_t base = (uint64_t)ncl * F;
                const int32_t b = c->start[base + fp + 1];
                const int32_t e = c->start[base + F - 1] + c->cnt[base + F - 1];
                for (int32_t pos = b; pos < e; pos++) {
                    const uint32_t q = c->sortedIdx[pos];
                    volatile float dx = c->x[q] - c->x[p];
                    volatile float dy = c->y[q] - c->y[p];
                    volatile float sx = dx * dx;
                    volatile float sy = dy * dy;
                    volatile float d2 = sx + sy;
                    const int canon = d2 <= c->r2;
                    const int ra = fmaf(dx, dx, sy) <= c->r2;
                    const int rb = fmaf(dy, dy, sx) <= c->r2;
                    volatile double dd = (double)dx * (double)dx + (double)dy * (double)dy;
                    const int rd = (float)dd <= c->r2;
                    t++;
                    nb += (uint64_t)canon;
                    fa += (uint64_t)(ra != canon);
                    fb += (uint64_t)(rb != canon);
                    db += (uint64_t)(rd != canon);
                    any += (uint64_t)(ra != canon || rb != canon || rd != canon);
                }
            }
    }
    out[0] = t; out[1] = nb; out[2] = fa; out[3] = fb; out[4] = db; out[5] = any;
}

/* ---------------------------------------------------------------- level 2 */

/* degree2TableSlotCounter_Kernel (K:12-55) for candidate (a, b): one pass
 * per a-instance.  bitmap: the reference's N_a + N_b 0/1 words with the
 * overlapping layout (bitIndex2 = q - start_b + N_a - 1, K:45-47).  exact:
 * N_a + N_b words without overlap (a at t, b at N_a + rank).  Either may be
 * NULL.  Returns the number of rows. */
uint64_t orc_join2_count(const orc_ctx *c, uint32_t a, uint32_t b, uint32_t *slots,
                         uint32_t *bitmap, uint32_t *exact) {
    uint32_t sa = c->fstart[a], ea = c->fend[a], Na = c->fcount[a];
    uint32_t sb = c->fstart[b];
    uint64_t total = 0;
    #pragma omp parallel for schedule(dynamic, 256) reduction(+:total)
    for (int64_t t = 0; t < (int64_t)Na; t++) {
        uint32_t g = sa + (uint32_t)t;
        uint32_t cur = c->cellId[g];
        uint32_t cnt = 0;
        for (int32_t j = -1; j <= 1; j++)
            for (int32_t i = -1; i <= 1; i++) {
                int32_t nc = orc_nbr_cell(c->gx, c->cells, cur, i, j);
                if (nc < 0) continue;
                uint64_t idx = (uint64_t)nc * c->F + b;
                int32_t k = c->cnt[idx];
                if (k <= 0) continue;
                int32_t s0 = c->start[idx];
                for (int32_t pos = s0; pos < s0 + k; pos++) {
                    uint32_t q = c->sortedIdx[pos];
                    if (d2_of(c, g, q) <= c->r2) {
                        cnt++;
                        if (bitmap) {
                            #pragma omp atomic write
                            bitmap[t] = 1;
                            #pragma omp atomic write
                            bitmap[(uint64_t)(q - sb) + (ea - sa)] = 1;
                        }
                        if (exact) {
                            #pragma omp atomic write
                            exact[t] = 1;
                            #pragma omp atomic write
                            exact[(uint64_t)Na + (q - sb)] = 1;
                        }
                    }
                }
            }
        if (slots) slots[t] = cnt;
        total += cnt;
    }
    (void)ea;
    return total;
}

/* Exclusive scan of slot counts scaled by the degree (host
 * thrust::exclusive_scan cpp:391 + scalebyConstant K:387-396), 64-bit. */
static uint64_t *scan_offsets(const uint32_t *slots, uint64_t n, uint32_t degree) {
    uint64_t *off = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    uint64_t run = 0;
    for (uint64_t i = 0; i < n; i++) { off[i] = run * degree; run += slots[i]; }
    return off;
}

/* degree2TableGenerator_Kernel (K:56-100): writes rows (g, q) of the
 * level-2 table in a-instance order, neighbours in scan order. */
void orc_join2_write(const orc_ctx *c, uint32_t a, uint32_t b, const uint32_t *slots, uint32_t *table) {
    uint32_t sa = c->fstart[a], Na = c->fcount[a];
    uint64_t *off = scan_offsets(slots, Na, 2);
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t t = 0; t < (int64_t)Na; t++) {
        if (slots[t] == 0) continue;
        uint32_t g = sa + (uint32_t)t;
        uint64_t w = off[t];
        uint32_t cur = c->cellId[g];
        for (int32_t j = -1; j <= 1; j++)
            for (int32_t i = -1; i <= 1; i++) {
                int32_t nc = orc_nbr_cell(c->gx, c->cells, cur, i, j);
                if (nc < 0) continue;
                uint64_t idx = (uint64_t)nc * c->F + b;
                int32_t k = c->cnt[idx];
                if (k <= 0) continue;
                int32_t s0 = c->start[idx];
                for (int32_t pos = s0; pos < s0 + k; pos++) {
                    uint32_t q = c->sortedIdx[pos];
                    if (d2_of(c, g, q) <= c->r2) { table[w++] = g; table[w++] = q; }
                }
            }
    }
    free(off);
}

/* ---------------------------------------------------------------- level k */

/* generalInstanceTableSlotCounter_Kernel (K:255-316): one pass per parent
 * row; the anchor is the row's first instance; a last-feature instance q is
 * accepted when d2(row[m], q) <= r2 for every m (tested in order, K:282-290).
 * bitmap: sum_j N_{p_j} 0/1 words, member m at value - start + offset
 * (K:298-310).  The same layout has no overlap, so it serves both PI modes. */
uint64_t orc_joink_count(const orc_ctx *c, const uint32_t *parent, uint64_t R, uint32_t k,
                         const uint32_t *cand, uint32_t *slots, uint32_t *bitmap) {
    uint32_t km1 = k - 1, last = cand[k - 1];
    uint64_t offs[64];
    uint64_t o = 0;
    for (uint32_t m = 0; m < k; m++) { offs[m] = o; o += c->fcount[cand[m]]; }
    uint64_t total = 0;
    #pragma omp parallel for schedule(dynamic, 256) reduction(+:total)
    for (int64_t rI = 0; rI < (int64_t)R; rI++) {
        const uint32_t *row = parent + (uint64_t)rI * km1;
        uint32_t cur = c->cellId[row[0]];
        uint32_t cnt = 0;
        for (int32_t j = -1; j <= 1; j++)
            for (int32_t i = -1; i <= 1; i++) {
                int32_t nc = orc_nbr_cell(c->gx, c->cells, cur, i, j);
                if (nc < 0) continue;
                uint64_t idx = (uint64_t)nc * c->F + last;
                int32_t kk = c->cnt[idx];
                if (kk <= 0) continue;
                int32_t s0 = c->start[idx];
                for (int32_t pos = s0; pos < s0 + kk; pos++) {
                    uint32_t q = c->sortedIdx[pos];
                    int ok = 1;
                    for (uint32_t m = 0; m < km1; m++)
                        if (d2_of(c, row[m], q) > c->r2) { ok = 0; break; }
                    if (!ok) continue;
                    cnt++;
                    if (bitmap) {
                        for (uint32_t m = 0; m < km1; m++) {
                            #pragma omp atomic write
                            bitmap[offs[m] + (row[m] - c->fstart[cand[m]])] = 1;
                        }
                        #pragma omp atomic write
                        bitmap[offs[km1] + (q - c->fstart[last])] = 1;
                    }
                }
            }
        if (slots) slots[rI] = cnt;
        total += cnt;
    }
    return total;
}

/* generateInstanceTableGeneral_Kernel (K:318-385): rows row ++ [q]. */
void orc_joink_write(const orc_ctx *c, const uint32_t *parent, uint64_t R, uint32_t k,
                     const uint32_t *cand, const uint32_t *slots, uint32_t *table) {
    uint32_t km1 = k - 1, last = cand[k - 1];
    uint64_t *off = scan_offsets(slots, R, k);
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t rI = 0; rI < (int64_t)R; rI++) {
        if (slots[rI] == 0) continue;
        const uint32_t *row = parent + (uint64_t)rI * km1;
        uint64_t w = off[rI];
        uint32_t cur = c->cellId[row[0]];
        for (int32_t j = -1; j <= 1; j++)
            for (int32_t i = -1; i <= 1; i++) {
                int32_t nc = orc_nbr_cell(c->gx, c->cells, cur, i, j);
                if (nc < 0) continue;
                uint64_t idx = (uint64_t)nc * c->F + last;
                int32_t kk = c->cnt[idx];
                if (kk <= 0) continue;
                int32_t s0 = c->start[idx];
                for (int32_t pos = s0; pos < s0 + kk; pos++) {
                    uint32_t q = c->sortedIdx[pos];
                    int ok = 1;
                    for (uint32_t m = 0; m < km1; m++)
                        if (d2_of(c, row[m], q) > c->r2) { ok = 0; break; }
                    if (!ok) continue;
                    for (uint32_t m = 0; m < km1; m++) table[w++] = row[m];
                    table[w++] = q;
                }
            }
    }
    free(off);
}

/* -------------------------------------------------------------------- PI */

/* getParticipationIndex (cu:108-132): window j sums words [from, to-1) with
 * from/to advanced as in cu:116-123 (off by one, the compat numerators);
 * pr_j = fl32(sum) / fl32(N_j); PI = min(1, min_j pr_j). */
float orc_pi_compat(const orc_ctx *c, const uint32_t *bitmap, uint32_t k, const uint32_t *cand, uint64_t *num) {
    float PI = 1.0f;
    uint32_t from = 0, to = 0;
    for (uint32_t j = 0; j < k; j++) {
        uint32_t N = c->fcount[cand[j]];
        if (j == 0) { from = 0; to = N; }
        else { from = to - 1; to = from + N + 1; }
        uint32_t s = 0;
        for (uint32_t w = from; w < to - 1; w++) s += bitmap[w];
        if (num) num[j] = s;
        volatile float pr = (float)s / (float)N;
        if (pr < PI) PI = pr;
    }
    return PI;
}

/* Exact numerators: full per-feature windows of a non-overlapping bitmap. */
float orc_pi_exact(const orc_ctx *c, const uint32_t *bitmap, uint32_t k, const uint32_t *cand, uint64_t *num) {
    float PI = 1.0f;
    uint64_t o = 0;
    for (uint32_t j = 0; j < k; j++) {
        uint32_t N = c->fcount[cand[j]];
        uint64_t s = 0;
        for (uint64_t w = 0; w < N; w++) s += bitmap[o + w];
        o += N;
        if (num) num[j] = s;
        volatile float pr = (float)s / (float)N;
        if (pr < PI) PI = pr;
    }
    return PI;
}

/* ---------------------------------------------------------------- filter */

static int threshold_counts(const orc_ctx *c, const uint32_t *cand, uint32_t k, float theta, const uint32_t *cnt_out) {
    float PI = 1.0f;
    for (uint32_t t = 0; t < k; t++) {
        volatile float ub = (float)((double)cnt_out[t] / (double)(float)c->fcount[cand[t]]);
        if (t == 0) PI = ub; else if (PI > ub) PI = ub;
    }
    return PI >= theta ? 1 : 0;
}

/* filter2_Kernel (K:469-509) + threshold_Kernel (K:545-573) for one
 * candidate: quadrants (qx,qy) in [0,gx-2]x[0,gy-2]; a quadrant qualifies
 * when every feature has a non-zero count in one of its 4 cells; each cell of
 * a qualifying quadrant is counted once (isCellProcessed); ub_t =
 * fl32(double(cnt_t) / double(fl32(N_t))); pass iff min_t ub_t >= theta. */
static int filter_one(const orc_ctx *c, const uint32_t *cand, uint32_t k, float theta,
                      uint32_t *stamp, uint32_t stampv, uint32_t *cnt_out) {
    for (uint32_t t = 0; t < k; t++) cnt_out[t] = 0;
    if (c->gx < 2 || c->gy < 2) goto thr;
    for (uint32_t qy = 0; qy + 1 < c->gy; qy++)
        for (uint32_t qx = 0; qx + 1 < c->gx; qx++) {
            int done = 0;
            for (uint32_t t = 0; t < k && !done; t++) {
                int32_t s = 0;
                for (uint32_t dy = 0; dy < 2; dy++)
                    for (uint32_t dx = 0; dx < 2; dx++) {
                        uint32_t cid = (qy + dy) * c->gx + qx + dx;
                        s += c->cnt[(uint64_t)cid * c->F + cand[t]];
                    }
                if (!s) done = 1;
            }
            if (done) continue;
            for (uint32_t dy = 0; dy < 2; dy++)
                for (uint32_t dx = 0; dx < 2; dx++) {
                    uint32_t cid = (qy + dy) * c->gx + qx + dx;
                    if (stamp[cid] == stampv) continue;
                    stamp[cid] = stampv;
                    for (uint32_t t = 0; t < k; t++) cnt_out[t] += (uint32_t)c->cnt[(uint64_t)cid * c->F + cand[t]];
                }
        }
thr:
    return threshold_counts(c, cand, k, theta, cnt_out);
}

/* Same result as filter_one, visiting only quadrants that contain a cell
 * where the first feature occurs (every qualifying quadrant does).
 * f0cells: the non-empty cells of cand[0], nf0 of them. */
static int filter_fast(const orc_ctx *c, const uint32_t *cand, uint32_t k, float theta,
                       const uint32_t *f0cells, uint64_t nf0,
                       uint32_t *qstamp, uint32_t *cstamp, uint32_t stampv, uint32_t *cnt_out) {
    for (uint32_t t = 0; t < k; t++) cnt_out[t] = 0;
    if (c->gx >= 2 && c->gy >= 2) {
        for (uint64_t u = 0; u < nf0; u++) {
            uint32_t cid = f0cells[u];
            uint32_t cy = cid / c->gx, cx = cid - cy * c->gx;
            for (uint32_t dqy = 0; dqy < 2; dqy++)
                for (uint32_t dqx = 0; dqx < 2; dqx++) {
                    if (cy < dqy || cx < dqx) continue;
                    uint32_t qy = cy - dqy, qx = cx - dqx;
                    if (qy + 1 >= c->gy || qx + 1 >= c->gx) continue;
                    uint32_t qid = qy * (c->gx - 1) + qx;
                    if (qstamp[qid] == stampv) continue;
                    qstamp[qid] = stampv;
                    int ok = 1;
                    for (uint32_t t = 0; t < k && ok; t++) {
                        int32_t s = 0;
                        for (uint32_t dy = 0; dy < 2; dy++)
                            for (uint32_t dx = 0; dx < 2; dx++)
                                s += c->cnt[(uint64_t)((qy + dy) * c->gx + qx + dx) * c->F + cand[t]];
                        if (!s) ok = 0;
                    }
                    if (!ok) continue;
                    for (uint32_t dy = 0; dy < 2; dy++)
                        for (uint32_t dx = 0; dx < 2; dx++) {
                            uint32_t cc = (qy + dy) * c->gx + qx + dx;
                            if (cstamp[cc] == stampv) continue;
                            cstamp[cc] = stampv;
                            for (uint32_t t = 0; t < k; t++) cnt_out[t] += (uint32_t)c->cnt[(uint64_t)cc * c->F + cand[t]];
                        }
                }
        }
    }
    return threshold_counts(c, cand, k, theta, cnt_out);
}

/* Literal quadrant sweep over the whole grid (K:469-509); O(C * cells). */
void orc_filter_literal(const orc_ctx *c, const uint32_t *cands, uint32_t C, uint32_t k, float theta,
                        uint32_t *counts, uint8_t *pass) {
    uint32_t *stamp = (uint32_t *)calloc((uint64_t)c->cells + 1, sizeof(uint32_t));
    for (uint32_t i = 0; i < C; i++)
        pass[i] = (uint8_t)filter_one(c, cands + (uint64_t)i * k, k, theta, stamp, i + 1, counts + (uint64_t)i * k);
    free(stamp);
}

/* Runs the filter for C candidates of size k; counts: C*k, pass: C. */
void orc_filter(const orc_ctx *c, const uint32_t *cands, uint32_t C, uint32_t k, float theta,
                uint32_t *counts, uint8_t *pass) {
    /* non-empty cells per feature, ascending */
    uint64_t *fo = (uint64_t *)calloc((uint64_t)c->F + 1, sizeof(uint64_t));
    for (uint64_t cid = 0; cid < c->cells; cid++)
        for (uint32_t f = 0; f < c->F; f++) if (c->cnt[cid * c->F + f] > 0) fo[f + 1]++;
    for (uint32_t f = 0; f < c->F; f++) fo[f + 1] += fo[f];
    uint32_t *fc = (uint32_t *)malloc(sizeof(uint32_t) * (fo[c->F] + 1));
    uint64_t *fw = (uint64_t *)malloc(sizeof(uint64_t) * (c->F + 1));
    memcpy(fw, fo, sizeof(uint64_t) * (c->F + 1));
    for (uint64_t cid = 0; cid < c->cells; cid++)
        for (uint32_t f = 0; f < c->F; f++) if (c->cnt[cid * c->F + f] > 0) fc[fw[f]++] = (uint32_t)cid;
    free(fw);
    #pragma omp parallel
    {
        uint32_t *qstamp = (uint32_t *)calloc((uint64_t)c->cells + 1, sizeof(uint32_t));
        uint32_t *cstamp = (uint32_t *)calloc((uint64_t)c->cells + 1, sizeof(uint32_t));
        #pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < (int64_t)C; i++) {
            const uint32_t *cand = cands + (uint64_t)i * k;
            pass[i] = (uint8_t)filter_fast(c, cand, k, theta, fc + fo[cand[0]], fo[cand[0] + 1] - fo[cand[0]],
                                           qstamp, cstamp, (uint32_t)i + 1, counts + (uint64_t)i * k);
        }
        free(qstamp);
        free(cstamp);
    }
    free(fo);
    free(fc);
}

/* -------------------------------------------------------- candidate gen */

/* subsetGen (cpp:598-616) for n = degree, size degree-1: lexicographic by
 * position, so subset s omits position degree-1-s. checkSubset (cpp:618-634)
 * is a linear search of the prevalent list. */
static int in_list(const uint32_t *L, uint64_t P, uint32_t w, const uint32_t *v) {
    for (uint64_t i = 0; i < P; i++) {
        uint32_t j = 0;
        for (; j < w; j++) if (L[i * w + j] != v[j]) break;
        if (j == w) return 1;
    }
    return 0;
}

/* candidateColocationGeneral (cpp:517-557): L holds P patterns of size
 * km1 = degree-1 in list order; out receives the size-degree candidates.
 * The j loop ends at the first j whose prefix differs, and also at the
 * first candidate with a missing subset (cpp:541-542).  Returns the count. */
uint64_t orc_candgen(const uint32_t *L, uint64_t P, uint32_t km1, uint32_t *out, uint64_t cap) {
    uint32_t degree = km1 + 1, lastIndex = degree - 2;
    uint64_t nout = 0;
    uint32_t inter[64], sub[64];
    for (uint64_t i = 0; i < P; i++) {
        for (uint64_t j = i + 1; j < P; j++) {
            int same = 1;
            for (uint32_t k = 0; k < lastIndex; k++)
                if (L[i * km1 + k] != L[j * km1 + k]) { same = 0; break; }
            if (!(same && L[i * km1 + lastIndex] < L[j * km1 + lastIndex])) break;
            for (uint32_t t = 0; t < km1; t++) inter[t] = L[i * km1 + t];
            inter[km1] = L[j * km1 + lastIndex];
            int ok = 1;
            for (uint32_t s = 2; s < degree && ok; s++) {
                uint32_t omit = degree - 1 - s, w = 0;
                for (uint32_t t = 0; t < degree; t++) if (t != omit) sub[w++] = inter[t];
                ok = in_list(L, P, km1, sub);
            }
            if (!ok) break;
            if (nout < cap) memcpy(out + nout * degree, inter, sizeof(uint32_t) * degree);
            nout++;
        }
    }
    return nout;
}

/* ---------------------------------------------------------------- hashes */

/* Table checksum independent of row order: the 64-bit sum over rows of a
 * hash of the row's instance ids in column order (splitmix64 finaliser as
 * the mixing step).  A report carries it per instance table, so two miners
 * that build the same rows in different orders report the same value. */
static inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
uint64_t orc_row_hash(const uint32_t *row, uint32_t k) {
    uint64_t h = 0x2545F4914F6CDD1Dull ^ k;
    for (uint32_t m = 0; m < k; m++) h = mix64(h ^ (((uint64_t)m << 32) | row[m]));
    return h;
}
uint64_t orc_table_hash(const uint32_t *t, uint64_t rows, uint32_t k) {
    uint64_t h = 0;
    for (uint64_t i = 0; i < rows; i++) h += orc_row_hash(t + i * k, k);
    return h;
}

/* ----------------------------------------------------------------- miner */

typedef struct { char *s; size_t n, cap; } sbuf;
static void sb_printf(sbuf *b, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
#include <stdarg.h>
static void sb_printf(sbuf *b, const char *fmt, ...) {
    va_list ap;
    for (;;) {
        va_start(ap, fmt);
        int k = vsnprintf(b->s + b->n, b->cap - b->n, fmt, ap);
        va_end(ap);
        if (k >= 0 && (size_t)k < b->cap - b->n) { b->n += (size_t)k; return; }
        b->cap = b->cap * 2 + (size_t)(k > 0 ? k : 256) + 256;
        b->s = (char *)realloc(b->s, b->cap);
    }
}

static uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

typedef struct {
    uint32_t *pat;      /* k features */
    uint32_t *slots;    /* per anchor / parent row */
    uint64_t nslots;
    uint64_t rows;
    uint32_t *table;    /* generated table, rows*k, or NULL */
    uint8_t *part;      /* exact participation, 0/1 per instance of each
                           member in input rank order, members back to back
                           (kept with subset_bound) */
} prev_t;

static void write_patterns(const orc_ctx *c, const char *prefix, uint32_t k, const prev_t *L, uint64_t P) {
    if (!prefix) return;
    char path[4096];
    snprintf(path, sizeof path, "%sColocationDegree_%u.txt", prefix, k);
    FILE *fp = fopen(path, "wb");
    if (!fp) return;
    for (uint64_t i = 0; i < P; i++) {          /* savetoFile2/savetoFileGen, cpp:960-997 */
        fputs("( ", fp);
        for (uint32_t j = 0; j < k; j++) {
            fputs(c->labels[L[i].pat[j]], fp);
            fputs(" ", fp);
            fputs(j != k - 1 ? "| " : ")\n", fp);
        }
    }
    fclose(fp);
}

/* tableGenRequired (cpp:479-515): join over the current candidates with the
 * subset check against the candidates themselves. */
static int ref_need_table(const uint32_t *C, uint64_t nC, uint32_t k) {
    uint32_t tmp[64];
    return orc_candgen(C, nC, k, tmp, 0) > 0 ? 1 : 0;
}

static int64_t find_pattern(const prev_t *L, uint64_t P, uint32_t w, const uint32_t *v) {
    for (uint64_t i = 0; i < P; i++) {          /* getIndex, cpp:703-718 */
        uint32_t j = 0;
        for (; j < w; j++) if (L[i].pat[j] != v[j]) break;
        if (j == w) return (int64_t)i;
    }
    return -1;
}

/* Begin level loop (cpp:1104-1137) with degree2Processing (cpp:256-425),
 * generatePrevalentPatternsGeneral (cpp:736-854) and
 * generateInstanceTableGeneral (cpp:856-909).  Tables are built for the
 * prefixes of the next level's candidates.  When the reference's
 * needInstanceTable test (Q10) would skip tables that the next level needs,
 * the report records it on a "q10" line and the tables are built anyway.
 * pi_mode 0 = compat numerators (cu:108-132), 1 = exact.  max_k = 0 means no
 * cap.  subset_bound = 1 skips, from level 3 on, the candidates whose subset
 * bound (subset_bound_q) is below theta: their report line ends with the
 * bound after "q", and every candidate that passes both tests carries it
 * too.  Returns a malloc'd report (free with orc_free_str). */
static double wall_s(void) { return omp_get_wtime(); }

/* Subset bound of a size-k candidate (k >= 3).  An instance of member f
 * that takes part in a row of the candidate's table takes part in a row of
 * every (k-1)-subset S containing f (drop the other column), so
 *   num_exact(f) <= cnt_f = |intersection over S of part_f(S)|.
 * A reference numerator (pi_mode 0, cu:108-132 with the back-to-back layout
 * of K:298-310) is num_j = P_{j-1}[N_{j-1}-1] + pop(P_j) - P_j[N_j-1] for
 * j >= 1 and pop(P_0) - P_0[N_0-1] for j = 0, so
 *   num_compat(j) <= cnt_j + last_{j-1},
 * last_{j-1} = whether instance N-1 of member j-1 lies in its intersection.
 * Both PIs are therefore at most
 *   q = min(1, min_j fl32(cnt_j + t_j) / fl32(N_j)),  t_j = 0 (exact) or
 *   last_{j-1} for j >= 1 (reference numerators),
 * and a candidate with q < theta is not prevalent.  Every (k-1)-subset of a
 * candidate is a prevalent pattern of the previous level (the Apriori
 * subset test of cpp:517-557), whose participation Lprev holds.  Returns q;
 * no counterpart in the reference. */
static float subset_bound_q(const orc_ctx *c, const prev_t *Lprev, uint64_t Pprev, uint32_t k, const uint32_t *cand,
                            int pi_mode) {
    const uint8_t *win[64];
    uint32_t sub[64];
    uint64_t cnt[64];
    int last[64];
    for (uint32_t j = 0; j < k; j++) {
        uint32_t nw = 0;
        for (uint32_t i = 0; i < k; i++) {
            if (i == j) continue;
            uint32_t w = 0;
            for (uint32_t t = 0; t < k; t++) if (t != i) sub[w++] = cand[t];
            int64_t s = find_pattern(Lprev, Pprev, k - 1, sub);
            if (s < 0 || !Lprev[s].part) return 1.0f;   /* not reached: candgen keeps only such candidates */
            uint32_t pos = j < i ? j : j - 1;
            uint64_t off = 0;
            for (uint32_t t = 0; t < pos; t++) off += c->fcount[sub[t]];
            win[nw++] = Lprev[s].part + off;
        }
        uint32_t N = c->fcount[cand[j]];
        uint64_t n1 = 0;
        for (uint32_t w = 0; w < N; w++) {
            uint8_t a = 1;
            for (uint32_t t = 0; t < nw && a; t++) a &= win[t][w];
            n1 += a;
        }
        cnt[j] = n1;
        uint8_t a = N > 0;
        for (uint32_t t = 0; t < nw && a; t++) a &= win[t][N - 1];
        last[j] = a;
    }
    float q = 1.0f;
    for (uint32_t j = 0; j < k; j++) {
        uint64_t v = cnt[j] + (pi_mode == 0 && j >= 1 ? (uint64_t)last[j - 1] : 0);
        volatile float b = (float)v / (float)c->fcount[cand[j]];
        if (b < q) q = b;
    }
    return q;
}

char *orc_mine(orc_ctx *c, float r, float theta, uint32_t max_k, int pi_mode, int filter_on,
               int hash_tables, int subset_bound, const char *out_prefix, int *err) {
    sbuf b = {0};
    b.cap = 4096; b.s = (char *)malloc(b.cap); b.s[0] = 0;
    *err = ORC_OK;
    if (!c->binned || c->r != r) {
        int e = orc_bin(c, r);
        if (e) { *err = e; free(b.s); return NULL; }
    }
    sb_printf(&b, "colo-report v1\n");
    sb_printf(&b, "grid gx=%u gy=%u cells=%u n=%u F=%u r2=%08x theta=%08x\n", c->gx, c->gy, c->cells,
              c->n, c->F, fbits(c->r2), fbits(theta));
    uint32_t F = c->F;
    /* level-2 candidates: all pairs in order (cpp:75-82) */
    uint64_t nC = F >= 2 ? (uint64_t)F * (F - 1) / 2 : 0;
    uint32_t *C = (uint32_t *)malloc(sizeof(uint32_t) * (nC ? nC * 2 : 1));
    { uint64_t w = 0; for (uint32_t i = 0; i + 1 < F; i++) for (uint32_t j = i + 1; j < F; j++) { C[w++] = i; C[w++] = j; } }
    prev_t *Lprev = NULL; uint64_t Pprev = 0;
    uint32_t k = 2;
    uint64_t total_prev = 0;
    uint32_t levels = 0;
    /* ORC_PROGRESS set: a line on stderr at least every 10 s of candidates
     * (long runs under a supervisor that kills silent processes) */
    const int progress = getenv("ORC_PROGRESS") != NULL;
    double last_note = wall_s();
    while (nC > 0 && k <= F && (max_k == 0 || k <= max_k)) {
        levels++;
        uint32_t *fcnt = (uint32_t *)calloc(nC * k, sizeof(uint32_t));
        uint8_t *pass = (uint8_t *)malloc(nC);
        if (filter_on) orc_filter(c, C, (uint32_t)nC, k, theta, fcnt, pass);
        else memset(pass, 1, nC);
        uint64_t npass = 0;
        for (uint64_t i = 0; i < nC; i++) npass += pass[i];
        float *qv = NULL;
        if (subset_bound && k >= 3) {
            qv = (float *)malloc(sizeof(float) * nC);
#pragma omp parallel for schedule(dynamic, 1)
            for (int64_t i = 0; i < (int64_t)nC; i++)
                qv[i] = pass[i] ? subset_bound_q(c, Lprev, Pprev, k, C + i * k, pi_mode) : 1.0f;
        }
        prev_t *L = (prev_t *)calloc(nC, sizeof(prev_t));
        uint64_t P = 0;
        sbuf lb = {0}; lb.cap = 4096; lb.s = (char *)malloc(lb.cap); lb.s[0] = 0;
        for (uint64_t i = 0; i < nC; i++) {
            const uint32_t *cand = C + i * k;
            if (progress && wall_s() - last_note >= 10.0) {
                last_note = wall_s();
                fprintf(stderr, "[oracle] level %u: candidate %llu of %llu\n", k, (unsigned long long)i,
                        (unsigned long long)nC);
            }
            sb_printf(&lb, "c");
            for (uint32_t m = 0; m < k; m++) sb_printf(&lb, " %u", cand[m]);
            sb_printf(&lb, " f");
            for (uint32_t m = 0; m < k; m++) sb_printf(&lb, " %u", fcnt[i * k + m]);
            sb_printf(&lb, " p %u", (unsigned)pass[i]);
            if (!pass[i]) { sb_printf(&lb, "\n"); continue; }
            if (qv) {
                sb_printf(&lb, " q %08x", fbits(qv[i]));
                if (qv[i] < theta) { sb_printf(&lb, "\n"); continue; }
            }
            uint64_t wsz = 0;
            for (uint32_t m = 0; m < k; m++) wsz += c->fcount[cand[m]];
            uint32_t *bm = (uint32_t *)calloc(wsz + 1, sizeof(uint32_t));
            uint32_t *ex = NULL;
            uint32_t *slots;
            uint64_t nslots, rows;
            if (k == 2) {
                ex = (uint32_t *)calloc(wsz + 1, sizeof(uint32_t));
                nslots = c->fcount[cand[0]];
                slots = (uint32_t *)malloc(sizeof(uint32_t) * (nslots ? nslots : 1));
                rows = orc_join2_count(c, cand[0], cand[1], slots, bm, ex);
            } else {
                int64_t pi = find_pattern(Lprev, Pprev, k - 1, cand);
                if (pi < 0 || !Lprev[pi].table) { *err = ORC_EARG; free(bm); free(lb.s); free(b.s); return NULL; }
                nslots = Lprev[pi].rows;
                slots = (uint32_t *)malloc(sizeof(uint32_t) * (nslots ? nslots : 1));
                rows = orc_joink_count(c, Lprev[pi].table, nslots, k, cand, slots, bm);
                ex = bm;
            }
            uint64_t num[64], exn[64];
            float pic = orc_pi_compat(c, bm, k, cand, num);
            float pie = orc_pi_exact(c, ex, k, cand, exn);
            float pim = pi_mode ? pie : pic;
            int prevalent = !(pim < theta);     /* cpp:335, 827 */
            sb_printf(&lb, " r %llu n", (unsigned long long)rows);
            for (uint32_t m = 0; m < k; m++) sb_printf(&lb, " %llu", (unsigned long long)num[m]);
            sb_printf(&lb, " e");
            for (uint32_t m = 0; m < k; m++) sb_printf(&lb, " %llu", (unsigned long long)exn[m]);
            sb_printf(&lb, " pi %08x %08x v %d\n", fbits(pic), fbits(pie), prevalent);
            if (prevalent && subset_bound) {
                L[P].part = (uint8_t *)malloc(wsz + 1);
                for (uint64_t w = 0; w < wsz; w++) L[P].part[w] = ex[w] ? 1 : 0;
            }
            if (ex != bm) free(ex);
            free(bm);
            if (prevalent) {
                L[P].pat = (uint32_t *)malloc(sizeof(uint32_t) * k);
                memcpy(L[P].pat, cand, sizeof(uint32_t) * k);
                L[P].slots = slots;
                L[P].nslots = nslots;
                L[P].rows = rows;
                P++;
            } else {
                free(slots);
            }
        }
        sb_printf(&b, "level %u cand %llu pass %llu prev %llu\n", k, (unsigned long long)nC,
                  (unsigned long long)npass, (unsigned long long)P);
        if (progress)
            fprintf(stderr, "[oracle] level %u done: %llu candidates, %llu prevalent\n", k, (unsigned long long)nC,
                    (unsigned long long)P);
        sb_printf(&b, "%s", lb.s);
        free(lb.s);
        total_prev += P;
        write_patterns(c, out_prefix, k, L, P);
        /* next-level candidates */
        uint64_t nN = 0;
        uint32_t *N = NULL;
        if (max_k == 0 || k + 1 <= max_k) {
            uint32_t *flat = (uint32_t *)malloc(sizeof(uint32_t) * (P * k + 1));
            for (uint64_t i = 0; i < P; i++) memcpy(flat + i * k, L[i].pat, sizeof(uint32_t) * k);
            nN = orc_candgen(flat, P, k, NULL, 0);
            N = (uint32_t *)malloc(sizeof(uint32_t) * (nN * (k + 1) + 1));
            orc_candgen(flat, P, k, N, nN);
            free(flat);
            if (k >= 3 && nN > 0 && !ref_need_table(C, nC, k))
                sb_printf(&b, "q10 level %u\n", k);
        }
        /* tables for prefixes of next-level candidates (cpp:364-414, 1042-1066) */
        uint8_t *need = (uint8_t *)calloc(P + 1, 1);
        for (uint64_t i = 0; i < nN; i++) {
            int64_t t = find_pattern(L, P, k, N + i * (k + 1));
            if (t >= 0) need[t] = 1;
        }
        for (uint64_t i = 0; i < P; i++) {
            if (!need[i]) continue;
            L[i].table = (uint32_t *)malloc(sizeof(uint32_t) * (L[i].rows * k + 1));
            if (k == 2) orc_join2_write(c, L[i].pat[0], L[i].pat[1], L[i].slots, L[i].table);
            else {
                int64_t pi = find_pattern(Lprev, Pprev, k - 1, L[i].pat);
                orc_joink_write(c, Lprev[pi].table, Lprev[pi].rows, k, L[i].pat, L[i].slots, L[i].table);
            }
            sb_printf(&b, "t");
            for (uint32_t m = 0; m < k; m++) sb_printf(&b, " %u", L[i].pat[m]);
            sb_printf(&b, " rows %llu", (unsigned long long)L[i].rows);
            if (hash_tables) sb_printf(&b, " h %016llx", (unsigned long long)orc_table_hash(L[i].table, L[i].rows, k));
            sb_printf(&b, "\n");
        }
        free(need);
        for (uint64_t i = 0; i < P; i++) { free(L[i].slots); L[i].slots = NULL; }
        for (uint64_t i = 0; i < Pprev; i++) { free(Lprev[i].pat); free(Lprev[i].table); free(Lprev[i].part); }
        free(Lprev);
        Lprev = L; Pprev = P;
        free(fcnt); free(pass); free(qv); free(C);
        C = N; nC = nN;
        k++;
    }
    for (uint64_t i = 0; i < Pprev; i++) { free(Lprev[i].pat); free(Lprev[i].table); free(Lprev[i].part); }
    free(Lprev);
    free(C);
    sb_printf(&b, "end levels %u patterns %llu\n", levels, (unsigned long long)total_prev);
    return b.s;
}

void orc_free_str(char *s) { free(s); }

int orc_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void orc_set_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}
