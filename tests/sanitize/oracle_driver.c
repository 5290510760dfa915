/* Drives the CPU oracle (oracle/colo_oracle.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer: clustered random points, binning, complete
 * mining with table checksums and pattern files, both filter forms, the
 * joins, both PI forms, candidate generation, the boundary-pair count and
 * the CSV loader (valid and malformed files).  Exits non-zero on a failed
 * check; the sanitizers abort on a memory or UB error. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct orc_ctx orc_ctx;
orc_ctx *orc_load_csv(const char *path, int *err);
orc_ctx *orc_from_arrays(uint32_t n, uint32_t F, const float *x, const float *y, const uint32_t *fcount,
                         const float *bounds);
void orc_free(orc_ctx *c);
int orc_bin(orc_ctx *c, float r);
void orc_boundary_pairs(const orc_ctx *c, uint64_t *out);
uint64_t orc_join2_count(const orc_ctx *c, uint32_t a, uint32_t b, uint32_t *slots, uint32_t *bitmap, uint32_t *exact);
void orc_join2_write(const orc_ctx *c, uint32_t a, uint32_t b, const uint32_t *slots, uint32_t *table);
uint64_t orc_joink_count(const orc_ctx *c, const uint32_t *parent, uint64_t R, uint32_t k, const uint32_t *cand,
                         uint32_t *slots, uint32_t *bitmap);
void orc_joink_write(const orc_ctx *c, const uint32_t *parent, uint64_t R, uint32_t k, const uint32_t *cand,
                     const uint32_t *slots, uint32_t *table);
float orc_pi_compat(const orc_ctx *c, const uint32_t *bitmap, uint32_t k, const uint32_t *cand, uint64_t *num);
float orc_pi_exact(const orc_ctx *c, const uint32_t *bitmap, uint32_t k, const uint32_t *cand, uint64_t *num);
void orc_filter(const orc_ctx *c, const uint32_t *cands, uint32_t C, uint32_t k, float theta, uint32_t *counts,
                uint8_t *pass);
void orc_filter_literal(const orc_ctx *c, const uint32_t *cands, uint32_t C, uint32_t k, float theta,
                        uint32_t *counts, uint8_t *pass);
uint64_t orc_candgen(const uint32_t *L, uint64_t P, uint32_t km1, uint32_t *out, uint64_t cap);
char *orc_mine(orc_ctx *c, float r, float theta, uint32_t max_k, int pi_mode, int filter_on, int hash_tables,
               int subset_bound, const char *out_prefix, int *err);
void orc_free_str(char *s);

#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            return 1;                                                     \
        }                                                                 \
    } while (0)

static uint64_t rng = 88172645463325252ull;
static double urand(void) {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (double)(rng >> 11) / 9007199254740992.0;
}

int main(int argc, char **argv) {
    const char *tmp = argc > 1 ? argv[1] : "/tmp";
    enum { F = 5, PER = 400 };
    const uint32_t n = F * PER;
    float *x = malloc(sizeof(float) * n), *y = malloc(sizeof(float) * n);
    uint32_t fc[F];
    for (uint32_t f = 0; f < F; f++) {
        fc[f] = PER;
        for (uint32_t i = 0; i < PER; i++) {
            uint32_t p = f * PER + i;
            /* one instance in eight in a 60 x 60 cluster, the rest uniform */
            const int hot = i % 8 == 0;
            double cx = hot ? 200.0 : urand() * 950.0, cy = hot ? 300.0 : urand() * 950.0;
            x[p] = (float)(cx + (hot ? urand() * 60.0 : 0.0));
            y[p] = (float)(cy + (hot ? urand() * 60.0 : 0.0));
        }
    }
    const float bounds[4] = {0, 0, 1000, 1000};
    orc_ctx *c = orc_from_arrays(n, F, x, y, fc, bounds);
    CHECK(c && orc_bin(c, 12.0f) == 0);
    uint64_t bp[6];
    orc_boundary_pairs(c, bp);
    CHECK(bp[0] > 0 && bp[1] <= bp[0]);
    char prefix[512];
    snprintf(prefix, sizeof prefix, "%s/san_", tmp);
    int err = 0;
    char *rep = orc_mine(c, 12.0f, 0.05f, 0, 0, 1, 1, 0, prefix, &err);
    CHECK(rep && !err && strstr(rep, "end levels 4"));
    char *rep2 = orc_mine(c, 12.0f, 0.05f, 0, 1, 0, 1, 0, NULL, &err);
    CHECK(rep2 && !err);
    char *rep3 = orc_mine(c, 12.0f, 0.05f, 0, 0, 1, 1, 1, NULL, &err);
    CHECK(rep3 && !err && strstr(rep3, "end levels 4"));
    orc_free_str(rep);
    orc_free_str(rep2);
    orc_free_str(rep3);
    /* filter: both forms agree */
    uint32_t cands[2 * 10], C = 0;
    for (uint32_t a = 0; a < F; a++)
        for (uint32_t b = a + 1; b < F; b++) { cands[2 * C] = a; cands[2 * C + 1] = b; C++; }
    uint32_t c1[20], c2[20];
    uint8_t p1[10], p2[10];
    orc_filter(c, cands, C, 2, 0.2f, c1, p1);
    orc_filter_literal(c, cands, C, 2, 0.2f, c2, p2);
    CHECK(!memcmp(c1, c2, sizeof c1) && !memcmp(p1, p2, sizeof p1));
    /* level-2 and level-3 joins, both PI forms */
    uint32_t *slots = malloc(4 * PER), *bm = calloc(2 * PER + 1, 4), *ex = calloc(2 * PER + 1, 4);
    uint64_t rows = orc_join2_count(c, 0, 1, slots, bm, ex);
    uint32_t *t2 = malloc(4 * (2 * rows + 1));
    orc_join2_write(c, 0, 1, slots, t2);
    uint64_t num[3];
    uint32_t cand2[2] = {0, 1}, cand3[3] = {0, 1, 2};
    (void)orc_pi_compat(c, bm, 2, cand2, num);
    (void)orc_pi_exact(c, ex, 2, cand2, num);
    uint32_t *s3 = malloc(4 * (rows + 1)), *b3 = calloc(3 * PER + 1, 4);
    uint64_t r3 = orc_joink_count(c, t2, rows, 3, cand3, s3, b3);
    uint32_t *t3 = malloc(4 * (3 * r3 + 1));
    orc_joink_write(c, t2, rows, 3, cand3, s3, t3);
    (void)orc_pi_compat(c, b3, 3, cand3, num);
    (void)orc_pi_exact(c, b3, 3, cand3, num);
    /* candidate generation with the first-failure break */
    uint32_t L[10] = {0, 1, 0, 2, 0, 3, 1, 3, 2, 3}, out[9];
    CHECK(orc_candgen(L, 5, 2, out, 3) == 1 && out[0] == 0 && out[1] == 2 && out[2] == 3);
    free(slots); free(bm); free(ex); free(t2); free(s3); free(b3); free(t3);
    orc_free(c);
    /* CSV loader: a valid file and malformed ones */
    char path[512];
    snprintf(path, sizeof path, "%s/san.csv", tmp);
    const char *good = "4\n7,1.5,2.5\n7,3,4\n3,5,6\n3,\"7\",8\n\n0,0,10,10";
    const char *bad[] = {"3\n7,1,2\n\n0,0,10,10", "2\n7,1\n7,2,3\n\n0,0,1,1", "2\n7,1,2\n7,2,3\n", "x\n", ""};
    FILE *fp = fopen(path, "wb");
    fputs(good, fp);
    fclose(fp);
    orc_ctx *h = orc_load_csv(path, &err);
    CHECK(h && orc_bin(h, 2.0f) == 0);
    orc_free(h);
    for (unsigned i = 0; i < sizeof bad / sizeof bad[0]; i++) {
        fp = fopen(path, "wb");
        fputs(bad[i], fp);
        fclose(fp);
        h = orc_load_csv(path, &err);
        CHECK(!h);
    }
    free(x);
    free(y);
    printf("oracle sanitizer run ok (%llu tests, %llu rows at level 2)\n", (unsigned long long)bp[0],
           (unsigned long long)rows);
    return 0;
}
