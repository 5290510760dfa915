// colo_host.cpp -- host half of the C ABI that needs no GPU: the dataset CSV
// parser (populateData, Codes/colocationFinder.cpp:139-254) with its
// multi-threaded form and binary cache, the host dataset handle, and the
// Apriori candidate generation (candidateColocationGeneral, cpp:517-557).
// Plain C++17: built into libcolo.so by hipcc and, for the sanitizer tests,
// by g++ alone.
#include <errno.h>
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/colo.h"
#include "colo_host.h"

namespace colo_host {

// candidateColocationGeneral (cpp:517-557): L holds P patterns of size km1
// in list order.  For each i, j runs over the following patterns while they
// share L[i]'s prefix; a candidate whose (km1)-subsets at positions 2..km1
// of the lexicographic subset enumeration (cpp:598-616, 669-679) are not
// all in L ends the j loop (cpp:541-542).
std::vector<uint32_t> gen_candidates(const std::vector<uint32_t> &L, uint32_t km1) {
    std::vector<uint32_t> out;
    const uint64_t P = km1 ? L.size() / km1 : 0;
    if (P < 2) return out;
    PatSet S(L.data(), P, km1);
    const uint32_t deg = km1 + 1, li = deg - 2;
    std::vector<uint32_t> inter(deg), sub(km1);
    for (uint64_t i = 0; i < P; i++) {
        const uint32_t *a = &L[i * km1];
        for (uint64_t j = i + 1; j < P; j++) {
            const uint32_t *b = &L[j * km1];
            bool same = std::equal(a, a + li, b);
            if (!(same && a[li] < b[li])) break;
            std::copy(a, a + km1, inter.begin());
            inter[km1] = b[li];
            bool ok = true;
            for (uint32_t s = 2; s < deg && ok; s++) {
                uint32_t omit = deg - 1 - s, w = 0;
                for (uint32_t t = 0; t < deg; t++)
                    if (t != omit) sub[w++] = inter[t];
                ok = S.has(sub);
            }
            if (!ok) break;
            out.insert(out.end(), inter.begin(), inter.end());
        }
    }
    return out;
}

// stoi (libstdc++: strtol base 10) on a field.
bool parse_long(const char *s, long *v) {
    char *e;
    errno = 0;
    long x = strtol(s, &e, 10);
    if (e == s || errno) return false;
    *v = x;
    return true;
}
bool parse_flt(const char *s, float *v) {
    char *e;
    float x = strtof(s, &e);
    if (e == s) return false;
    *v = x;
    return true;
}

// Field split of one line (boost::escaped_list_separator on numeric rows,
// cpp:156,166): comma separated, surrounding double quotes dropped.
int split(char *line, char **f, int maxf) {
    int k = 0;
    char *p = line;
    while (k < maxf) {
        if (*p == '"') {
            char *e = strchr(p + 1, '"');
            if (!e) return -1;
            *e = 0;
            f[k++] = p + 1;
            p = e + 1;
            if (*p == ',') { p++; continue; }
            break;
        }
        f[k++] = p;
        char *cm = strchr(p, ',');
        if (!cm) break;
        *cm = 0;
        p = cm + 1;
    }
    return k;
}

}  // namespace colo_host

using colo_host::gen_candidates;
using colo_host::parse_flt;
using colo_host::parse_long;
using colo_host::split;

struct colo_host_data {
    std::vector<float> x, y;
    std::vector<uint32_t> fcount;
    std::vector<std::string> labels;
    std::vector<const char *> label_ptrs;
    float b[4] = {0, 0, 0, 0};
    uint32_t n = 0;
};

extern "C" {

void colo_free_str(char *s) { free(s); }

uint64_t colo_candidates(const uint32_t *L, uint64_t P, uint32_t km1, uint32_t *out, uint64_t cap) {
    if (!km1 || (!L && P)) return 0;
    std::vector<uint32_t> l(L, L + P * km1);
    std::vector<uint32_t> c = gen_candidates(l, km1);
    uint64_t n = c.size() / (km1 + 1);
    if (out) std::copy(c.begin(), c.begin() + std::min<uint64_t>(n, cap) * (km1 + 1), out);
    return n;
}

// Binary cache of a parsed dataset: magic, source size and mtime, n, F,
// bounds, per-feature counts and label lengths, label bytes, x, y.
static const char kCacheMagic[8] = {'C', 'O', 'L', 'O', 'B', 'I', 'N', '1'};

static bool source_stamp(const char *path, uint64_t *size, int64_t *mtime_ns) {
    struct stat st;
    if (stat(path, &st) != 0) return false;
    *size = (uint64_t)st.st_size;
    *mtime_ns = (int64_t)st.st_mtim.tv_sec * 1000000000ll + st.st_mtim.tv_nsec;
    return true;
}

static bool cache_read(const char *cache, uint64_t size, int64_t mtime, colo_host_data *h) {
    FILE *fp = fopen(cache, "rb");
    if (!fp) return false;
    bool ok = false;
    do {
        char mg[8];
        uint64_t sz;
        int64_t mt;
        uint32_t n, F;
        if (fread(mg, 1, 8, fp) != 8 || memcmp(mg, kCacheMagic, 8)) break;
        if (fread(&sz, 8, 1, fp) != 1 || fread(&mt, 8, 1, fp) != 1 || sz != size || mt != mtime) break;
        if (fread(&n, 4, 1, fp) != 1 || fread(&F, 4, 1, fp) != 1 || fread(h->b, 4, 4, fp) != 4) break;
        h->n = n;
        h->fcount.resize(F);
        std::vector<uint32_t> ll(F);
        if (F && (fread(h->fcount.data(), 4, F, fp) != F || fread(ll.data(), 4, F, fp) != F)) break;
        h->labels.clear();
        bool bad = false;
        for (uint32_t f = 0; f < F && !bad; f++) {
            std::string l(ll[f], '\0');
            if (ll[f] && fread(&l[0], 1, ll[f], fp) != ll[f]) bad = true;
            h->labels.push_back(std::move(l));
        }
        if (bad) break;
        h->x.resize(n);
        h->y.resize(n);
        if (n && (fread(h->x.data(), 4, n, fp) != n || fread(h->y.data(), 4, n, fp) != n)) break;
        ok = true;
    } while (false);
    fclose(fp);
    return ok;
}

static void cache_write(const char *cache, uint64_t size, int64_t mtime, const colo_host_data *h) {
    std::string tmp = std::string(cache) + ".tmp";
    FILE *fp = fopen(tmp.c_str(), "wb");
    if (!fp) return;
    uint32_t F = (uint32_t)h->fcount.size();
    bool ok = fwrite(kCacheMagic, 1, 8, fp) == 8 && fwrite(&size, 8, 1, fp) == 1 && fwrite(&mtime, 8, 1, fp) == 1 &&
              fwrite(&h->n, 4, 1, fp) == 1 && fwrite(&F, 4, 1, fp) == 1 && fwrite(h->b, 4, 4, fp) == 4;
    std::vector<uint32_t> ll;
    for (const std::string &l : h->labels) ll.push_back((uint32_t)l.size());
    if (ok && F) ok = fwrite(h->fcount.data(), 4, F, fp) == F && fwrite(ll.data(), 4, F, fp) == F;
    for (const std::string &l : h->labels)
        if (ok && !l.empty()) ok = fwrite(l.data(), 1, l.size(), fp) == l.size();
    if (ok && h->n) ok = fwrite(h->x.data(), 4, h->n, fp) == h->n && fwrite(h->y.data(), 4, h->n, fp) == h->n;
    ok = (fclose(fp) == 0) && ok;
    if (ok) ok = rename(tmp.c_str(), cache) == 0;
    if (!ok) remove(tmp.c_str());
}

// populateData (cpp:139-254) on T threads.  Line 1 is N; the rows run to the
// first empty line; a feature starts wherever the label differs from the
// previous row's; the line after the empty one holds the bounds.  Threads
// take byte ranges starting at line starts: pass 1 counts each range's rows
// up to its first empty line, pass 2 parses the rows at their global index,
// and the feature starts are merged in row order.  Any malformed row, a row
// count other than N or a missing bounds line is COLO_EPARSE, as in a
// sequential parse.
static int parse_csv(std::string &buf, uint32_t threads, colo_host_data *h) {
    char *const base = &buf[0];
    char *const end = base + buf.size() - 1;  // buf ends with an added '\0'
    char *p = base;
    char *nl = (char *)memchr(p, '\n', (size_t)(end - p));
    if (!nl) return COLO_EPARSE;
    *nl = 0;
    long nval;
    if (!parse_long(p, &nval) || nval < 0 || nval > 0xFFFFFFFFL) return COLO_EPARSE;
    h->n = (uint32_t)nval;
    char *body = nl + 1;
    // default: one range per MiB up to the hardware threads; an explicit
    // thread count only keeps ranges of at least 16 bytes
    const uint64_t bytes = (uint64_t)(end - body);
    uint64_t T = threads ? std::min<uint64_t>(threads, bytes / 16 + 1)
                         : std::min<uint64_t>(std::max(1u, std::thread::hardware_concurrency()), bytes / (1 << 20) + 1);
    T = std::max<uint64_t>(1, std::min<uint64_t>(T, 64));
    std::vector<char *> cs(T + 1);
    cs[0] = body;
    cs[T] = end;
    for (uint64_t t = 1; t < T; t++) {
        char *q = body + (uint64_t)(end - body) * t / T;
        if (q < cs[t - 1]) q = cs[t - 1];
        char *e = q > body ? (char *)memchr(q - 1, '\n', (size_t)(end - (q - 1))) : q - 1;
        cs[t] = e ? e + 1 : end;  // start of the first line beginning at or after q
        if (q == body) cs[t] = body;
    }
    // pass 1: rows of each range up to its first empty line
    std::vector<uint64_t> rows(T, 0);
    std::vector<char *> empty_at(T, nullptr);
    auto pass1 = [&](uint64_t t) {
        char *q = cs[t];
        while (q < cs[t + 1]) {
            char *e = (char *)memchr(q, '\n', (size_t)(end - q));
            char *le = e ? e : end;
            if (le == q) {
                empty_at[t] = q;
                return;
            }
            rows[t]++;
            q = e ? e + 1 : end;
        }
    };
    {
        std::vector<std::thread> th;
        for (uint64_t t = 1; t < T; t++) th.emplace_back(pass1, t);
        pass1(0);
        for (auto &x : th) x.join();
    }
    uint64_t tstop = T, nrows = 0;
    for (uint64_t t = 0; t < T; t++) {
        nrows += rows[t];
        if (empty_at[t]) {
            tstop = t;
            break;
        }
    }
    if (tstop == T || nrows != h->n) return COLO_EPARSE;  // no empty line, or not N rows
    h->x.assign(h->n, 0.f);
    h->y.assign(h->n, 0.f);
    std::vector<long> lab(h->n);
    // pass 2: parse; keep the label text where a range starts or the label changes
    std::vector<std::vector<std::pair<uint64_t, std::string>>> starts(tstop + 1);
    std::vector<int> bad(tstop + 1, 0);
    std::vector<uint64_t> row0(tstop + 1, 0);
    for (uint64_t t = 1; t <= tstop; t++) row0[t] = row0[t - 1] + rows[t - 1];
    auto pass2 = [&](uint64_t t) {
        char *q = cs[t];
        uint64_t r = row0[t];
        for (uint64_t i = 0; i < rows[t]; i++) {
            char *e = (char *)memchr(q, '\n', (size_t)(end - q));
            if (e) *e = 0;
            char *fld[3];
            long lb;
            if (split(q, fld, 3) < 3 || !parse_long(fld[0], &lb) || !parse_flt(fld[1], &h->x[r]) ||
                !parse_flt(fld[2], &h->y[r])) {
                bad[t] = 1;
                return;
            }
            lab[r] = lb;
            if (i == 0 || lb != lab[r - 1]) starts[t].emplace_back(r, std::string(fld[0]));
            r++;
            q = e ? e + 1 : end;
        }
    };
    {
        std::vector<std::thread> th;
        for (uint64_t t = 1; t <= tstop; t++) th.emplace_back(pass2, t);
        pass2(0);
        for (auto &x : th) x.join();
    }
    for (int b : bad)
        if (b) return COLO_EPARSE;
    // features in row order (cpp:168-183)
    h->fcount.clear();
    h->labels.clear();
    uint64_t prev_start = 0;
    for (uint64_t t = 0; t <= tstop; t++)
        for (auto &st : starts[t]) {
            const uint64_t r = st.first;
            if (r != 0 && lab[r] == lab[r - 1]) continue;  // a range boundary, not a feature start
            if (!h->fcount.empty()) h->fcount.back() = (uint32_t)(r - prev_start);
            h->fcount.push_back(0);
            h->labels.push_back(std::move(st.second));
            prev_start = r;
        }
    if (!h->fcount.empty()) h->fcount.back() = (uint32_t)(h->n - prev_start);
    // bounds: the line after the first empty one
    char *q = empty_at[tstop] + 1;
    if (q > end) return COLO_EPARSE;
    char *e = (char *)memchr(q, '\n', (size_t)(end - q));
    if (e) *e = 0;
    char *fld[4];
    if (split(q, fld, 4) < 4) return COLO_EPARSE;
    for (int i = 0; i < 4; i++)
        if (!parse_flt(fld[i], &h->b[i])) return COLO_EPARSE;
    return COLO_OK;
}

int colo_load_csv_opts(const char *path, const char *cache_path, uint32_t threads, colo_host_data **out) {
    if (!path || !out) return COLO_EARG;
    *out = nullptr;
    auto h = std::make_unique<colo_host_data>();
    uint64_t size = 0;
    int64_t mtime = 0;
    const bool stamped = source_stamp(path, &size, &mtime);
    if (!(cache_path && stamped && cache_read(cache_path, size, mtime, h.get()))) {
        FILE *fp = fopen(path, "rb");
        if (!fp) return COLO_EFILE;
        fseek(fp, 0, SEEK_END);
        long sz = ftell(fp);
        fseek(fp, 0, SEEK_SET);
        std::string buf((size_t)(sz > 0 ? sz : 0), '\0');
        size_t got = sz > 0 ? fread(&buf[0], 1, (size_t)sz, fp) : 0;
        fclose(fp);
        buf.resize(got);
        buf.push_back('\0');
        h = std::make_unique<colo_host_data>();
        int rc = parse_csv(buf, threads, h.get());
        if (rc) return rc;
        if (cache_path && stamped) cache_write(cache_path, size, mtime, h.get());
    }
    for (auto &l : h->labels) h->label_ptrs.push_back(l.c_str());
    *out = h.release();
    return COLO_OK;
}

int colo_load_csv(const char *path, colo_host_data **out) { return colo_load_csv_opts(path, nullptr, 0, out); }

int colo_host_dataset(const colo_host_data *h, colo_dataset *ds) {
    if (!h || !ds) return COLO_EARG;
    ds->n = h->n;
    ds->F = (uint32_t)h->fcount.size();
    ds->x = h->x.data();
    ds->y = h->y.data();
    ds->fcount = h->fcount.data();
    ds->minX = h->b[0]; ds->minY = h->b[1]; ds->maxX = h->b[2]; ds->maxY = h->b[3];
    ds->labels = h->label_ptrs.data();
    return COLO_OK;
}

const char *colo_host_label(const colo_host_data *h, uint32_t f) {
    return (h && f < h->labels.size()) ? h->labels[f].c_str() : "";
}

void colo_host_free(colo_host_data *h) { delete h; }

}  // extern "C"
