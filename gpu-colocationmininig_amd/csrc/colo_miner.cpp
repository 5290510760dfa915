// colo_miner.cpp -- host side of the miner: the colocationFinder role
// (Codes/colocationFinder.cpp) on top of the device layer.
//
//  - input: config.txt (cpp:1077-1093), parameter.txt (cpp:57-74), the
//    dataset CSV (cpp:139-254);
//  - the Apriori level loop (cpp:1104-1137), candidate generation with the
//    reference's first-failure break (cpp:517-557, 598-701);
//  - per level: filter, count pass, PI, prevalence, instance tables for the
//    prefixes of the next level's candidates (cpp:256-425, 736-909);
//  - output: ColocationDegree_<k>.txt (cpp:960-997) and a report.
//
// Two schedules share the kernels: PLAIN launches and synchronises per
// candidate (the reference's host loops cpp:315-348, 780-842, 860-908);
// PERSISTENT runs each phase of a level as one launch over a device work
// queue of (candidate, anchor range) items.
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>

#include "colo_host.h"
#include "colo_internal.h"

using namespace colo;
using colo_host::PatSet;
using colo_host::gen_candidates;
using colo_host::parse_flt;

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

struct Str {
    std::string s;
    void printf(const char *fmt, ...) __attribute__((format(printf, 2, 3))) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        int k = vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        if (k < (int)sizeof buf) {
            s.append(buf, k > 0 ? k : 0);
            return;
        }
        std::vector<char> big(k + 1);
        va_start(ap, fmt);
        vsnprintf(big.data(), big.size(), fmt, ap);
        va_end(ap);
        s.append(big.data(), k);
    }
};

uint32_t fbits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

// Device buffer from the context's caching pool (or plain hipMalloc when
// made without one).  Users synchronise before a buffer is released.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevPool *pool = nullptr;
    DevBuf() = default;
    explicit DevBuf(DevPool *pl) : pool(pl) {}
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            if (pool) pool->free(p, bytes);
            else (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
    }
    bool ensure(size_t b) {
        if (b <= bytes && p) return true;
        release();
        if (pool) {
            p = pool->alloc(b, &bytes);
            return p != nullptr;
        }
        if (hipMalloc(&p, b ? b : 8) != hipSuccess) {
            p = nullptr;
            return false;
        }
        bytes = b ? b : 8;
        return true;
    }
    template <typename T>
    T *as() const { return (T *)p; }
};

// A prevalent pattern of the current level.
struct Prev {
    std::vector<uint32_t> pat;
    std::shared_ptr<DevBuf> slots;  // count-pass slots (until its table is built)
    uint64_t slots_off = 0;         // offset inside a shared slots arena
    uint64_t nslots = 0;
    uint64_t rows = 0;
    int64_t parent = -1;            // index into the previous level's list
    std::shared_ptr<DevBuf> table;  // memory holding the instance table (own buffer or a level arena)
    uint32_t *tptr = nullptr;       // instance table, rows x k
    std::shared_ptr<DevBuf> offs;   // row offsets of the table (nslots + 1 entries from offs_at)
    std::shared_ptr<DevBuf> nbr;    // level 2: memory of the neighbour list
    uint32_t *nptr = nullptr;       // level 2: second column as sorted positions
    uint64_t offs_at = 0;           // index of the first offset inside offs
    uint64_t off_base = 0;          // value of the first offset
    // exact participation (subset bound): the members' windows of one bit
    // per instance in sorted-rank order, back to back (seg_layout), from
    // word part_off of part
    std::shared_ptr<DevBuf> part;
    uint64_t part_off = 0;
};

// Results that do not depend on theta, kept across the runs of a threshold
// sweep (ascending theta): every candidate's counted rows, numerators and
// PIs, and the instance tables built (a table's rows do not depend on
// theta; the prevalent pairs at a higher theta are a subset of those at a
// lower one).
struct SweepMemo {
    struct Res {
        uint64_t rows;
        uint64_t numc[kMaxK], nume[kMaxK];
        float pic, pie;
    };
    std::map<std::vector<uint32_t>, Res> res;
    std::map<std::pair<uint32_t, uint32_t>, Prev> l2;
    // instance tables of size 3 and more built so far (their rows do not
    // depend on theta): memory and device pointer
    std::map<std::vector<uint32_t>, std::pair<std::shared_ptr<DevBuf>, uint32_t *>> tk;
    std::map<std::vector<uint32_t>, std::vector<uint32_t>> fcnt;  // filter counts (theta-independent)
    // subset bounds and the participation of every prevalent pattern so far
    // (both theta-independent): a candidate first met at a higher theta
    // finds its subsets' participation here when they were remembered
    // instead of counted
    std::map<std::vector<uint32_t>, float> q;
    std::map<std::vector<uint32_t>, std::pair<std::shared_ptr<DevBuf>, uint64_t>> part;
    uint64_t hits = 0, misses = 0;
};

struct Miner {
    colo_ctx *c;
    const colo_mine_opts &o;
    colo_mine_stats st{};
    hipStream_t s;
    Str rep;
    colo_ctx *ctx_for_err;

    SweepMemo *memo = nullptr;

    Miner(colo_ctx *c_, const colo_mine_opts &o_) : c(c_), o(o_), s(c_->stream), ctx_for_err(c_) {
        for (DevBuf *b : {&arena_slots, &arena_bits, &d_desc, &d_items, &d_segs, &d_segout, &d_queue, &d_idx, &d_rowsel,
                          &d_pipre})
            b->pool = &c->pool;
    }
    ~Miner() {
        (void)hipStreamSynchronize(s);
        for (auto &x : execs_) (void)hipGraphExecDestroy(x);
        for (auto &e : pool_) (void)hipEventDestroy(e);
    }

    // HIP-event timing of launches on the miner's stream: mark() records an
    // event, span() pairs two marks with a category, harvest() (after a
    // stream synchronisation) adds the elapsed times to the statistics.
    enum Cat { kCount, kWrite, kFilter, kPi, kL2Count, kL2Write, kScan };
    std::vector<hipEvent_t> pool_;
    size_t used_ = 0;
    struct Span { size_t a, b; int cat; uint32_t k; };
    std::vector<Span> spans_;
    size_t mark() {
        if (used_ == pool_.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return SIZE_MAX;
            pool_.push_back(e);
        }
        (void)hipEventRecord(pool_[used_], s);
        return used_++;
    }
    void span(size_t a, size_t b, int cat, uint32_t k = 0) {
        if (a != SIZE_MAX && b != SIZE_MAX) spans_.push_back(Span{a, b, cat, k});
    }
    void harvest() {
        for (const Span &sp : spans_) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, pool_[sp.a], pool_[sp.b]) != hipSuccess) {
                (void)hipGetLastError();  // an untimed span must not fail a later check
                continue;
            }
            if ((sp.cat == kCount || sp.cat == kL2Count) && sp.k < 16) st.level_count_kernel_ms[sp.k] += ms;
            switch (sp.cat) {
                case kCount: st.count_kernel_ms += ms; break;
                case kWrite: st.write_kernel_ms += ms; break;
                case kFilter: st.filter_kernel_ms += ms; break;
                case kPi: st.pi_kernel_ms += ms; break;
                case kL2Count: st.l2_count_kernel_ms += ms; st.count_kernel_ms += ms; break;
                case kL2Write: st.l2_write_kernel_ms += ms; st.write_kernel_ms += ms; break;
                case kScan: st.scan_kernel_ms += ms; break;
            }
        }
        spans_.clear();
        used_ = 0;
    }

    // Algorithmic bytes (each datum once): level 2 count reads both features'
    // coordinates (8 B each), writes 4 B of slots per anchor and one bit per
    // instance; level-k count reads the last feature's coordinates and
    // writes slots (when kept) and bits -- its parent rows (4(k-1) B per
    // row) are charged once per run of candidates sharing them, by the
    // caller; writes add the offsets (8 B) and the table (4k B per row).
    double count_bytes(uint32_t k, const uint32_t *cand, uint64_t n_anchor, bool slots_kept = true) const {
        double Nsum = 0;
        for (uint32_t j = 0; j < k; j++) Nsum += c->fcount[cand[j]];
        const double sl = slots_kept ? 4.0 * n_anchor : 0.0;
        if (k == 2) return 8.0 * Nsum + sl + Nsum / 8.0;
        return 8.0 * c->fcount[cand[k - 1]] + sl + Nsum / 8.0;
    }
    double write_bytes(uint32_t k, const uint32_t *cand, uint64_t n_anchor, uint64_t rows) const {
        if (k == 2) return 8.0 * (c->fcount[cand[0]] + c->fcount[cand[1]]) + 12.0 * n_anchor + 8.0 * rows;
        return 4.0 * (k - 1) * n_anchor + 12.0 * n_anchor + 8.0 * c->fcount[cand[k - 1]] + 4.0 * k * rows;
    }

    // Out-of-memory error naming the buffer, the request and the device's
    // free memory.
    uint32_t cur_k_ = 0;
    int oom(const char *what, double bytes) {
        size_t fr = 0, tot = 0;
        (void)hipMemGetInfo(&fr, &tot);
        (void)hipGetLastError();
        char b[256];
        snprintf(b, sizeof b, "%s: %.3f GB requested at level %u; device free %.3f of %.3f GB, pool cache %.3f GB", what,
                 bytes / 1e9, cur_k_, fr / 1e9, tot / 1e9, c->pool.cached / 1e9);
        return set_error(c, COLO_ENOMEM, b);
    }

    // words per row of a batched level-k table in its level arena: k
    // positions, plus the neighbour list at level 2
    static uint64_t table_words(uint32_t k) { return k + (k == 2 ? 1 : 0); }

    int sync() {
        COLO_HIP(hipStreamSynchronize(s));
        for (auto &x : execs_) (void)hipGraphExecDestroy(x);
        execs_.clear();
        harvest();
        return COLO_OK;
    }

    // Each pass of a level (count, PI popcounts, offsets scan, table write)
    // is enqueued as a captured HIP graph when the context's graphs are on.
    // `launches` may only launch kernels (no host waits, allocations, host
    // copies or timing marks: the uploads come before, the marks around it;
    // an event recorded inside a capture cannot be timed by the host here).
    // Tuning runs stay eager; a stream that cannot be captured runs the
    // launches eagerly from then on.  Executables are destroyed at the next
    // synchronisation.
    std::vector<hipGraphExec_t> execs_;
    int as_graph(const std::function<int()> &launches) {
        if (!c->use_graphs || c->level_graphs_off || tuning_) return launches();
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            c->level_graphs_off = true;
            return launches();
        }
        const int rc = launches();
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        hipGraphExec_t x = nullptr;
        if (rc == COLO_OK && e == hipSuccess && g && hipGraphInstantiate(&x, g, nullptr, nullptr, 0) == hipSuccess) {
            (void)hipGraphDestroy(g);
            execs_.push_back(x);
            COLO_HIP(hipGraphLaunch(x, s));
            st.graphs++;
            return COLO_OK;
        }
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        c->level_graphs_off = true;  // capture unavailable here: eager from now on
        return launches();
    }

    void write_patterns(uint32_t k, const std::vector<Prev> &L) {
        if (!o.out_prefix) return;
        std::string path = std::string(o.out_prefix) + "ColocationDegree_" + std::to_string(k) + ".txt";
        FILE *fp = fopen(path.c_str(), "wb");
        if (!fp) return;
        std::string buf;
        for (const Prev &p : L) {  // savetoFile2 / savetoFileGen, cpp:960-997
            buf += "( ";
            for (uint32_t j = 0; j < k; j++) {
                buf += c->labels[p.pat[j]];
                buf += " ";
                buf += j != k - 1 ? "| " : ")\n";
            }
        }
        fwrite(buf.data(), 1, buf.size(), fp);
        fclose(fp);
    }

    // getIndex (cpp:703-718) as a hash lookup: position of pattern v (w ids)
    // in the previous level's prevalent list, -1 if absent.  The index is
    // rebuilt whenever the run moves to the next level (index_prev).
    std::unordered_map<std::vector<uint32_t>, int64_t, PatSet::H> prev_index_{16, PatSet::H{0}};
    const std::vector<Prev> *prev_indexed_ = nullptr;
    void index_prev(const std::vector<Prev> &L) {
        prev_index_.clear();
        prev_index_.reserve(L.size() * 2 + 1);
        for (size_t i = 0; i < L.size(); i++) prev_index_.emplace(L[i].pat, (int64_t)i);
        prev_indexed_ = &L;
    }
    int64_t find(const std::vector<Prev> &L, const uint32_t *v, uint32_t w) const {
        if (prev_indexed_ != &L || prev_index_.size() != L.size()) {  // not the indexed list
            for (size_t i = 0; i < L.size(); i++)
                if (std::equal(v, v + w, L[i].pat.begin())) return (int64_t)i;
            return -1;
        }
        auto it = prev_index_.find(std::vector<uint32_t>(v, v + w));
        return it == prev_index_.end() ? -1 : it->second;
    }

    // Per-candidate result of the count pass.
    struct CandRes {
        bool pass = false;
        uint64_t rows = 0;
        uint64_t numc[kMaxK], nume[kMaxK];
        float pic = 1.f, pie = 1.f;
        bool prevalent = false;
    };

    int run_level_plain(uint32_t k, const std::vector<uint32_t> &C, const std::vector<uint8_t> &pass,
                        const std::vector<Prev> &Lprev, std::vector<CandRes> &res, std::vector<Prev> &L) {
        const uint64_t nC = C.size() / k;
        DevBuf bits(&c->pool);
        uint64_t max_words = 0;
        for (uint64_t i = 0; i < nC; i++)
            if (pass[i]) max_words = std::max(max_words, colo_bits_words(c, k, &C[i * k]));
        if (!bits.ensure(max_words * 4)) return oom("bits buffer", max_words * 4);
        for (uint64_t i = 0; i < nC; i++) {
            if (!pass[i]) continue;
            const uint32_t *cand = &C[i * k];
            CandRes &r = res[i];
            r.pass = true;
            auto slots = std::make_shared<DevBuf>(&c->pool);
            uint64_t nslots;
            int64_t parent = -1;
            auto t0 = Clock::now();
            int rc;
            if (k == 2) {
                nslots = c->fcount[cand[0]];
                if (!slots->ensure(nslots * 4)) return oom("slots", nslots * 4);
                rc = colo_join2_count(c, cand[0], cand[1], slots->as<uint32_t>(), bits.as<uint32_t>(), &r.rows, s);
            } else {
                parent = find(Lprev, cand, k - 1);
                if (parent < 0 || !Lprev[parent].tptr) return set_error(c, COLO_EARG, "missing parent table");
                nslots = Lprev[parent].rows;
                if (!slots->ensure(nslots * 4)) return oom("slots", nslots * 4);
                const Prev *nb = nbr_list(cand[0], cand[k - 1]);
                if (!nb) return set_error(c, COLO_EARG, "internal: missing neighbour list");
                rc = colo_joink_count_nbr(c, Lprev[parent].tptr, nslots, k, cand,
                                          nb->nptr, nb->offs->as<uint64_t>() + nb->offs_at,
                                          slots->as<uint32_t>(), bits.as<uint32_t>(), &r.rows, s);
            }
            if (rc) return rc;
            st.count_ms += ms_since(t0);
            st.launches += 2;
            auto t1 = Clock::now();
            rc = colo_pi(c, bits.as<uint32_t>(), k, cand, r.numc, r.nume, &r.pic, &r.pie, s);
            if (rc) return rc;
            st.pi_ms += ms_since(t1);
            st.launches += 1;
            if (k == 2) st.l2_pairs += r.rows;
            float pim = o.pi_mode ? r.pie : r.pic;
            r.prevalent = !(pim < o.theta);  // cpp:335, 827
            if (r.prevalent) {
                Prev p;
                p.pat.assign(cand, cand + k);
                p.slots = slots;
                p.nslots = nslots;
                p.rows = r.rows;
                p.parent = parent;
                if (keep_parts_ && !force_slots_) {
                    // the count pass indexes members 0..k-2 by input rank
                    // and the last by sorted rank; the kept copy is all
                    // sorted rank
                    uint64_t so[kMaxK], words;
                    seg_layout(c, k, cand, 0, so, &words);
                    p.part = std::make_shared<DevBuf>(&c->pool);
                    if (!p.part->ensure(words * 4)) return oom("participation", words * 4);
                    if ((rc = fill_u32(p.part->as<uint32_t>(), words, 0, s))) return rc;
                    for (uint32_t j = 0; j < k && !rc; j++) {
                        if (j + 1 < k)
                            rc = rank_to_sorted(c, bits.as<uint32_t>() + so[j], cand[j], p.part->as<uint32_t>() + so[j], s);
                        else
                            COLO_HIP(hipMemcpyAsync(p.part->as<uint32_t>() + so[j], bits.as<uint32_t>() + so[j],
                                                    (c->fcount[cand[j]] + 31) / 32 * 4, hipMemcpyDeviceToDevice, s));
                    }
                    if (rc) return rc;
                    st.launches += k;
                }
                L.push_back(std::move(p));
            }
        }
        return COLO_OK;
    }

    int build_tables_plain(uint32_t k, std::vector<Prev> &L, const std::vector<uint8_t> &build,
                           const std::vector<Prev> &Lprev) {
        for (size_t i = 0; i < L.size(); i++) {
            if (!build[i]) continue;
            Prev &p = L[i];
            auto t0 = Clock::now();
            p.offs = std::make_shared<DevBuf>(&c->pool);
            if (!p.offs->ensure((p.nslots + 1) * 8)) return oom("offsets", (p.nslots + 1) * 8);
            p.offs_at = 0;
            p.off_base = 0;
            uint64_t total = 0;
            int rc = colo_scan_slots(c, p.slots->as<uint32_t>() + p.slots_off, p.nslots, p.offs->as<uint64_t>(), &total, s);
            if (rc) return rc;
            st.scan_ms += ms_since(t0);
            auto t1 = Clock::now();
            p.table = std::make_shared<DevBuf>(&c->pool);
            if (!p.table->ensure(std::max<uint64_t>(total * k, 1) * 4)) return oom("table", std::max<uint64_t>(total * k, 1) * 4);
            p.tptr = p.table->as<uint32_t>();
            if (k == 2) {
                p.nbr = std::make_shared<DevBuf>(&c->pool);
                if (!p.nbr->ensure(std::max<uint64_t>(total, 1) * 4)) return oom("neighbour list", std::max<uint64_t>(total, 1) * 4);
                p.nptr = p.nbr->as<uint32_t>();
                rc = colo_join2_write(c, p.pat[0], p.pat[1], p.slots->as<uint32_t>() + p.slots_off, p.offs->as<uint64_t>(),
                                      p.tptr, p.nptr, s);
            } else {
                const Prev *nb = nbr_list(p.pat[0], p.pat[k - 1]);
                if (!nb) return set_error(c, COLO_EARG, "internal: missing neighbour list");
                rc = colo_joink_write_nbr(c, Lprev[p.parent].tptr, p.nslots, k, p.pat.data(),
                                          nb->nptr, nb->offs->as<uint64_t>() + nb->offs_at,
                                          p.slots->as<uint32_t>() + p.slots_off, p.offs->as<uint64_t>(),
                                          p.tptr, s);
            }
            if (rc) return rc;
            if ((rc = sync())) return rc;
            st.write_ms += ms_since(t1);
            st.launches += 2;
            if (k == 2) st.l2_rows += total;
        }
        return COLO_OK;
    }

    // Level-2 patterns with tables, kept for the whole run as the anchors'
    // neighbour lists of the level-k passes.
    std::vector<Prev> l2_;
    std::vector<int64_t> l2_index_;  // F x F -> index into l2_, -1 if none
    const Prev *nbr_list(uint32_t a, uint32_t b) const {
        if (l2_index_.empty()) return nullptr;
        int64_t i = l2_index_[(size_t)a * c->F + b];
        return (i >= 0 && l2_[i].nptr && l2_[i].offs) ? &l2_[i] : nullptr;
    }

    // PERSISTENT: one count launch, one PI launch, one scan for the level.
    std::shared_ptr<DevBuf> arena_offs;
    DevBuf arena_slots, arena_bits, d_desc, d_items, d_segs, d_segout, d_queue, d_idx, d_rowsel, d_pipre;

    // Write items of chunk offsets start at multiples of 64 rows.
    static uint64_t write_chunk(uint64_t rows, bool chunked) { return chunked ? (rows + 63) / 64 * 64 : rows; }
    static void make_items(uint64_t cand, uint64_t n, uint64_t chunk, std::vector<WorkItem> &items, uint32_t ncand = 1) {
        for (uint64_t b = 0; b < n; b += chunk) items.push_back(WorkItem{(uint32_t)cand, ncand, b, std::min(n, b + chunk)});
    }

    // 8 contiguous item segments of about equal weight (rows x candidates,
    // or predicted cost), one per XCD, appended to the item upload.
    static std::vector<uint64_t> xcd_segments(const std::vector<WorkItem> &items, const std::vector<double> *w) {
        std::vector<uint64_t> seg(9, items.size());
        double tot = 0;
        std::vector<double> wt(items.size());
        for (size_t i = 0; i < items.size(); i++) {
            wt[i] = w ? (*w)[i] : (double)(items[i].end - items[i].begin) * items[i].ncand;
            tot += wt[i];
        }
        seg[0] = 0;
        double acc = 0;
        size_t x = 1;
        for (size_t i = 0; i < items.size() && x < 8; i++) {
            acc += wt[i];
            while (x < 8 && acc >= tot * x / 8.0) seg[x++] = i + 1;
        }
        for (; x < 8; x++) seg[x] = items.size();
        seg[8] = items.size();
        return seg;
    }
    int upload_items(const std::vector<WorkItem> &items, const std::vector<double> *w) {
        std::vector<uint64_t> seg = xcd_segments(items, w);
        size_t ib = items.size() * sizeof(WorkItem);
        if (!ensure_buf(d_items, ib + 9 * 8)) return oom("work items", ib + 9 * 8);
        COLO_HIP(hipMemcpyAsync(d_items.p, items.data(), ib, hipMemcpyHostToDevice, s));
        COLO_HIP(hipMemcpyAsync((char *)d_items.p + ib, seg.data(), 9 * 8, hipMemcpyHostToDevice, s));
        d_seg_ = (const uint64_t *)((const char *)d_items.p + ib);
        return COLO_OK;
    }
    const uint64_t *d_seg_ = nullptr;

    // Runs of consecutive descriptors that share a parent table (level k);
    // level 2 descriptors are runs of one.
    static std::vector<std::pair<uint64_t, uint32_t>> parent_runs(const std::vector<CandDesc> &desc, uint32_t k) {
        std::vector<std::pair<uint64_t, uint32_t>> runs;
        for (uint64_t u = 0; u < desc.size();) {
            uint64_t v = u + 1;
            if (k > 2)
                while (v < desc.size() && desc[v].parent == desc[u].parent && v - u < 64) v++;
            runs.emplace_back(u, (uint32_t)(v - u));
            u = v;
        }
        return runs;
    }

    int run_level_persistent(uint32_t k, const std::vector<uint32_t> &C, const std::vector<uint8_t> &pass,
                             const std::vector<Prev> &Lprev, std::vector<CandRes> &res, std::vector<Prev> &L,
                             std::shared_ptr<DevBuf> &slots_arena, const std::vector<float> *pred) {
        const uint64_t nC = C.size() / k;
        std::vector<uint64_t> idx;  // passing candidates
        for (uint64_t i = 0; i < nC; i++)
            if (pass[i]) idx.push_back(i);
        const uint64_t m = idx.size();
        if (!m) return COLO_OK;
        std::vector<CandDesc> desc(m);
        std::vector<int64_t> parent(m, -1);
        std::vector<uint64_t> cwords(m);  // bitmap words of each candidate
        uint64_t slots_tot = 0, bits_tot = 0;
        for (uint64_t u = 0; u < m; u++) {
            const uint32_t *cand = &C[idx[u] * k];
            CandDesc &d = desc[u];
            memset(&d, 0, sizeof d);
            d.k = k;
            for (uint32_t j = 0; j < k; j++) d.f[j] = cand[j];
            if (k == 2) {
                d.n_anchor = c->fcount[cand[0]];
            } else {
                parent[u] = find(Lprev, cand, k - 1);
                if (parent[u] < 0 || !Lprev[parent[u]].tptr) return set_error(c, COLO_EARG, "missing parent table");
                d.n_anchor = Lprev[parent[u]].rows;
                d.parent = Lprev[parent[u]].tptr;
                const Prev *nb = nbr_list(cand[0], cand[k - 1]);
                if (!nb) return set_error(c, COLO_EARG, "internal: missing neighbour list");
                d.nb_table = nb->nptr;
                d.nb_off = nb->offs->as<uint64_t>() + nb->offs_at;
                d.nb_base = nb->off_base;
            }
            d.slots_off = slots_tot;
            slots_tot += d.n_anchor;
            uint64_t words;
            seg_layout(c, k, cand, 0, d.seg_off, &words);
            d.bits_off = bits_tot;
            bits_tot += words;
            cwords[u] = words;
        }
        // The slots arena holds one u32 per parent row and candidate, but
        // only the tables built afterwards read slots.  Above the budget some
        // candidates keep none (their rows and bits are still counted); a
        // table built for one of them is recounted with slots
        // (recount_missing_slots).  Slots go first to the candidates most
        // likely to be prevalent -- the highest filter bound -- and among
        // equal bounds to the smallest.
        const uint64_t budget = force_slots_ ? ~0ull : slots_budget_words();
        if (slots_tot > budget) {
            std::vector<float> ub(m, 0.f);
            if (level_fcnt_ && level_fcnt_->size() >= C.size())
                for (uint64_t u = 0; u < m; u++) {
                    float b = 1.f;
                    for (uint32_t j = 0; j < k; j++) {
                        const float v = (float)((double)(*level_fcnt_)[idx[u] * k + j] /
                                                (double)(float)c->fcount[C[idx[u] * k + j]]);
                        b = j == 0 || v < b ? v : b;
                    }
                    ub[u] = b;
                }
            std::vector<uint64_t> ord(m);
            for (uint64_t u = 0; u < m; u++) ord[u] = u;
            std::stable_sort(ord.begin(), ord.end(), [&](uint64_t a, uint64_t b) {
                return ub[a] != ub[b] ? ub[a] > ub[b] : desc[a].n_anchor < desc[b].n_anchor;
            });
            std::vector<uint8_t> keep(m, 0);
            uint64_t kept = 0;
            for (uint64_t u : ord)
                if (kept + desc[u].n_anchor <= budget) {
                    kept += desc[u].n_anchor;
                    keep[u] = 1;
                }
            slots_tot = 0;
            for (uint64_t u = 0; u < m; u++) {
                if (keep[u]) {
                    desc[u].slots_off = slots_tot;
                    slots_tot += desc[u].n_anchor;
                } else {
                    desc[u].slots_off = kNoSlots;
                    st.slots_deferred++;
                }
            }
        }
        // work queue: candidates sharing a parent table form one item stream;
        // costliest runs first so the tail of the launch is short.  Without
        // a model a run costs its rows x candidates and is cut into
        // 1024-row items; with the model (learned schedule) a run costs its
        // parent rows plus the predicted refinement tests and table rows of
        // its candidates, and items are cut to an equal predicted cost.
        auto runs = parent_runs(desc, k);
        std::vector<double> cost(m, 0.0);
        for (auto &r : runs) {
            double cst = (double)desc[r.first].n_anchor * r.second;
            if (pred) {
                cst = (double)desc[r.first].n_anchor;
                for (uint32_t j = 0; j < r.second; j++) {
                    const float *y = &(*pred)[(r.first + j) * COLO_MLP_OUT];
                    cst += std::expm1((double)std::max(0.f, y[0])) + std::expm1((double)std::max(0.f, y[1]));
                }
            }
            cost[r.first] = cst;
        }
        std::stable_sort(runs.begin(), runs.end(), [&](const std::pair<uint64_t, uint32_t> &a,
                                                       const std::pair<uint64_t, uint32_t> &b) {
            return cost[a.first] > cost[b.first];
        });
        // work items of rows [begin, end) x one run, for an item size of
        // item_rows; with the model, cut at equal predicted cost
        std::vector<WorkItem> items;
        std::vector<double> iw;  // predicted cost of each item (learned schedule)
        auto build_items = [&](uint32_t item_rows) {
            items.clear();
            iw.clear();
            if (pred) {
                double total = 0;
                for (auto &r : runs) total += cost[r.first];
                // about 8 items per workgroup at equal predicted cost, each
                // between 1/4 and 8x the tuned item size
                double target = std::max(1.0, total / (double)(num_cus() * c->tune.blocks_per_cu * 8));
                const double lo = std::max(64.0, item_rows / 4.0), hi = item_rows * 8.0;
                for (auto &r : runs) {
                    uint64_t n = desc[r.first].n_anchor;
                    double per_row = n ? cost[r.first] / (double)n : 1.0;
                    uint64_t chunk = (uint64_t)std::min(hi, std::max(lo, target / std::max(per_row, 1e-9)));
                    size_t before = items.size();
                    make_items(r.first, n, chunk, items, r.second);
                    for (size_t i = before; i < items.size(); i++) iw.push_back(per_row * (items[i].end - items[i].begin));
                }
            } else {
                for (auto &r : runs) make_items(r.first, desc[r.first].n_anchor, item_rows, items, r.second);
            }
        };
        build_items(c->tune.item_rows_of(k));
        slots_arena = std::make_shared<DevBuf>(&c->pool);
        if (!ensure_buf(*slots_arena, std::max<uint64_t>(slots_tot, 1) * 4))
            return oom("slots arena", std::max<uint64_t>(slots_tot, 1) * 4);
        if (!ensure_buf(arena_bits, std::max<uint64_t>(bits_tot, 1) * 4))
            return oom("bits arena", std::max<uint64_t>(bits_tot, 1) * 4);
        if (!ensure_buf(d_desc, m * sizeof(CandDesc)) || !ensure_buf(d_queue, 64))
            return set_error(c, COLO_ENOMEM, "descriptors");
        COLO_HIP(hipMemcpyAsync(d_desc.p, desc.data(), m * sizeof(CandDesc), hipMemcpyHostToDevice, s));
        {
            int rc0 = upload_items(items, pred ? &iw : nullptr);
            if (rc0) return rc0;
        }
        if (!ensure_buf(d_rowsel, (m + 1) * 8)) return oom("row totals", (m + 1) * 8);
        if (tuning_ && !force_slots_) {
            int rc0 = tune_level(k, items, iw, pred != nullptr, build_items, slots_arena->as<uint32_t>(), bits_tot, m);
            if (rc0) return rc0;
            build_items(c->tune.item_rows_of(k));
            if ((rc0 = upload_items(items, pred ? &iw : nullptr))) return rc0;
        }
        auto t0 = Clock::now();
        // PI segments of every candidate, uploaded before the level's graph
        std::vector<SegDesc> segs(m * k);
        for (uint64_t u = 0; u < m; u++) seg_descs(c, k, &C[idx[u] * k], desc[u].bits_off, &segs[u * k], true);
        if (!ensure_buf(d_segs, segs.size() * sizeof(SegDesc)) || !ensure_buf(d_segout, segs.size() * sizeof(SegOut)) ||
            !ensure_buf(d_pipre, segs.size() * 8))
            return set_error(c, COLO_ENOMEM, "segments");
        std::vector<uint64_t> pi_prefix;  // read by the upload before the sync below
        const uint64_t pi_chunks = pi_chunk_prefix(segs.data(), segs.size(), pi_prefix);
        COLO_HIP(hipMemcpyAsync(d_segs.p, segs.data(), segs.size() * sizeof(SegDesc), hipMemcpyHostToDevice, s));
        COLO_HIP(hipMemcpyAsync(d_pipre.p, pi_prefix.data(), segs.size() * 8, hipMemcpyHostToDevice, s));
        // count pass, then the PI popcounts
        int rc = fill_u32(arena_bits.as<uint32_t>(), bits_tot, 0, s);
        if (!rc) rc = fill_u32(d_rowsel.as<uint32_t>(), 2 * m, 0, s);
        if (rc) return rc;
        size_t e0 = mark();
        rc = as_graph([&]() {
            return count_batch(c, k, d_desc.as<CandDesc>(), d_items.as<WorkItem>(), d_seg_, items.size(),
                               slots_arena->as<uint32_t>(), arena_bits.as<uint32_t>(), d_rowsel.as<uint64_t>(),
                               d_queue.as<uint32_t>(), s);
        });
        if (rc) return rc;
        span(e0, mark(), k == 2 ? kL2Count : kCount, k);
        size_t e1 = mark();
        rc = as_graph([&]() {
            return pi_batch_launch(c, arena_bits.as<uint32_t>(), d_segs.as<SegDesc>(), d_pipre.as<uint64_t>(),
                                   segs.size(), pi_chunks, d_segout.as<SegOut>(), s);
        });
        if (rc) return rc;
        span(e1, mark(), kPi);
        if (pred) {
            // pre-size the level's table arena from the predicted rows while
            // the count pass runs (the exact sizes are known only after it).
            // Every row is a passed test, so a prediction is capped by the
            // predicted tests and, for a pair, by its exact 3x3 join work.
            // A reservation is a guess: it takes at most half of the pool's
            // headroom, so that an over-prediction cannot push the cached
            // blocks of earlier levels out of the pool.
            double words = 0;
            for (uint64_t u = 0; u < m; u++) {
                const float *y = &(*pred)[u * COLO_MLP_OUT];
                double rows = std::expm1((double)std::max(0.f, std::min(y[0], y[1])));
                if (k == 2 && !c->pairwork.empty())
                    rows = std::min(rows, (double)c->pairwork[(size_t)desc[u].f[0] * c->F + desc[u].f[1]]);
                words += rows * table_words(k);
            }
            words *= 1.10;
            const double headroom = (double)c->pool.cap_bytes() - (double)c->pool.in_use;
            if (words * 4 <= headroom / 2) {
                arena_reserved_ = std::make_shared<DevBuf>(&c->pool);
                if (!arena_reserved_->ensure((size_t)std::max(words, 1.0) * 4)) arena_reserved_.reset();
                else st.arena_bytes += (double)arena_reserved_->bytes;
            }
        }
        {
            // a run's candidates read their shared parent rows once
            double by = 0;
            for (uint64_t u = 0; u < m; u++)
                by += count_bytes(k, &C[idx[u] * k], desc[u].n_anchor, desc[u].slots_off != kNoSlots);
            if (k > 2)
                for (auto &r : runs) by += 4.0 * (k - 1) * desc[r.first].n_anchor;
            st.count_bytes += by;
            if (k == 2) st.l2_count_bytes += by;
            st.count_launches++;
        }
        st.launches += 2;
        // rows per candidate, summed by the count pass (offsets are built
        // later, only for the tables that are written)
        std::vector<uint64_t> rows_of(m, 0);
        COLO_HIP(hipMemcpyAsync(rows_of.data(), d_rowsel.p, m * 8, hipMemcpyDeviceToHost, s));
        std::vector<SegOut> so(segs.size());
        COLO_HIP(hipMemcpyAsync(so.data(), d_segout.p, so.size() * sizeof(SegOut), hipMemcpyDeviceToHost, s));
        if ((rc = sync())) return rc;
        st.count_ms += ms_since(t0);
        st.launches += 4;
        for (uint64_t u = 0; u < m; u++) {
            const uint32_t *cand = &C[idx[u] * k];
            CandRes &r = res[idx[u]];
            r.pass = true;
            r.rows = rows_of[u];
            pi_from_segs(c, k, cand, &so[u * k], r.numc, r.nume, &r.pic, &r.pie);
            if (k == 2) st.l2_pairs += r.rows;
            float pim = o.pi_mode ? r.pie : r.pic;
            r.prevalent = !(pim < o.theta);
            if (r.prevalent) {
                Prev p;
                p.pat.assign(cand, cand + k);
                if (desc[u].slots_off != kNoSlots) p.slots = slots_arena;  // else recounted if its table is built
                p.slots_off = desc[u].slots_off;
                p.nslots = desc[u].n_anchor;
                p.rows = r.rows;
                p.parent = parent[u];
                L.push_back(std::move(p));
            }
        }
        if (keep_parts_ && !force_slots_) {
            // the prevalent candidates' bitmaps, copied out of the arena
            // (which the next pass reuses) for the next level's subset bound
            std::vector<CopyRange> cr;
            uint64_t tot = 0;
            for (uint64_t u = 0; u < m; u++)
                if (res[idx[u]].prevalent) {
                    cr.push_back(CopyRange{desc[u].bits_off, tot, cwords[u]});
                    tot += cwords[u];
                }
            if (!cr.empty()) {
                auto store = std::make_shared<DevBuf>(&c->pool);
                if (!store->ensure(tot * 4)) return oom("participation store", tot * 4);
                if ((rc = gather_ranges(c, arena_bits.as<uint32_t>(), store->as<uint32_t>(), cr.data(), cr.size(), s)))
                    return rc;
                st.launches += 1;
                const size_t j0 = L.size() - cr.size();
                for (size_t t = 0; t < cr.size(); t++) {
                    L[j0 + t].part = store;
                    L[j0 + t].part_off = cr[t].dst;
                }
            }
        }
        return COLO_OK;
    }

    // Subset bounds of the level's candidates that pass the filter
    // (k >= 3): q[i] = min over members j of fl32(cnt_j + t_j) / fl32(N_j),
    // cnt_j = |AND over the (k-1)-subsets S containing member j of its
    // participation in S|, t_j = 0 with exact numerators and, with the
    // reference's (cu:108-132), the last-instance bit of member j-1's AND
    // for j >= 1.  A candidate's participation in a member is a subset of
    // that in every such S, and a reference numerator exceeds the exact one
    // only by member j-1's last-instance bit, so both PIs are at most q.
    // Candidates not computed keep NaN.
    int subset_bounds(uint32_t k, const std::vector<uint32_t> &C, const std::vector<uint8_t> &pass,
                      const std::vector<Prev> &Lprev, std::vector<float> &q) {
        const uint64_t nC = C.size() / k;
        q.assign(nC, std::numeric_limits<float>::quiet_NaN());
        std::vector<BoundTask> tasks;
        std::vector<uint64_t> owner;
        uint32_t sub[kMaxK];
        for (uint64_t i = 0; i < nC; i++) {
            if (!pass[i]) continue;
            const uint32_t *cand = &C[i * k];
            if (memo) {
                auto it = memo->q.find(std::vector<uint32_t>(cand, cand + k));
                if (it != memo->q.end()) {
                    q[i] = it->second;
                    continue;
                }
            }
            // window of member j inside S_ip = the candidate without member ip
            const uint32_t *base[kMaxK];
            uint64_t so[kMaxK][kMaxK];
            for (uint32_t ip = 0; ip < k; ip++) {
                uint32_t w = 0;
                for (uint32_t t = 0; t < k; t++)
                    if (t != ip) sub[w++] = cand[t];
                const int64_t pi = find(Lprev, sub, k - 1);
                const DevBuf *buf = nullptr;
                uint64_t off = 0;
                if (pi >= 0 && Lprev[pi].part) {
                    buf = Lprev[pi].part.get();
                    off = Lprev[pi].part_off;
                } else if (memo) {
                    auto it = memo->part.find(std::vector<uint32_t>(sub, sub + k - 1));
                    if (it != memo->part.end()) {
                        buf = it->second.first.get();
                        off = it->second.second;
                    }
                }
                if (!buf) return set_error(c, COLO_EARG, "internal: subset participation missing");
                base[ip] = buf->as<uint32_t>();
                seg_layout(c, k - 1, sub, off, so[ip], nullptr);
            }
            for (uint32_t j = 0; j < k; j++) {
                BoundTask t{};
                for (uint32_t ip = 0; ip < k; ip++)
                    if (ip != j) t.w[t.nw++] = base[ip] + so[ip][j < ip ? j : j - 1];
                t.words = (c->fcount[cand[j]] + 31) / 32;
                t.last_bit = c->last_sorted[cand[j]];
                tasks.push_back(t);
            }
            owner.push_back(i);
        }
        std::vector<uint64_t> out(tasks.size() * 2);
        int rc = subset_bound_batch(c, tasks.data(), tasks.size(), out.data(), s);
        if (rc) return rc;
        if (!tasks.empty()) st.launches += 1;
        for (size_t g = 0; g < owner.size(); g++) {
            const uint32_t *cand = &C[owner[g] * k];
            float qq = 1.0f;
            for (uint32_t j = 0; j < k; j++) {
                const uint64_t v = out[2 * (g * k + j)] + (o.pi_mode == 0 && j >= 1 ? out[2 * (g * k + j - 1) + 1] : 0);
                volatile float b = (float)v / (float)c->fcount[cand[j]];
                if (b < qq) qq = b;
            }
            q[owner[g]] = qq;
            if (memo) memo->q.emplace(std::vector<uint32_t>(cand, cand + k), qq);
        }
        return COLO_OK;
    }


    int build_tables_persistent(uint32_t k, std::vector<Prev> &L, const std::vector<uint8_t> &build,
                                const std::vector<Prev> &Lprev, const std::vector<uint32_t> &C) {
        // the offsets arena still holds the level's global row offsets
        std::vector<CandDesc> desc;
        std::vector<WorkItem> items;
        auto t0 = Clock::now();
        (void)C;
        // one table arena for the level: tables (k words per row) and, at
        // level 2, the neighbour lists (1 word per row)
        uint64_t words = 0;
        for (size_t i = 0; i < L.size(); i++)
            if (build[i]) words += L[i].rows * table_words(k);
        // row offsets of the written tables: one exclusive scan per table
        // into one offsets arena (n_anchor + 1 entries each, the total
        // last).  Write variant 3 (levels >= 3) keeps one offset per 64
        // rows instead and rebuilds the rest from the slots; level 2 keeps
        // every row's, which serve the later levels as list bounds.
        const bool chunked = k > 2 && c->tune.write_var == 3;
        auto n_offs = [&](uint64_t n) { return chunked ? (n + 63) / 64 + 1 : n + 1; };
        uint64_t owords = 0;
        for (size_t i = 0; i < L.size(); i++)
            if (build[i]) owords += n_offs(L[i].nslots);
        arena_offs = std::make_shared<DevBuf>(&c->pool);
        if (!ensure_buf(*arena_offs, std::max<uint64_t>(owords, 1) * 8))
            return oom("offsets arena", std::max<uint64_t>(owords, 1) * 8);
        std::vector<uint64_t> offs_at(L.size(), 0);
        ScanPlan plan;  // the scan runs with the write, in the level's second graph
        {
            std::vector<ScanSeg> segs;
            uint64_t oat = 0;
            for (size_t i = 0; i < L.size(); i++) {
                if (!build[i]) continue;
                offs_at[i] = oat;
                segs.push_back(ScanSeg{L[i].slots->as<uint32_t>() + L[i].slots_off, arena_offs->as<uint64_t>() + oat,
                                       L[i].nslots, chunked});
                oat += n_offs(L[i].nslots);
            }
            int rc0 = scan_prepare(c, segs.data(), segs.size(), plan, s);
            if (rc0) return rc0;
        }
        // the offsets scan (before the tuning's timed writes, which need it)
        auto scan = [&]() -> int {
            size_t e0 = mark();
            int r = as_graph([&]() { return scan_launch(c, plan, s); });
            if (r) return r;
            span(e0, mark(), kScan, k);
            return COLO_OK;
        };
        std::shared_ptr<DevBuf> arena;
        if (arena_reserved_ && arena_reserved_->bytes >= words * 4) {
            arena = arena_reserved_;      // pre-sized before the count pass
        } else {
            if (arena_reserved_) st.overflows++;  // prediction too small: exact size
            arena_reserved_.reset();               // given back before the exact allocation
            arena = std::make_shared<DevBuf>(&c->pool);
            if (!arena->ensure(std::max<uint64_t>(words, 1) * 4)) return oom("table arena", std::max<uint64_t>(words, 1) * 4);
        }
        arena_reserved_.reset();
        st.arenas++;
        uint64_t at = 0;
        for (size_t i = 0; i < L.size(); i++) {
            if (!build[i]) continue;
            Prev &p = L[i];
            p.table = arena;
            p.tptr = arena->as<uint32_t>() + at;
            at += p.rows * k;
            CandDesc d;
            memset(&d, 0, sizeof d);
            d.k = k;
            for (uint32_t j = 0; j < k; j++) d.f[j] = p.pat[j];
            d.n_anchor = p.nslots;
            d.wslots = p.slots->as<uint32_t>() + p.slots_off;
            if (k > 2) {
                d.parent = Lprev[p.parent].tptr;
                const Prev *nb = nbr_list(p.pat[0], p.pat[k - 1]);
                if (!nb) return set_error(c, COLO_EARG, "internal: missing neighbour list");
                d.nb_table = nb->nptr;
                d.nb_off = nb->offs->as<uint64_t>() + nb->offs_at;
                d.nb_base = nb->off_base;
            }
            d.offsets = arena_offs->as<uint64_t>() + offs_at[i];
            d.off_base = 0;
            p.offs = arena_offs;
            p.offs_at = offs_at[i];
            p.off_base = 0;
            if (k == 2) {
                p.nbr = arena;
                p.nptr = arena->as<uint32_t>() + at;
                at += p.rows;
                d.nbr_out = p.nptr;
            }
            d.table = p.tptr;
            desc.push_back(d);
            if (k == 2) st.l2_rows += p.rows;
        }
        if (desc.empty()) return COLO_OK;
        if (!ensure_buf(d_desc, desc.size() * sizeof(CandDesc)) || !ensure_buf(d_queue, 64))
            return set_error(c, COLO_ENOMEM, "descriptors");
        COLO_HIP(hipMemcpyAsync(d_desc.p, desc.data(), desc.size() * sizeof(CandDesc), hipMemcpyHostToDevice, s));
        const auto runs = parent_runs(desc, k);
        if (tuning_ && k < 16) {
            // the level's item size of the write: each size writes the
            // whole tables (the same rows every time), the fastest is kept
            if (int r = scan()) return r;
            static const uint32_t sizes[] = {1024, 4096, 16384};
            hipEvent_t ev[2];
            COLO_HIP(hipEventCreate(&ev[0]));
            if (hipEventCreate(&ev[1]) != hipSuccess) {
                (void)hipEventDestroy(ev[0]);
                return set_error(c, COLO_EHIP, "event");
            }
            double best = 1e300;
            uint32_t bs = c->tune.write_rows_of(k);
            int rc0 = COLO_OK;
            for (uint32_t sz : sizes) {
                items.clear();
                for (auto &r : runs) make_items(r.first, desc[r.first].n_anchor, write_chunk(sz, chunked), items, r.second);
                if ((rc0 = upload_items(items, nullptr))) break;
                float ms = 0;
                if (hipEventRecord(ev[0], s) != hipSuccess) { rc0 = set_error(c, COLO_EHIP, "event"); break; }
                if ((rc0 = write_batch(c, k, d_desc.as<CandDesc>(), d_items.as<WorkItem>(), d_seg_, items.size(), nullptr,
                                       d_queue.as<uint32_t>(), s)))
                    break;
                if (hipEventRecord(ev[1], s) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
                    hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) {
                    rc0 = set_error(c, COLO_EHIP, "tuning launch");
                    break;
                }
                if (c->verbose)
                    fprintf(stderr, "[colo tune] level %u write, %u-row items: %.2f ms\n", k, sz, ms);
                if (ms < best) {
                    best = ms;
                    bs = sz;
                }
            }
            (void)hipEventDestroy(ev[0]);
            (void)hipEventDestroy(ev[1]);
            if (rc0) return rc0;
            c->tune.write_rows_k[k] = bs;
            items.clear();
        }
        for (auto &r : runs)
            make_items(r.first, desc[r.first].n_anchor, write_chunk(c->tune.write_rows_of(k), chunked), items, r.second);
        {
            int rc0 = upload_items(items, nullptr);
            if (rc0) return rc0;
        }
        // offsets scan (already run when tuning), then the table write
        int rc = COLO_OK;
        if (!(tuning_ && k < 16) && (rc = scan())) return rc;
        size_t e0 = mark();
        rc = as_graph([&]() {
            return write_batch(c, k, d_desc.as<CandDesc>(), d_items.as<WorkItem>(), d_seg_, items.size(), nullptr,
                               d_queue.as<uint32_t>(), s);
        });
        if (rc) return rc;
        span(e0, mark(), k == 2 ? kL2Write : kWrite);
        {
            double by = 0;
            for (size_t i = 0; i < L.size(); i++)
                if (build[i]) by += write_bytes(k, L[i].pat.data(), L[i].nslots, L[i].rows);
            st.write_bytes += by;
            if (k == 2) st.l2_write_bytes += by;
            st.write_launches++;
        }
        if ((rc = sync())) return rc;
        st.write_ms += ms_since(t0);
        st.launches += 2;
        (void)Lprev;
        return COLO_OK;
    }

    // Online tuning of a level's count pass (the autotuner's per-level
    // step): every (count variant, item size) pair is timed with HIP events
    // on a sample of the level's work items -- blocks of 16 consecutive
    // items, one block in 8, so neighbouring items still share their L2 --
    // and the pairs within 10% of the fastest sample (at most four) are
    // timed again on all of the level's items -- a sample of about one item
    // per workgroup measures its slowest item more than the variant's
    // throughput -- and the fastest of those is kept for the level in
    // c->tune.  The tuning launches write into the level's slots, bits and
    // row totals, which the real pass clears and rewrites: results do not
    // depend on the tuning.
    bool tuning_ = false;
    int tune_level(uint32_t k, std::vector<WorkItem> &items, std::vector<double> &iw, bool has_pred,
                   const std::function<void(uint32_t)> &build, uint32_t *d_slots, uint64_t bits_tot, uint64_t m) {
        if (k >= 16) return COLO_OK;
        static const uint32_t vars[] = {0, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13}, vars2[] = {0, 1},
                              sizes[] = {1024, 4096, 16384};
        const uint32_t nv = k == 2 ? 2 : 12;  // level 2: one anchor per lane, or packed anchors
        const uint32_t *vs = k == 2 ? vars2 : vars;
        hipEvent_t ev[2];
        COLO_HIP(hipEventCreate(&ev[0]));
        if (hipEventCreate(&ev[1]) != hipSuccess) {
            (void)hipEventDestroy(ev[0]);
            return set_error(c, COLO_EHIP, "event");
        }
        double best = 1e300;
        uint32_t bv = c->tune.count_var_of(k), bs = c->tune.item_rows_of(k);
        struct Sample {
            float ms;
            uint32_t var, rows;
        };
        std::vector<Sample> samples;
        int rc = COLO_OK;
        // one count launch over the uploaded items (n of them), in ms
        auto timed = [&](size_t n, float &ms) {
            int r = fill_u32(arena_bits.as<uint32_t>(), bits_tot, 0, s);
            if (!r) r = fill_u32(d_rowsel.as<uint32_t>(), 2 * m, 0, s);
            if (r) return r;
            if (hipEventRecord(ev[0], s) != hipSuccess) return set_error(c, COLO_EHIP, "event");
            r = count_batch(c, k, d_desc.as<CandDesc>(), d_items.as<WorkItem>(), d_seg_, n, d_slots,
                            arena_bits.as<uint32_t>(), d_rowsel.as<uint64_t>(), d_queue.as<uint32_t>(), s);
            if (r) return r;
            if (hipEventRecord(ev[1], s) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
                hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess)
                return set_error(c, COLO_EHIP, "tuning launch");
            return COLO_OK;
        };
        for (uint32_t si = 0; si < 3 && !rc; si++) {
            build(sizes[si]);
            std::vector<WorkItem> sub;
            std::vector<double> subw;
            for (size_t i = 0; i < items.size(); i++)
                if ((i / 16) % 8 == 0) {
                    sub.push_back(items[i]);
                    if (has_pred) subw.push_back(iw[i]);
                }
            if ((rc = upload_items(sub, has_pred ? &subw : nullptr))) break;
            for (uint32_t vi = 0; vi < nv && !rc; vi++) {
                c->tune.count_var_k[k] = vs[vi];
                // the faster of two timed launches (one launch of a sample
                // varies by several percent)
                float ms = 1e30f;
                for (int rep = 0; rep < 2 && !rc; rep++) {
                    float one = 0;
                    if (!(rc = timed(sub.size(), one))) ms = std::min(ms, one);
                }
                if (rc) break;
                if (c->verbose)
                    fprintf(stderr, "[colo tune] level %u count var %u, %u-row items: %.2f ms on %zu of %zu items\n", k,
                            vs[vi], sizes[si], ms, sub.size(), items.size());
                samples.push_back(Sample{ms, vs[vi], sizes[si]});
                if (ms < best) {
                    best = ms;
                    bv = vs[vi];
                    bs = sizes[si];
                }
            }
        }
        // the close samples again, on every item of the level
        std::stable_sort(samples.begin(), samples.end(), [](const Sample &a, const Sample &b) { return a.ms < b.ms; });
        size_t nclose = 0;
        while (nclose < samples.size() && nclose < 4 && samples[nclose].ms <= 1.1 * best) nclose++;
        if (!rc && nclose > 1) {
            float full_best = 1e30f;
            for (size_t i = 0; i < nclose && !rc; i++) {
                build(samples[i].rows);
                if ((rc = upload_items(items, has_pred ? &iw : nullptr))) break;
                c->tune.count_var_k[k] = samples[i].var;
                float ms = 0;
                if ((rc = timed(items.size(), ms))) break;
                if (c->verbose)
                    fprintf(stderr, "[colo tune] level %u count var %u, %u-row items: %.2f ms on all %zu items\n", k,
                            samples[i].var, samples[i].rows, ms, items.size());
                if (ms < full_best) {
                    full_best = ms;
                    bv = samples[i].var;
                    bs = samples[i].rows;
                }
            }
        }
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        if (rc) return rc;
        c->tune.count_var_k[k] = bv;
        c->tune.item_rows_k[k] = bs;
        return COLO_OK;
    }

    // Recounts, with slots, the candidates whose tables are to be built but
    // whose count pass kept no slots (deferred above the slots budget, or
    // results remembered from an earlier threshold of a sweep).  The
    // recount's results equal the first count's; only its slots are used.
    int recount_missing_slots(uint32_t k, const std::vector<uint32_t> &C, std::vector<Prev> &L,
                              const std::vector<uint8_t> &build, const std::vector<Prev> &Lprev) {
        const uint64_t nC = C.size() / k;
        std::vector<uint8_t> again(nC, 0);
        bool any = false;
        for (size_t i = 0; i < L.size() && !any; i++) any = build[i] && !L[i].slots;
        if (!any) return COLO_OK;
        std::unordered_map<std::vector<uint32_t>, uint64_t, PatSet::H> at(nC * 2 + 1, PatSet::H{k});
        for (uint64_t ci = 0; ci < nC; ci++) at.emplace(std::vector<uint32_t>(&C[ci * k], &C[ci * k] + k), ci);
        for (size_t i = 0; i < L.size(); i++)
            if (build[i] && !L[i].slots) {
                auto it = at.find(L[i].pat);
                if (it == at.end()) return set_error(c, COLO_EARG, "internal: table without a candidate");
                again[it->second] = 1;
                st.recounts++;
            }
        std::vector<CandRes> res2(nC);
        std::vector<Prev> L2;
        std::shared_ptr<DevBuf> arena2;
        force_slots_ = true;
        int rc = o.sched == COLO_SCHED_PLAIN ? run_level_plain(k, C, again, Lprev, res2, L2)
                                             : run_level_persistent(k, C, again, Lprev, res2, L2, arena2, nullptr);
        force_slots_ = false;
        if (rc) return rc;
        std::unordered_map<std::vector<uint32_t>, size_t, PatSet::H> li(L.size() * 2 + 1, PatSet::H{k});
        for (size_t i = 0; i < L.size(); i++) li.emplace(L[i].pat, i);
        for (Prev &q : L2) {
            auto it = li.find(q.pat);
            if (it == li.end()) continue;
            Prev &p = L[it->second];
            p.slots = q.slots;
            p.slots_off = q.slots_off;
            p.nslots = q.nslots;
            p.parent = q.parent;
        }
        for (size_t i = 0; i < L.size(); i++)
            if (build[i] && !L[i].slots) return set_error(c, COLO_EARG, "internal: recount kept no slots");
        return COLO_OK;
    }

    // ---------------------------------------------------- cost model
    // The learned schedule's table arena, reserved from the predicted rows
    // while a level's count pass runs; taken by build_tables_persistent when
    // large enough.
    std::shared_ptr<DevBuf> arena_reserved_;
    // Allocation that gives the reservation back before it reports a
    // failure: the reservation only saves an allocation, so it must never
    // make a run fail that the exact sizes fit.
    // Slots arena budget of a level's count pass in u32 words: 1/16 of the
    // device's memory, or the tuned limit (KiB) when one is set.
    bool force_slots_ = false;
    // Per-row slots kept by a count pass: 1/16 of device memory.  A larger
    // budget spares recounts (c5 sweep 763 -> 742 ms with all 70 GB of
    // level-3 slots kept) but leaves the pool too little room to keep its
    // blocks between runs: a budget sized from the pool's headroom made
    // repeated runs free and reallocate arenas (c5 persistent 0.76 -> 6.1 s).
    const std::vector<uint32_t> *level_fcnt_ = nullptr;  // the level's filter counts (C x k), while it runs
    bool keep_parts_ = false;  // keep the prevalent patterns' participation (a next level's subset bound)
    uint64_t slots_budget_words() {
        if (c->tune.slots_budget_kb) return (uint64_t)c->tune.slots_budget_kb << 8;
        return (uint64_t)(c->pool.total_bytes() / 16 / 4);
    }
    bool ensure_buf(DevBuf &b, size_t bytes) {
        if (b.ensure(bytes)) return true;
        if (!arena_reserved_) return false;
        arena_reserved_.reset();
        return b.ensure(bytes);
    }

    // Features of a candidate (COLO_MLP_IN floats): size, instance counts,
    // anchors / parent rows, filter counts and bound, non-empty cells, the
    // 3x3 work of (first, last), level-2 rows of (first, last) and
    // (second-to-last, last), mean cell occupancy of first and last.
    void cand_features(uint32_t k, const uint32_t *cand, const uint32_t *fcnt, uint64_t n_anchor, uint64_t pairwork,
                       float *f) const {
        auto L1 = [](double v) { return (float)std::log1p(v > 0 ? v : 0); };
        uint32_t a = cand[0], z = cand[k - 1];
        double Nmin = 1e30, Nmax = 0, ub = 1e30;
        for (uint32_t j = 0; j < k; j++) {
            double N = c->fcount[cand[j]];
            Nmin = std::min(Nmin, N);
            Nmax = std::max(Nmax, N);
            double u = fcnt ? (float)((double)fcnt[j] / (double)(float)N) : 1.0;
            ub = std::min(ub, u);
        }
        auto H = [&](uint32_t ff) { return (double)(c->run_off[ff + 1] - c->run_off[ff]); };
        auto R2 = [&](uint32_t x, uint32_t y) -> double {
            if (l2_index_.empty()) return 0;
            int64_t i = l2_index_[(size_t)x * c->F + y];
            return i >= 0 ? (double)l2_[i].rows : 0;
        };
        f[0] = (float)k;
        f[1] = L1(c->fcount[a]);
        f[2] = L1(c->fcount[z]);
        f[3] = L1(Nmin);
        f[4] = L1(Nmax);
        f[5] = L1((double)n_anchor);
        f[6] = L1(fcnt ? fcnt[0] : 0);
        f[7] = L1(fcnt ? fcnt[k - 1] : 0);
        f[8] = (float)ub;
        f[9] = L1(H(a));
        f[10] = L1(H(z));
        f[11] = L1((double)pairwork);
        f[12] = k > 2 ? L1(R2(a, z)) : 0.f;
        f[13] = k > 2 ? L1(R2(cand[k - 2], z)) : 0.f;
        f[14] = (float)std::log(c->fcount[a] / std::max(1.0, H(a)));
        f[15] = (float)std::log(c->fcount[z] / std::max(1.0, H(z)));
    }

    // Features of the level's passing candidates (rows of COLO_MLP_IN).
    int level_features(uint32_t k, const std::vector<uint32_t> &C, const std::vector<uint8_t> &pass,
                       const std::vector<uint32_t> &fcnt, const std::vector<Prev> &Lprev,
                       std::vector<uint64_t> &idx, std::vector<float> &X) {
        const uint64_t nC = C.size() / k;
        idx.clear();
        for (uint64_t i = 0; i < nC; i++)
            if (pass[i]) idx.push_back(i);
        // 3x3 join work of every feature pair, computed once per binning in
        // one launch.  The work is symmetric (the neighbour-cell relation is),
        // so each pair is summed once over the cells of its feature with
        // fewer non-empty cells.
        if (c->pairwork.empty() && c->F) {
            std::vector<uint32_t> pairs;
            auto runs = [&](uint32_t f) { return c->run_off[f + 1] - c->run_off[f]; };
            for (uint32_t a = 0; a < c->F; a++)
                for (uint32_t b = a + 1; b < c->F; b++) {
                    const bool swap = runs(b) < runs(a);
                    pairs.push_back(swap ? b : a);
                    pairs.push_back(swap ? a : b);
                }
            std::vector<uint64_t> out(pairs.size() / 2, 0);
            int rc = pairwork_batch(c, pairs.data(), (uint32_t)out.size(), out.data(), s);
            if (rc) return rc;
            std::vector<uint64_t> pw((size_t)c->F * c->F, 0);
            for (size_t i = 0; i < out.size(); i++) {
                const uint32_t a = pairs[2 * i], b = pairs[2 * i + 1];
                pw[(size_t)a * c->F + b] = pw[(size_t)b * c->F + a] = out[i];
            }
            c->pairwork = std::move(pw);
        }
        std::vector<uint64_t> pw(idx.size(), 0);
        for (size_t u = 0; u < idx.size(); u++)
            pw[u] = c->pairwork[(size_t)C[idx[u] * k] * c->F + C[idx[u] * k + k - 1]];
        X.assign(idx.size() * COLO_MLP_IN, 0.f);
        for (size_t u = 0; u < idx.size(); u++) {
            const uint32_t *cand = &C[idx[u] * k];
            uint64_t na = c->fcount[cand[0]];
            if (k > 2) {
                int64_t pi = find(Lprev, cand, k - 1);
                na = pi >= 0 ? Lprev[pi].rows : 0;
            }
            cand_features(k, cand, &fcnt[idx[u] * k], na, pw[u], &X[u * COLO_MLP_IN]);
        }
        return COLO_OK;
    }

    // Refinement tests of the counted candidates idx (the cost model's
    // second target, measured for the log only): at size 2 the second
    // feature's instances in the 3x3 cells of each anchor (the pair's join
    // work), at size k the neighbours in the last feature of each parent
    // row's first instance.
    int refine_tests(uint32_t k, const std::vector<uint32_t> &C, const std::vector<uint64_t> &idx,
                     const std::vector<Prev> &Lprev, std::vector<uint64_t> &out) {
        out.assign(idx.size(), 0);
        if (k == 2) {
            for (size_t u = 0; u < idx.size(); u++)
                out[u] = c->pairwork[(size_t)C[idx[u] * 2] * c->F + C[idx[u] * 2 + 1]];
            return COLO_OK;
        }
        std::vector<RefineItem> items(idx.size());
        for (size_t u = 0; u < idx.size(); u++) {
            const uint32_t *cand = &C[idx[u] * k];
            int64_t pi = find(Lprev, cand, k - 1);
            const Prev *nb = nbr_list(cand[0], cand[k - 1]);
            if (pi < 0 || !Lprev[pi].tptr || !nb) return set_error(c, COLO_EARG, "internal: refinement tests");
            items[u] = RefineItem{Lprev[pi].tptr, nb->offs->as<uint64_t>() + nb->offs_at, Lprev[pi].rows, k - 1,
                                  c->fstart[cand[0]]};
        }
        return refine_tests_batch(c, items.data(), (uint32_t)items.size(), out.data(), s);
    }

    int log_records(uint32_t k, const std::vector<uint32_t> &C, const std::vector<uint64_t> &idx,
                    const std::vector<float> &X, const std::vector<CandRes> &res, const std::vector<Prev> &Lprev) {
        if (!o.log_path || idx.empty()) return COLO_OK;
        std::vector<uint64_t> tests;
        int rc = refine_tests(k, C, idx, Lprev, tests);
        if (rc) return rc;
        FILE *fp = fopen(o.log_path, "ab");
        if (!fp) return set_error(c, COLO_EFILE, "cannot open the cost-model log");
        std::vector<float> rec(COLO_MLP_RECORD);
        for (size_t u = 0; u < idx.size(); u++) {
            memcpy(rec.data(), &X[u * COLO_MLP_IN], COLO_MLP_IN * 4);
            rec[COLO_MLP_IN] = (float)std::log1p((double)res[idx[u]].rows);
            rec[COLO_MLP_IN + 1] = (float)std::log1p((double)tests[u]);
            rec[COLO_MLP_IN + 2] = (float)k;
            fwrite(rec.data(), 4, rec.size(), fp);
        }
        return fclose(fp) == 0 ? COLO_OK : set_error(c, COLO_EFILE, "cannot write the cost-model log");
    }

    // The filter's counts and pass flags of a level's candidates (cpp:297,
    // 760); in a sweep, counts known from an earlier threshold need only the
    // threshold test.
    DevBuf d_cands{&c->pool}, d_counts{&c->pool}, d_pass{&c->pool};
    int level_filter(uint32_t k, const std::vector<uint32_t> &C, std::vector<uint32_t> &fcnt,
                     std::vector<uint8_t> &pass) {
        auto tf = Clock::now();
        const uint64_t nC = C.size() / k;
        fcnt.assign(nC * k, 0);
        pass.assign(nC, 1);
        if (o.filter_on && nC) {
            if (!d_cands.ensure(C.size() * 4) || !d_counts.ensure(C.size() * 4) || !d_pass.ensure(nC))
                return set_error(c, COLO_ENOMEM, "filter buffers");
            COLO_HIP(hipMemcpyAsync(d_cands.p, C.data(), C.size() * 4, hipMemcpyHostToDevice, s));
            bool cached = memo != nullptr;
            if (memo)
                for (uint64_t i = 0; i < nC && cached; i++) {
                    auto it = memo->fcnt.find(std::vector<uint32_t>(&C[i * k], &C[i * k] + k));
                    if (it == memo->fcnt.end()) cached = false;
                    else memcpy(&fcnt[i * k], it->second.data(), k * 4);
                }
            size_t e0 = mark();
            int rc;
            if (cached) {
                COLO_HIP(hipMemcpyAsync(d_counts.p, fcnt.data(), fcnt.size() * 4, hipMemcpyHostToDevice, s));
                rc = threshold_batch(c, d_cands.as<uint32_t>(), (uint32_t)nC, k, o.theta, d_counts.as<uint32_t>(),
                                     d_pass.as<uint8_t>(), s);
            } else {
                rc = filter_batch(c, d_cands.as<uint32_t>(), (uint32_t)nC, k, o.theta, d_counts.as<uint32_t>(),
                                  d_pass.as<uint8_t>(), s);
            }
            if (rc) return rc;
            span(e0, mark(), kFilter);
            if (!cached) COLO_HIP(hipMemcpyAsync(fcnt.data(), d_counts.p, fcnt.size() * 4, hipMemcpyDeviceToHost, s));
            COLO_HIP(hipMemcpyAsync(pass.data(), d_pass.p, nC, hipMemcpyDeviceToHost, s));
            if ((rc = sync())) return rc;
            st.launches += cached ? 1 : 2;
            if (memo && !cached)
                for (uint64_t i = 0; i < nC; i++)
                    memo->fcnt[std::vector<uint32_t>(&C[i * k], &C[i * k] + k)] =
                        std::vector<uint32_t>(&fcnt[i * k], &fcnt[i * k] + k);
        }
        st.filter_ms += ms_since(tf);
        return COLO_OK;
    }
    // Subset bounds of a level's candidates that pass the filter (k >= 3).
    int level_bounds(uint32_t k, const std::vector<uint32_t> &C, const std::vector<uint8_t> &pass,
                     const std::vector<Prev> &Lprev, std::vector<float> &sbq) {
        auto tb = Clock::now();
        int rc = subset_bounds(k, C, pass, Lprev, sbq);
        st.bound_ms += ms_since(tb);
        return rc;
    }
    // The next level's filter results and bounds, computed before this
    // level's tables are built (only the tables that surviving candidates
    // extend are built when no table checksums are asked for).
    bool pre_valid_ = false;
    std::vector<uint32_t> pre_fcnt_;
    std::vector<uint8_t> pre_pass_;
    std::vector<float> pre_sbq_;

    int run(char **report) {
        auto T0 = Clock::now();
        const double sys_ms0 = c->pool.sys_ms, sys_bytes0 = c->pool.sys_bytes;
        c->pool.peak = c->pool.in_use;
        const uint32_t F = c->F;
        rep.printf("colo-report v1\n");
        rep.printf("grid gx=%u gy=%u cells=%u n=%u F=%u r2=%08x theta=%08x\n", c->gx, c->gy, c->cells, c->n, F,
                   fbits(c->r2), fbits(o.theta));
        std::vector<uint32_t> C;
        for (uint32_t i = 0; i + 1 < F; i++)  // cpp:75-82
            for (uint32_t j = i + 1; j < F; j++) {
                C.push_back(i);
                C.push_back(j);
            }
        std::vector<Prev> Lprev;
        uint32_t k = 2;
        uint64_t total_prev = 0;
        uint32_t levels = 0;
        while (!C.empty() && k <= F && (o.max_k == 0 || k <= o.max_k)) {
            levels++;
            cur_k_ = k;
            auto TL = Clock::now();
            const uint64_t nC = C.size() / k;
            st.candidates += nC;
            // filter (cpp:297, 760) and, from level 3, subset bounds: computed
            // here, or already by the previous level (to choose its tables)
            std::vector<uint32_t> fcnt;
            std::vector<uint8_t> pass;
            std::vector<float> sbq;
            if (pre_valid_) {
                fcnt.swap(pre_fcnt_);
                pass.swap(pre_pass_);
                sbq.swap(pre_sbq_);
                pre_valid_ = false;
            } else {
                int rc0 = level_filter(k, C, fcnt, pass);
                if (!rc0 && o.subset_bound && k >= 3) rc0 = level_bounds(k, C, pass, Lprev, sbq);
                if (rc0) return rc0;
            }
            uint64_t npass = 0;
            for (uint8_t p : pass) npass += p;
            std::vector<CandRes> res(nC);
            std::vector<Prev> L;
            std::shared_ptr<DevBuf> arena;
            auto tl = Clock::now();
            // subset bound (from level 3): candidates whose bound is below
            // theta cannot be prevalent and are not counted
            std::vector<uint8_t> live = pass;
            if (!sbq.empty())
                for (uint64_t i = 0; i < nC; i++)
                    if (pass[i] && sbq[i] < o.theta) {
                        live[i] = 0;
                        st.pruned++;
                    }
            keep_parts_ = o.subset_bound && (o.max_k == 0 || k + 1 <= o.max_k);
            // candidates to count: those passing both tests whose results
            // are not known from an earlier run of a sweep
            std::vector<uint8_t> todo = live;
            if (memo)
                for (uint64_t i = 0; i < nC; i++) {
                    if (!live[i]) continue;
                    auto it = memo->res.find(std::vector<uint32_t>(&C[i * k], &C[i * k] + k));
                    if (it == memo->res.end()) {
                        memo->misses++;
                        continue;
                    }
                    memo->hits++;
                    todo[i] = 0;
                    CandRes &r = res[i];
                    r.pass = true;
                    r.rows = it->second.rows;
                    memcpy(r.numc, it->second.numc, sizeof r.numc);
                    memcpy(r.nume, it->second.nume, sizeof r.nume);
                    r.pic = it->second.pic;
                    r.pie = it->second.pie;
                    r.prevalent = !((o.pi_mode ? r.pie : r.pic) < o.theta);
                }
            // cost-model features of the passing candidates (logging and the
            // learned schedule); predictions of log1p(rows)
            const bool learned = o.sched == COLO_SCHED_LEARNED && o.model;
            std::vector<uint64_t> fidx;
            std::vector<float> FX, pred;
            auto tm = Clock::now();
            if (learned || o.log_path) {
                int rc0 = level_features(k, C, todo, fcnt, Lprev, fidx, FX);
                if (rc0) return rc0;
            }
            if (learned && !fidx.empty()) {
                pred.resize(fidx.size() * COLO_MLP_OUT);
                int rc0 = mlp_predict_host(o.model, FX.data(), (uint32_t)fidx.size(), pred.data());
                if (rc0) return set_error(c, rc0, "cost model prediction");
            }
            st.model_ms += ms_since(tm);
            level_fcnt_ = o.filter_on ? &fcnt : nullptr;
            int rc = o.sched == COLO_SCHED_PLAIN
                         ? run_level_plain(k, C, todo, Lprev, res, L)
                         : run_level_persistent(k, C, todo, Lprev, res, L, arena, learned ? &pred : nullptr);
            level_fcnt_ = nullptr;
            if (rc) return rc;
            if (memo) {
                for (uint64_t i = 0; i < nC; i++)
                    if (todo[i]) {
                        SweepMemo::Res m{};
                        m.rows = res[i].rows;
                        memcpy(m.numc, res[i].numc, sizeof m.numc);
                        memcpy(m.nume, res[i].nume, sizeof m.nume);
                        m.pic = res[i].pic;
                        m.pie = res[i].pie;
                        memo->res.emplace(std::vector<uint32_t>(&C[i * k], &C[i * k] + k), m);
                    }
                // the prevalent list in candidate order: counted ones carry
                // their slots, remembered ones are added without
                std::vector<Prev> all;
                size_t j = 0;
                for (uint64_t i = 0; i < nC; i++) {
                    if (!res[i].prevalent) continue;
                    if (todo[i]) {
                        all.push_back(std::move(L[j++]));
                        continue;
                    }
                    Prev p;
                    p.pat.assign(&C[i * k], &C[i * k] + k);
                    p.rows = res[i].rows;
                    if (k == 2) {
                        p.nslots = c->fcount[C[i * k]];
                    } else {
                        p.parent = find(Lprev, &C[i * k], k - 1);
                        if (p.parent < 0) return set_error(c, COLO_EARG, "internal: missing parent");
                        p.nslots = Lprev[p.parent].rows;
                    }
                    all.push_back(std::move(p));
                }
                L = std::move(all);
            }
            double level_count_ms = ms_since(tl);
            if (learned)
                for (size_t u = 0; u < fidx.size(); u++) {
                    st.model_abs_err +=
                        std::fabs((double)pred[u * COLO_MLP_OUT] - std::log1p((double)res[fidx[u]].rows));
                    st.model_n += 1;
                }
            if ((rc = log_records(k, C, fidx, FX, res, Lprev))) return rc;
            rep.printf("level %u cand %llu pass %llu prev %llu\n", k, (unsigned long long)nC,
                       (unsigned long long)npass, (unsigned long long)L.size());
            for (uint64_t i = 0; i < nC; i++) {
                const uint32_t *cand = &C[i * k];
                rep.printf("c");
                for (uint32_t m = 0; m < k; m++) rep.printf(" %u", cand[m]);
                rep.printf(" f");
                for (uint32_t m = 0; m < k; m++) rep.printf(" %u", fcnt[i * k + m]);
                rep.printf(" p %u", (unsigned)pass[i]);
                if (!pass[i]) {
                    rep.printf("\n");
                    continue;
                }
                if (!sbq.empty()) {
                    rep.printf(" q %08x", fbits(sbq[i]));
                    if (!live[i]) {
                        rep.printf("\n");
                        continue;
                    }
                }
                const CandRes &r = res[i];
                rep.printf(" r %llu n", (unsigned long long)r.rows);
                for (uint32_t m = 0; m < k; m++) rep.printf(" %llu", (unsigned long long)r.numc[m]);
                rep.printf(" e");
                for (uint32_t m = 0; m < k; m++) rep.printf(" %llu", (unsigned long long)r.nume[m]);
                rep.printf(" pi %08x %08x v %d\n", fbits(r.pic), fbits(r.pie), r.prevalent ? 1 : 0);
            }
            total_prev += L.size();
            auto th = Clock::now();
            write_patterns(k, L);
            // next level candidates (cpp:356, 1124)
            std::vector<uint32_t> N;
            if (o.max_k == 0 || k + 1 <= o.max_k) {
                std::vector<uint32_t> flat;
                flat.reserve(L.size() * k);
                for (const Prev &p : L) flat.insert(flat.end(), p.pat.begin(), p.pat.end());
                N = gen_candidates(flat, k);
                // tableGenRequired (cpp:479-515) joins the candidates themselves;
                // record where it would have skipped tables the next level needs
                if (k >= 3 && !N.empty() && gen_candidates(C, k).empty()) rep.printf("q10 level %u\n", k);
            }
            std::vector<uint8_t> need(L.size(), 0);
            {
                // prefixes of the next level's candidates (cpp:364-372, 1055-1065)
                std::unordered_set<std::vector<uint32_t>, PatSet::H> pref(16, PatSet::H{k});
                for (uint64_t i = 0; i < N.size() / (k + 1); i++)
                    pref.emplace(N.begin() + i * (k + 1), N.begin() + i * (k + 1) + k);
                for (size_t i = 0; i < L.size(); i++) need[i] = pref.count(L[i].pat) ? 1 : 0;
            }
            // tables to build: the reference's (prefixes of the next level's
            // candidates); at level 2 every prevalent pair, whose tables serve
            // as the anchors' neighbour lists at the later levels.  Without
            // table checksums, only the tables that the next level's
            // candidates passing its filter and subset bound read: their
            // prefixes and, at level 2, also their (first, last) pairs.  Every
            // (first, last) pair of a candidate counted at a later level is
            // one of a counted size-3 candidate (its size-3 subsets are
            // prevalent).  The others' tables would never be read; the report
            // lists the reference's tables and their rows either way.
            std::vector<uint8_t> build = need;
            if (k == 2 && !N.empty())
                for (auto &b : build) b = 1;
            if (!N.empty() && !o.hash_tables) {
                const auto tp = Clock::now();  // filter and bound time, not host time
                index_prev(L);
                int rc0 = level_filter(k + 1, N, pre_fcnt_, pre_pass_);
                pre_sbq_.clear();
                if (!rc0 && o.subset_bound) rc0 = level_bounds(k + 1, N, pre_pass_, L, pre_sbq_);
                index_prev(Lprev);  // this level's parent lookups (recounts)
                if (rc0) return rc0;
                pre_valid_ = true;
                std::unordered_set<std::vector<uint32_t>, PatSet::H> live_pref(16, PatSet::H{k});
                for (uint64_t i = 0; i < N.size() / (k + 1); i++)
                    if (pre_pass_[i] && !(!pre_sbq_.empty() && pre_sbq_[i] < o.theta)) {
                        live_pref.emplace(N.begin() + i * (k + 1), N.begin() + i * (k + 1) + k);
                        if (k == 2) live_pref.emplace(std::vector<uint32_t>{N[i * 3], N[i * 3 + 2]});
                    }
                for (size_t i = 0; i < L.size(); i++)
                    if (build[i] && !live_pref.count(L[i].pat)) {
                        build[i] = 0;
                        st.tables_skipped++;
                    }
                st.host_ms -= ms_since(tp);
            }
            st.host_ms += ms_since(th);
            auto tt = Clock::now();
            if (memo) {
                // tables from earlier runs
                for (size_t i = 0; i < L.size(); i++) {
                    if (!build[i]) continue;
                    if (k == 2) {
                        auto it = memo->l2.find({L[i].pat[0], L[i].pat[1]});
                        if (it == memo->l2.end()) continue;
                        L[i] = it->second;
                    } else {
                        auto it = memo->tk.find(L[i].pat);
                        if (it == memo->tk.end()) continue;
                        L[i].table = it->second.first;
                        L[i].tptr = it->second.second;
                    }
                    build[i] = 0;
                }
            }
            // tables to build whose slots were not kept (deferred above the
            // slots budget, or remembered from an earlier threshold)
            if ((rc = recount_missing_slots(k, C, L, build, Lprev))) return rc;
            rc = o.sched == COLO_SCHED_PLAIN ? build_tables_plain(k, L, build, Lprev)
                                             : build_tables_persistent(k, L, build, Lprev, C);
            if (rc) return rc;
            if (k == 2) st.l2_join_ms = level_count_ms + ms_since(tt);
            for (size_t i = 0; i < L.size(); i++) {
                if (!need[i]) continue;
                const Prev &p = L[i];
                rep.printf("t");
                for (uint32_t m = 0; m < k; m++) rep.printf(" %u", p.pat[m]);
                rep.printf(" rows %llu", (unsigned long long)p.rows);
                if (o.hash_tables) {
                    uint64_t h = 0;
                    // batched tables hold sorted positions, plain ones instance ids
                    const uint32_t *map = o.sched == COLO_SCHED_PLAIN ? nullptr : c->d_sidx;
                    if ((rc = table_hash(c, p.tptr, p.rows, k, map, &h, s))) return rc;
                    rep.printf(" h %016llx", (unsigned long long)h);
                }
                rep.printf("\n");
            }
            for (Prev &p : L) p.slots.reset();
            if (memo)
                for (const Prev &p : L)
                    if (p.part) memo->part.emplace(p.pat, std::make_pair(p.part, p.part_off));
            if (memo && k == 2)
                for (const Prev &p : L)
                    if (p.tptr && p.nptr) memo->l2.emplace(std::make_pair(p.pat[0], p.pat[1]), p);
            if (memo && k >= 3)
                for (const Prev &p : L)
                    if (p.tptr) memo->tk.emplace(p.pat, std::make_pair(p.table, p.tptr));
            if (k == 2) {
                l2_index_.assign((size_t)c->F * c->F, -1);
                for (size_t i = 0; i < L.size(); i++) {
                    l2_index_[(size_t)L[i].pat[0] * c->F + L[i].pat[1]] = (int64_t)l2_.size();
                    l2_.push_back(L[i]);  // shares the table and offsets
                    l2_.back().part.reset();
                }
            } else {
                for (Prev &p : L) p.offs.reset();
            }
            Lprev = std::move(L);
            index_prev(Lprev);
            C = std::move(N);
            if (k < 16) st.level_ms[k] = ms_since(TL);
            k++;
        }
        rep.printf("end levels %u patterns %llu\n", levels, (unsigned long long)total_prev);
        st.levels = levels;
        st.patterns = total_prev;
        st.total_ms = ms_since(T0);
        st.alloc_ms = c->pool.sys_ms - sys_ms0;
        st.alloc_bytes = c->pool.sys_bytes - sys_bytes0;
        st.pool_peak_bytes = (double)c->pool.peak;
        if (report) {
            *report = (char *)malloc(rep.s.size() + 1);
            memcpy(*report, rep.s.c_str(), rep.s.size() + 1);
        }
        return COLO_OK;
    }
};

}  // namespace

extern "C" {

int colo_mine(colo_ctx *c, const colo_mine_opts *opts, char **report, colo_mine_stats *stats) {
    if (!c || !opts) return COLO_EARG;
    if (report) *report = nullptr;
    Miner M(c, *opts);
    int rc = M.run(report);
    (void)hipStreamSynchronize(c->stream);  // before buffers return to the pool
    if (stats) *stats = M.st;
    return rc;
}

int colo_mine_sweep(colo_ctx *c, const colo_mine_opts *opts, const float *thetas, uint32_t n, char **reports,
                    colo_mine_stats *stats) {
    if (!c || !opts || (!thetas && n)) return COLO_EARG;
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return thetas[a] < thetas[b]; });
    SweepMemo memo;
    int rc = COLO_OK;
    for (uint32_t i : order) {
        colo_mine_opts o = *opts;
        o.theta = thetas[i];
        std::string prefix;
        if (opts->out_prefix) {
            char b[64];
            snprintf(b, sizeof b, "theta_%g_", (double)thetas[i]);
            prefix = std::string(opts->out_prefix) + (n > 1 ? b : "");
            o.out_prefix = prefix.c_str();
        }
        Miner M(c, o);
        M.memo = &memo;
        if (reports) reports[i] = nullptr;
        rc = M.run(reports ? &reports[i] : nullptr);
        (void)hipStreamSynchronize(c->stream);
        if (stats) stats[i] = M.st;
        if (rc) break;
    }
    return rc;
}

// ------------------------------------------------------------ autotuner

static std::string tune_key(const colo_ctx *c, const colo_mine_opts &o) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t p;
    std::string arch = "unknown";
    int cus = 0;
    if (hipGetDeviceProperties(&p, dev) == hipSuccess) {
        arch = p.gcnArchName;
        cus = p.multiProcessorCount;
    } else {
        (void)hipGetLastError();
    }
    for (char &ch : arch)
        if (ch == ' ') ch = '_';
    char b[256];
    snprintf(b, sizeof b, "v5/%s/%dcu/n%u/F%u/r%08x/t%08x/k%u/s%d/b%d", arch.c_str(), cus, c->n, c->F,
             fbits(c->r2), fbits(o.theta), o.max_k, o.sched, o.subset_bound ? 1 : 0);
    return b;
}

// Cache lines: key, count_var, write_var, blocks_per_cu, item_rows, the 16
// per-level count variants, the 16 per-level item sizes.
static bool tune_load(const char *path, const std::string &key, TuneParams *t) {
    FILE *fp = fopen(path, "rb");
    if (!fp) return false;
    char line[2048];
    bool found = false;
    while (fgets(line, sizeof line, fp)) {
        char k[256];
        int used = 0;
        if (sscanf(line, "%255s%n", k, &used) != 1 || key != k) continue;
        constexpr int nf = 4 + 16 + 16 + 16;
        unsigned long v[nf];
        const char *p = line + used;
        int n = 0;
        for (; n < nf; n++) {
            char *e;
            v[n] = strtoul(p, &e, 10);
            if (e == p) break;
            p = e;
        }
        if (n != nf) continue;
        t->count_var = (uint32_t)v[0];  // the last entry of a key wins
        t->write_var = (uint32_t)v[1];
        t->blocks_per_cu = (uint32_t)v[2];
        t->item_rows = (uint32_t)v[3];
        for (int i = 0; i < 16; i++) {
            t->count_var_k[i] = (uint32_t)v[4 + i];
            t->item_rows_k[i] = (uint32_t)v[20 + i];
        }
        for (int i = 0; i < 16; i++) t->write_rows_k[i] = (uint32_t)v[36 + i];
        found = true;
    }
    fclose(fp);
    return found;
}

int colo_autotune(colo_ctx *c, const colo_mine_opts *opts, const char *cache_path, colo_tune *out) {
    if (!c || !opts) return COLO_EARG;
    colo_mine_opts o = *opts;
    o.out_prefix = nullptr;
    o.log_path = nullptr;
    o.hash_tables = 0;
    if (o.sched == COLO_SCHED_PLAIN) o.sched = COLO_SCHED_PERSISTENT;
    const std::string key = tune_key(c, o);
    TuneParams best = c->tune;
    if (!(cache_path && tune_load(cache_path, key, &best))) {
        // kernel time of one run: {level-k count, level-k write, all batched
        // passes}; lvl: wall time per level
        auto measure = [&](const TuneParams &t, double *m3) -> int {
            c->tune = t;
            Miner M(c, o);
            int rc = M.run(nullptr);
            (void)hipStreamSynchronize(c->stream);
            if (rc) return rc;
            m3[0] = M.st.count_kernel_ms - M.st.l2_count_kernel_ms;
            // a write variant also decides the offsets scan it needs
            m3[1] = M.st.write_kernel_ms - M.st.l2_write_kernel_ms + M.st.scan_kernel_ms;
            m3[2] = M.st.count_kernel_ms + M.st.write_kernel_ms + M.st.scan_kernel_ms;
            if (c->verbose)
                fprintf(stderr, "[colo tune] run: write var %u, %u workgroups per CU: count %.1f ms, write %.1f ms\n",
                        t.write_var, t.blocks_per_cu, M.st.count_kernel_ms, M.st.write_kernel_ms);
            return COLO_OK;
        };
        auto fail = [&](int rc) {
            c->tune = best;
            return rc;
        };
        for (int k = 0; k < 16; k++) best.count_var_k[k] = best.item_rows_k[k] = best.write_rows_k[k] = COLO_TUNE_UNSET;
        double cur[3], m[3];
        int rc = measure(best, cur);  // warm-up (pool, cached features)
        if (rc) return fail(rc);
        // 1. per level: the count variant and item size, timed online inside
        //    one mining run (Miner::tune_level)
        {
            c->tune = best;
            Miner M(c, o);
            M.tuning_ = true;
            rc = M.run(nullptr);
            (void)hipStreamSynchronize(c->stream);
            if (rc) return fail(rc);
            best = c->tune;
            // levels not seen here take the deepest level's variant
            for (int k = 15; k >= 3; k--)
                if (best.count_var_k[k] != COLO_TUNE_UNSET) {
                    best.count_var = best.count_var_k[k];
                    break;
                }
        }
        if ((rc = measure(best, cur))) return fail(rc);
        // 2. whole-run parameters, one at a time, by full runs
        struct Axis {
            uint32_t TuneParams::*field;
            std::vector<uint32_t> values;
            int metric;
        };
        const std::vector<Axis> axes = {{&TuneParams::write_var, {0, 1, 2, 3}, 1},
                                        {&TuneParams::blocks_per_cu, {4, 8, 16}, 2}};
        for (const Axis &ax : axes) {
            const uint32_t start = best.*(ax.field);
            for (uint32_t v : ax.values) {
                if (v == start) continue;
                TuneParams t = best;
                t.*(ax.field) = v;
                if ((rc = measure(t, m))) return fail(rc);
                if (m[ax.metric] < cur[ax.metric]) {
                    best = t;
                    memcpy(cur, m, sizeof cur);
                }
            }
        }
        if (cache_path) {
            FILE *fp = fopen(cache_path, "ab");
            if (fp) {
                fprintf(fp, "%s %u %u %u %u", key.c_str(), best.count_var, best.write_var, best.blocks_per_cu,
                        best.item_rows);
                for (int k = 0; k < 16; k++) fprintf(fp, " %u", best.count_var_k[k]);
                for (int k = 0; k < 16; k++) fprintf(fp, " %u", best.item_rows_k[k]);
                for (int k = 0; k < 16; k++) fprintf(fp, " %u", best.write_rows_k[k]);
                fprintf(fp, "\n");
                fclose(fp);
            }
        }
    }
    c->tune = best;
    if (out) colo_get_tune(c, out);
    return COLO_OK;
}

int colo_run_config(const char *config_path, const colo_mine_opts *opts, char **timing_line) {
    using Clock = std::chrono::steady_clock;
    auto T0 = Clock::now();
    FILE *fp = fopen(config_path, "rb");
    if (!fp) return COLO_EFILE;
    std::vector<std::string> lines;
    {
        char b[8192];
        while (lines.size() < 5 && fgets(b, sizeof b, fp)) {
            size_t l = strlen(b);
            if (l && b[l - 1] == '\n') b[--l] = 0;
            lines.emplace_back(b);
        }
    }
    fclose(fp);
    while (lines.size() < 5) lines.emplace_back();
    const std::string &location = lines[0], &param = lines[1], &dataset = lines[2];
    // parameter.txt: r theta filterFlag (cpp:66-68)
    float r = 0, theta = 0;
    {
        FILE *pf = fopen(param.c_str(), "rb");
        if (!pf) return COLO_EFILE;
        char a[128], b2[128], d[128];
        int k = fscanf(pf, "%127s %127s %127s", a, b2, d);
        fclose(pf);
        if (k < 2 || !parse_flt(a, &r) || !parse_flt(b2, &theta)) return COLO_EPARSE;
    }
    auto tio = Clock::now();
    colo_host_data *h = nullptr;
    int rc = colo_load_csv(dataset.c_str(), &h);
    if (rc) return rc;
    double io_ms = std::chrono::duration<double, std::milli>(Clock::now() - tio).count();
    colo_dataset ds;
    colo_host_dataset(h, &ds);
    colo_ctx *ctx = nullptr;
    rc = colo_ctx_create(&ds, r, nullptr, &ctx);
    if (rc) {
        if (ctx) fprintf(stderr, "colo: %s\n", colo_last_error(ctx));
        colo_ctx_destroy(ctx);
        colo_host_free(h);
        return rc;
    }
    colo_mine_opts o{};
    if (opts) o = *opts;
    else { o.filter_on = 1; o.sched = COLO_SCHED_PERSISTENT; }
    o.theta = theta;
    o.out_prefix = location.c_str();
    colo_mine_stats st{};
    rc = colo_mine(ctx, &o, nullptr, &st);
    if (rc) fprintf(stderr, "colo: %s\n", colo_last_error(ctx));
    double e2e = std::chrono::duration<double, std::milli>(Clock::now() - T0).count();
    if (timing_line && !rc) {
        // the reference's summary line (cpp:1144), fields as its comment at
        // cpp:1143 names them, each as cout prints a float (%g): dataset,
        // executable, theta, io (CSV parse, cpp:203), transfer (host-to-device
        // copy of the dataset, cpp:240), grid (binning, cpp:251), filter (all
        // levels, cpp:307, 770), refinement (count passes and PI, cpp:422,
        // 850), table and candidate generation (t1, cpp:1129), end to end
        // (ms).  The leading newline of cpp:1144 is the caller's to print.
        auto g = [](double v) { return (double)(float)v; };  // the reference's times are float
        char b[1024];
        snprintf(b, sizeof b, "%s , colo_miner , %g , %g , %g , %g , %g , %g , %g , %g", dataset.c_str(),
                 (double)theta, g(io_ms), g(ctx->upload_ms), g(ctx->bin_ms), g(st.filter_ms), g(st.count_ms + st.pi_ms),
                 g(st.scan_ms + st.write_ms + st.host_ms), g(e2e));
        *timing_line = strdup(b);
    }
    colo_ctx_destroy(ctx);
    colo_host_free(h);
    return rc;
}

}  // extern "C"
