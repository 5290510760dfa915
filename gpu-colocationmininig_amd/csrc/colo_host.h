// colo_host.h -- host-only helpers shared by the miner (colo_miner.cpp) and
// the host library (colo_host.cpp): the Apriori candidate generation and the
// reference's number parsing.  No HIP: colo_host.cpp also builds with plain
// g++ (the sanitizer build in tests/sanitize).
#pragma once

#include <stdint.h>

#include <string>
#include <unordered_set>
#include <vector>

namespace colo_host {

// Hash set of patterns of one size (the subset checks of cpp:618-651).
struct PatSet {
    uint32_t w;
    struct H {
        uint32_t w;
        size_t operator()(const std::vector<uint32_t> &v) const {
            uint64_t h = 1469598103934665603ull;
            for (uint32_t x : v) h = (h ^ x) * 1099511628211ull;
            return (size_t)h;
        }
    };
    std::unordered_set<std::vector<uint32_t>, H> set;
    PatSet(const uint32_t *L, uint64_t P, uint32_t w_) : w(w_), set(P * 2 + 1, H{w_}) {
        for (uint64_t i = 0; i < P; i++) set.emplace(L + i * w, L + (i + 1) * w);
    }
    bool has(const std::vector<uint32_t> &v) const { return set.count(v) != 0; }
};

// candidateColocationGeneral (cpp:517-557) with its first-failure break; see
// colo_host.cpp.
std::vector<uint32_t> gen_candidates(const std::vector<uint32_t> &L, uint32_t km1);

// stof (strtof: decimal -> binary32 correctly rounded), as the reference
// reads parameter.txt and the CSV (cpp:66-68, 185-186).
bool parse_flt(const char *s, float *v);

}  // namespace colo_host
