// Drives the host-only half of libcolo (gpu-colocationmininig_amd/csrc/
// colo_host.cpp) under AddressSanitizer and UndefinedBehaviorSanitizer:
// the dataset CSV parser on 1..16 threads (valid files, malformed ones,
// files whose row ranges split inside features), its binary cache, and the
// Apriori candidate generation.  Exits non-zero on a failed check.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "colo.h"

#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            return 1;                                                        \
        }                                                                    \
    } while (0)

static void write(const std::string &path, const std::string &s) {
    FILE *fp = std::fopen(path.c_str(), "wb");
    std::fwrite(s.data(), 1, s.size(), fp);
    std::fclose(fp);
}

int main(int argc, char **argv) {
    const std::string tmp = argc > 1 ? argv[1] : "/tmp";
    const std::string csv = tmp + "/san_host.csv", cache = tmp + "/san_host.bin";
    std::string s = "3000\n";
    for (int i = 0; i < 3000; i++) {
        char b[96];
        std::snprintf(b, sizeof b, "%d,%.6f,\"%.6f\"\n", i / 700 * 3 + 1, (i * 7919 % 1000) / 3.0, (i * 104729 % 1000) / 7.0);
        s += b;
    }
    s += "\n0,0,400,200";
    write(csv, s);
    colo_host_data *ref = nullptr;
    CHECK(colo_load_csv_opts(csv.c_str(), nullptr, 1, &ref) == COLO_OK);
    colo_dataset d0;
    CHECK(colo_host_dataset(ref, &d0) == COLO_OK && d0.n == 3000 && d0.F == 5);
    for (uint32_t t : {2u, 3u, 7u, 16u}) {
        colo_host_data *h = nullptr;
        CHECK(colo_load_csv_opts(csv.c_str(), t == 7 ? cache.c_str() : nullptr, t, &h) == COLO_OK);
        colo_dataset d;
        colo_host_dataset(h, &d);
        CHECK(d.n == d0.n && d.F == d0.F && !std::memcmp(d.x, d0.x, 4 * d.n) && !std::memcmp(d.y, d0.y, 4 * d.n));
        CHECK(!std::memcmp(d.fcount, d0.fcount, 4 * d.F));
        for (uint32_t f = 0; f < d.F; f++) CHECK(!std::strcmp(colo_host_label(h, f), colo_host_label(ref, f)));
        colo_host_free(h);
    }
    colo_host_data *hc = nullptr;  // from the cache written above
    CHECK(colo_load_csv_opts(csv.c_str(), cache.c_str(), 4, &hc) == COLO_OK);
    colo_host_free(hc);
    colo_host_free(ref);
    const char *bad[] = {"3\n7,1,2\n\n0,0,10,10", "2\n7,1\n7,2,3\n\n0,0,1,1", "2\n7,1,2\n7,2,3\n", "x\n", "",
                         "2\n7,1,2\n7,\"2,3\n\n0,0,1,1"};
    for (const char *b : bad) {
        write(csv, b);
        for (uint32_t t : {1u, 4u}) {
            colo_host_data *h = nullptr;
            CHECK(colo_load_csv_opts(csv.c_str(), nullptr, t, &h) != COLO_OK && !h);
        }
    }
    const uint32_t L[] = {0, 1, 2, 0, 1, 3, 0, 1, 4, 0, 2, 3, 1, 2, 3, 0, 3, 4};
    std::vector<uint32_t> out(4 * 8);
    const uint64_t n = colo_candidates(L, 6, 3, out.data(), 8);
    CHECK(n == 1 && out[0] == 0 && out[1] == 1 && out[2] == 2 && out[3] == 3);
    CHECK(colo_candidates(nullptr, 0, 3, nullptr, 0) == 0);
    std::printf("host sanitizer run ok\n");
    return 0;
}
