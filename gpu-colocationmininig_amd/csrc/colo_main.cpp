// colo_main.cpp -- command-line drop-in for the reference executable
// (Codes/main.cpp:8-48 + colocationFinder::Begin, cpp:1068-1147):
//   colo_miner <config.txt> [--sched plain|persistent|learned] [--model F]
//              [--max-k K] [--pi-mode compat|exact] [--no-filter]
//              [--no-subset-bound]
// --sched learned uses the cost model F (default: models/cost_mlp.bin
// beside the executable).  The subset bound (on by default) skips
// candidates that it proves non-prevalent; the pattern files are the same
// either way.
// Reads the reference's config.txt / parameter.txt / CSV, writes
// <location>ColocationDegree_<k>.txt and prints the timing line.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/colo.h"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::printf("Configuration file missing...");
        return 0;
    }
    colo_mine_opts o{};
    o.filter_on = 1;
    o.subset_bound = 1;
    o.sched = COLO_SCHED_PERSISTENT;
    std::string model_path;
    {
        std::string exe(argv[0]);
        size_t sl = exe.find_last_of('/');
        model_path = (sl == std::string::npos ? std::string(".") : exe.substr(0, sl)) + "/models/cost_mlp.bin";
    }
    for (int i = 2; i < argc; i++) {
        if (!std::strcmp(argv[i], "--sched") && i + 1 < argc) {
            const char *v = argv[++i];
            o.sched = !std::strcmp(v, "plain")     ? COLO_SCHED_PLAIN
                      : !std::strcmp(v, "learned") ? COLO_SCHED_LEARNED
                                                   : COLO_SCHED_PERSISTENT;
        } else if (!std::strcmp(argv[i], "--model") && i + 1 < argc) {
            model_path = argv[++i];
        } else if (!std::strcmp(argv[i], "--max-k") && i + 1 < argc) {
            o.max_k = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
        } else if (!std::strcmp(argv[i], "--pi-mode") && i + 1 < argc) {
            o.pi_mode = !std::strcmp(argv[++i], "exact") ? 1 : 0;
        } else if (!std::strcmp(argv[i], "--no-filter")) {
            o.filter_on = 0;
        } else if (!std::strcmp(argv[i], "--no-subset-bound")) {
            o.subset_bound = 0;
        } else {
            std::fprintf(stderr, "unknown option %s\n", argv[i]);
            return 2;
        }
    }
    colo_mlp *model = nullptr;
    if (o.sched == COLO_SCHED_LEARNED) {
        if (colo_mlp_load(model_path.c_str(), &model)) {
            std::fprintf(stderr, "colo_miner: cannot load the cost model %s\n", model_path.c_str());
            return 1;
        }
        o.model = model;
    }
    char *line = nullptr;
    int rc = colo_run_config(argv[1], &o, &line);
    if (model) colo_mlp_destroy(model);
    if (rc) {
        std::fprintf(stderr, "colo_miner: error %d\n", rc);
        return 1;
    }
    std::printf("\n%s", line);
    colo_free_str(line);
    return 0;
}
