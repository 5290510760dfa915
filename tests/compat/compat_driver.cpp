// A colocationFinder-style driver written against DeviceProcess_compat.h:
// the reference's level-2 sequence (Codes/colocationFinder.cpp:256-425:
// filter2_2, then per passing pair degree2InstanceTableSlotCounter +
// getParticipationIndex, then getTotalInstances / scalebyConstant /
// degree2InstanceTableGenerator for the pairs that later candidates
// extend) and level-3 sequence (cpp:736-909: filter2, then per passing
// candidate generalInstanceTableSlotCounter + getParticipationIndex, then
// calcInstanceTableGeneral), all through the reference's entry-point names.
//
//   compat_driver <dataset.csv> <r> <theta> <outdir>
//
// Writes <outdir>/result.txt (one line per filter flag, PI and table) and
// each table as <outdir>/t<k>_<features>.bin (u32 instance ids, the
// reference's row order).  Used by tests/test_compat.py.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "DeviceProcess_compat.h"

using namespace colo_compat;

static int fail(const char *what) {
    fprintf(stderr, "compat_driver: %s\n", what);
    return 1;
}

static uint32_t bits_of(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

int main(int argc, char **argv) {
    if (argc != 5) return fail("usage: compat_driver <csv> <r> <theta> <outdir>");
    const Real r = strtof(argv[2], nullptr), theta = strtof(argv[3], nullptr);
    const std::string out = argv[4];
    colo_host_data *h = nullptr;
    if (colo_load_csv(argv[1], &h)) return fail("cannot read the dataset");
    colo_dataset ds{};
    colo_host_dataset(h, &ds);
    const size_t n = ds.n, F = ds.F;
    std::vector<Integer> fid(n), fstart(F, 0), fcount(ds.fcount, ds.fcount + F);
    for (size_t f = 1; f < F; f++) fstart[f] = fstart[f - 1] + fcount[f - 1];
    for (size_t f = 0; f < F; f++)
        for (Integer i = 0; i < fcount[f]; i++) fid[fstart[f] + i] = (Integer)f;
    FILE *res = fopen((out + "/result.txt").c_str(), "w");
    if (!res) return fail("cannot write result.txt");

    // initializeDeviceMemory (cpp:440-458) and setGrid (cpp:102-126)
    Integer *d_instanceList;
    Real *d_x, *d_y;
    if (hipMalloc(&d_instanceList, n * 4) || hipMalloc(&d_x, n * 4) || hipMalloc(&d_y, n * 4)) return fail("hipMalloc");
    (void)hipMemcpy(d_instanceList, fid.data(), n * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(d_x, ds.x, n * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(d_y, ds.y, n * 4, hipMemcpyHostToDevice);
    gridBox g{};
    g.computeZone.minBound = {ds.minX, ds.minY};
    g.computeZone.maxBound = {ds.maxX, ds.maxY + r};
    g.cellWidth = r;
    CountCellFeatures(g, F, n, d_instanceList, d_x, d_y, nullptr, nullptr, nullptr, nullptr);
    if (initializeDeviceMemConst(n, F) != ST_NOERR) return fail("initializeDeviceMemConst");

    Integer maxF = 0;
    for (Integer c : fcount) maxF = c > maxF ? c : maxF;
    Integer *d_slots, *d_bitmap, *d_intermediate;
    Index64 *d_idx;
    const size_t tcap = (size_t)64 << 20;  // table buffer, u32 ids
    if (hipMalloc(&d_slots, tcap * 4) || hipMalloc(&d_bitmap, (size_t)maxF * 8 * 4) ||
        hipMalloc(&d_intermediate, tcap * 4) || hipMalloc(&d_idx, tcap * 4))
        return fail("hipMalloc");

    // ---- degree 2 (cpp:256-425)
    std::vector<Integer> cand2;
    for (Integer a = 0; a < F; a++)
        for (Integer b = a + 1; b < F; b++) {
            cand2.push_back(a);
            cand2.push_back(b);
        }
    const Integer C2 = (Integer)(cand2.size() / 2);
    Integer *d_cand, *d_countMap;
    if (hipMalloc(&d_cand, cand2.size() * 4) || hipMalloc(&d_countMap, cand2.size() * 4)) return fail("hipMalloc");
    (void)hipMemcpy(d_cand, cand2.data(), cand2.size() * 4, hipMemcpyHostToDevice);
    (void)hipMemset(d_countMap, 0, cand2.size() * 4);
    filter2_2(nullptr, nullptr, g, F, nullptr, theta, d_cand, C2, 2, d_countMap);
    std::vector<Integer> countMap(cand2.size());
    (void)hipMemcpy(countMap.data(), d_countMap, countMap.size() * 4, hipMemcpyDeviceToHost);
    std::vector<Integer> prev2;
    std::vector<std::vector<Integer>> slots2;
    for (Integer i = 0; i < C2; i++) {
        const Integer a = cand2[2 * i], b = cand2[2 * i + 1];
        fprintf(res, "F2 %u %u %u\n", a, b, countMap[2 * i]);
        if (!countMap[2 * i]) continue;
        const size_t na = fcount[a];
        (void)hipMemset(d_slots, 0, na * 4);
        (void)hipMemset(d_bitmap, 0, (fcount[a] + fcount[b]) * 4);
        if (degree2InstanceTableSlotCounter(na, d_slots, fstart[a], fstart[a] + na - 1, fstart[b], d_bitmap, b))
            return fail("degree2InstanceTableSlotCounter");
        const Real PI = getParticipationIndex(d_bitmap, 2, i, cand2, d_instanceList, d_instanceList, d_instanceList);
        fprintf(res, "P2 %u %u %08x\n", a, b, bits_of(PI));
        if (PI < theta) continue;
        std::vector<Integer> hs(na);
        (void)hipMemcpy(hs.data(), d_slots, na * 4, hipMemcpyDeviceToHost);
        slots2.push_back(hs);
        prev2.push_back(a);
        prev2.push_back(b);
    }
    const uint64_t P2 = prev2.size() / 2;
    std::vector<Integer> cand3(colo_candidates(prev2.data(), P2, 2, nullptr, 0) * 3);
    colo_candidates(prev2.data(), P2, 2, cand3.data(), cand3.size() / 3);
    const Integer C3 = (Integer)(cand3.size() / 3);
    // tables of the pairs that candidates extend (hasInstanceTable, cpp:353-367)
    std::vector<std::vector<Integer>> table2(P2);
    std::vector<uint8_t> has(P2, 0);
    for (Integer i = 0; i < C3; i++)
        for (uint64_t p = 0; p < P2; p++)
            if (prev2[2 * p] == cand3[3 * i] && prev2[2 * p + 1] == cand3[3 * i + 1]) has[p] = 1;
    for (uint64_t p = 0; p < P2; p++) {
        if (!has[p]) continue;
        const Integer a = prev2[2 * p], b = prev2[2 * p + 1];
        const size_t na = fcount[a];
        const Integer total = getTotalInstances(slots2[p].data(), na);
        if ((size_t)total * 2 > tcap) return fail("table buffer too small");
        (void)scalebyConstant(2, d_idx, na);
        (void)hipMemcpy(d_slots, slots2[p].data(), na * 4, hipMemcpyHostToDevice);
        if (degree2InstanceTableGenerator(na, d_slots, d_idx, d_intermediate, fstart[a], fstart[a] + na - 1, b))
            return fail("degree2InstanceTableGenerator");
        table2[p].resize((size_t)total * 2);
        (void)hipMemcpy(table2[p].data(), d_intermediate, table2[p].size() * 4, hipMemcpyDeviceToHost);
        fprintf(res, "T2 %u %u %u\n", a, b, total);
        FILE *tf = fopen((out + "/t2_" + std::to_string(a) + "_" + std::to_string(b) + ".bin").c_str(), "wb");
        if (!tf) return fail("cannot write a table");
        fwrite(table2[p].data(), 4, table2[p].size(), tf);
        fclose(tf);
    }

    // ---- degree 3 (cpp:736-909)
    if (C3) {
        Integer *d_cand3, *d_countMap3, *d_table1;
        if (hipMalloc(&d_cand3, cand3.size() * 4) || hipMalloc(&d_countMap3, cand3.size() * 4) ||
            hipMalloc(&d_table1, tcap * 4))
            return fail("hipMalloc");
        (void)hipMemcpy(d_cand3, cand3.data(), cand3.size() * 4, hipMemcpyHostToDevice);
        (void)hipMemset(d_countMap3, 0, cand3.size() * 4);
        filter2(nullptr, nullptr, g, F, nullptr, theta, d_cand3, C3, 3, d_countMap3);
        std::vector<Integer> cm3(cand3.size());
        (void)hipMemcpy(cm3.data(), d_countMap3, cm3.size() * 4, hipMemcpyDeviceToHost);
        for (Integer i = 0; i < C3; i++) {
            const Integer *c = &cand3[3 * i];
            fprintf(res, "F3 %u %u %u %u\n", c[0], c[1], c[2], cm3[3 * i]);
            if (!cm3[3 * i]) continue;
            uint64_t p = 0;
            while (!(prev2[2 * p] == c[0] && prev2[2 * p + 1] == c[1])) p++;
            const size_t R1 = table2[p].size() / 2;
            if (R1 > tcap) return fail("slot buffer too small");
            (void)hipMemcpy(d_table1, table2[p].data(), table2[p].size() * 4, hipMemcpyHostToDevice);
            (void)hipMemset(d_slots, 0, R1 * 4);
            (void)hipMemset(d_bitmap, 0, (fcount[c[0]] + fcount[c[1]] + fcount[c[2]]) * 4);
            if (R1 && generalInstanceTableSlotCounter(R1, d_slots, 3, d_table1, d_bitmap, d_cand3 + 3 * i, c[2]))
                return fail("generalInstanceTableSlotCounter");
            const Real PI =
                getParticipationIndex(d_bitmap, 3, i, cand3, d_instanceList, d_instanceList, d_instanceList);
            fprintf(res, "P3 %u %u %u %08x\n", c[0], c[1], c[2], bits_of(PI));
            if (PI < theta || !R1) continue;
            std::vector<Integer> hs(R1);
            (void)hipMemcpy(hs.data(), d_slots, R1 * 4, hipMemcpyDeviceToHost);
            const Integer total = getTotalInstances(hs.data(), R1);
            if ((size_t)total * 3 > tcap) return fail("table buffer too small");
            (void)scalebyConstant(3, d_idx, R1);
            if (calcInstanceTableGeneral(R1, d_slots, d_idx, d_intermediate, 3, d_table1, 0, c[2]))
                return fail("calcInstanceTableGeneral");
            std::vector<Integer> t3((size_t)total * 3);
            (void)hipMemcpy(t3.data(), d_intermediate, t3.size() * 4, hipMemcpyDeviceToHost);
            fprintf(res, "T3 %u %u %u %u\n", c[0], c[1], c[2], total);
            FILE *tf = fopen((out + "/t3_" + std::to_string(c[0]) + "_" + std::to_string(c[1]) + "_" +
                              std::to_string(c[2]) + ".bin").c_str(), "wb");
            if (!tf) return fail("cannot write a table");
            fwrite(t3.data(), 4, t3.size(), tf);
            fclose(tf);
        }
        (void)hipFree(d_cand3);
        (void)hipFree(d_countMap3);
        (void)hipFree(d_table1);
    }
    fclose(res);
    (void)hipFree(d_cand);
    (void)hipFree(d_countMap);
    (void)hipFree(d_slots);
    (void)hipFree(d_bitmap);
    (void)hipFree(d_intermediate);
    (void)hipFree(d_idx);
    (void)hipFree(d_instanceList);
    (void)hipFree(d_x);
    (void)hipFree(d_y);
    colo_ctx_destroy(colo_compat_ctx());
    colo_host_free(h);
    return 0;
}
