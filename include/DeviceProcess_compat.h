/*
 * DeviceProcess_compat.h -- header-only C++ shim that keeps the names and
 * roles of the reference's device entry points (Codes/DeviceProcess.h:4-36)
 * on top of the C ABI in colo.h, so a colocationFinder-style driver
 * (Codes/colocationFinder.cpp) maps 1:1.  Include it instead of
 * DeviceProcess.h and link libcolo.so.
 *
 * Differences a caller sees:
 *  - The dataset, grid and binning live in one colo_ctx (colo_compat_ctx())
 *    instead of __constant__ globals; CountCellFeatures creates it.
 *  - Participation bitmaps use colo.h's 1-bit layout; a buffer sized as the
 *    reference sizes d_bitmap (8 x the largest feature, in words) is large
 *    enough.  getParticipationIndex reads that layout.
 *  - The write passes derive their row offsets from the slot counts
 *    (64-bit); the int32 `indexes` argument is accepted and ignored, and
 *    scalebyConstant does nothing.
 *  - The filters leave the pass flag of candidate c in countMap[c*degree]
 *    as the reference does (K:570), with the same rule at degree 2.
 *  - calcInstanceTableGeneral takes the members' features from the first
 *    row of table1 (it returns ST_CUDAERR for an empty table, as the
 *    reference does when there is nothing to launch).
 * Statuses: 0 on success, ST_CUDAERR (5) otherwise; nothing exits.
 */
#ifndef COLO_DEVICEPROCESS_COMPAT_H
#define COLO_DEVICEPROCESS_COMPAT_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "colo.h"

namespace colo_compat {

typedef unsigned int Integer;  // Codes/DataTypes.h:5
typedef int Index64;           // Codes/DataTypes.h:9
typedef int SInteger;          // Codes/DataTypes.h:10
typedef float Real;            // Codes/DataTypes.h:11
enum StatusType { ST_NOERR, ST_FILEERR, ST_PARAMERR, ST_CALCERR, ST_MEMERR, ST_CUDAERR, ST_ETCERR };

struct scalar2 { Real x, y; };
struct Integer2 { Integer x, y; };
struct cBox { scalar2 minBound, maxBound; };
struct gridBox {  // Codes/DataTypes.h:77-83
    cBox computeZone;
    Integer2 gridSize;
    scalar2 offsets;
    Integer totalCells;
    Real cellWidth;
};

inline colo_ctx *&colo_compat_ctx() {
    static colo_ctx *ctx = nullptr;
    return ctx;
}
inline std::vector<Integer> &colo_compat_fstart() {
    static std::vector<Integer> v;
    return v;
}

inline StatusType st(int rc) { return rc == COLO_OK ? ST_NOERR : ST_CUDAERR; }

// feature whose first instance is `start` (the reference passes ranges)
inline Integer feature_at(size_t start) {
    const std::vector<Integer> &fs = colo_compat_fstart();
    for (size_t f = 0; f < fs.size(); f++)
        if (fs[f] == start) return (Integer)f;
    return 0xFFFFFFFFu;
}

// CountCellFeatures (DeviceProcess.h:33-34): bins the dataset.  The device
// arrays are read back once to build the context; the four out-pointers are
// left untouched (the context owns the binning; colo_binning_export copies
// it).  gridStructure.cellWidth is the neighbour distance r.
inline void CountCellFeatures(gridBox g, size_t fNum, size_t n, Integer *d_instanceList, Real *d_x, Real *d_y,
                              Integer **, Integer **, SInteger **, SInteger **) {
    std::vector<Integer> fid(n);
    std::vector<Real> x(n), y(n);
    if (hipMemcpy(fid.data(), d_instanceList, n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(x.data(), d_x, n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(y.data(), d_y, n * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return;
    std::vector<uint32_t> fcount(fNum, 0);
    for (size_t i = 0; i < n; i++) fcount[fid[i]]++;
    std::vector<Integer> &fs = colo_compat_fstart();
    fs.assign(fNum, 0);
    for (size_t f = 1; f < fNum; f++) fs[f] = fs[f - 1] + fcount[f - 1];
    colo_dataset ds{};
    ds.n = (uint32_t)n;
    ds.F = (uint32_t)fNum;
    ds.x = x.data();
    ds.y = y.data();
    ds.fcount = fcount.data();
    ds.minX = g.computeZone.minBound.x;
    ds.minY = g.computeZone.minBound.y;
    ds.maxX = g.computeZone.maxBound.x;
    // setGrid stores maxY + r in the box (cpp:108); the context adds r itself
    ds.maxY = g.computeZone.maxBound.y - g.cellWidth;
    if (colo_compat_ctx()) colo_ctx_destroy(colo_compat_ctx());
    colo_compat_ctx() = nullptr;
    colo_ctx_create(&ds, g.cellWidth, nullptr, &colo_compat_ctx());
}

// initializeDeviceMemConst (DeviceProcess.h:11): the context already holds
// the dataset; nothing to upload.
template <typename... A>
inline StatusType initializeDeviceMemConst(A...) { return colo_compat_ctx() ? ST_NOERR : ST_CUDAERR; }

// filter2 / filter2_2 (DeviceProcess.h:18-21): countMap[c*degree] = pass,
// written on the device.
inline void filter2_any(Integer *d_coloc, Integer C, Integer degree, Real theta, Integer *countMap) {
    colo_ctx *c = colo_compat_ctx();
    if (!c || !C) return;
    (void)colo_filter(c, d_coloc, C, degree, theta, countMap, nullptr, nullptr);
    (void)hipStreamSynchronize(nullptr);
}
inline void filter2(int *, SInteger *, gridBox, Integer, Integer *, Real PIthreshold, Integer *d_coloc,
                    Integer candiColocCounter, Integer degree, Integer *countMap) {
    filter2_any(d_coloc, candiColocCounter, degree, PIthreshold, countMap);
}
inline void filter2_2(int *, SInteger *, gridBox, Integer, Integer *, Real PIthreshold, Integer *d_coloc,
                      Integer candiColocCounter, Integer degree, Integer *countMap) {
    filter2_any(d_coloc, candiColocCounter, degree, PIthreshold, countMap);
}

// degree2InstanceTableSlotCounter (DeviceProcess.h:5)
inline StatusType degree2InstanceTableSlotCounter(size_t, Integer *slots, size_t start, size_t, size_t secondstart,
                                                  Integer *d_bitmap, Integer) {
    return st(colo_join2_count(colo_compat_ctx(), feature_at(start), feature_at(secondstart), slots, d_bitmap,
                               nullptr, nullptr));
}

// degree2InstanceTableGenerator (DeviceProcess.h:6)
inline StatusType degree2InstanceTableGenerator(size_t maxCount, Integer *slots, Index64 *, Integer *d_intermediate,
                                                size_t start, size_t, Integer combiningFeatureID) {
    colo_ctx *c = colo_compat_ctx();
    uint64_t *offs = nullptr;
    if (hipMalloc(&offs, (maxCount + 1) * 8) != hipSuccess) return ST_CUDAERR;
    int rc = colo_scan_slots(c, slots, maxCount, offs, nullptr, nullptr);
    if (!rc) rc = colo_join2_write(c, feature_at(start), combiningFeatureID, slots, offs, d_intermediate, nullptr,
                                   nullptr);
    (void)hipStreamSynchronize(nullptr);
    (void)hipFree(offs);
    return st(rc);
}

// generalInstanceTableSlotCounter (DeviceProcess.h:8)
inline StatusType generalInstanceTableSlotCounter(size_t R, Integer *slots, Integer degree, Integer *table1,
                                                  Integer *d_bitmap, Integer *coloc, Integer) {
    std::vector<Integer> cand(degree);
    if (hipMemcpy(cand.data(), coloc, degree * 4, hipMemcpyDeviceToHost) != hipSuccess) return ST_CUDAERR;
    return st(colo_joink_count(colo_compat_ctx(), table1, R, degree, cand.data(), slots, d_bitmap, nullptr,
                               nullptr));
}

// calcInstanceTableGeneral (DeviceProcess.h:9): the reference passes only
// the last feature; the features of the other members are those of the
// instances in table1's first row (instance ids are grouped by feature).
inline Integer feature_of_instance(Integer id) {
    const std::vector<Integer> &fs = colo_compat_fstart();
    size_t lo = 0, hi = fs.size();
    while (hi - lo > 1) {
        const size_t mid = (lo + hi) / 2;
        if (fs[mid] <= id) lo = mid; else hi = mid;
    }
    return (Integer)lo;
}
inline StatusType calcInstanceTableGeneral(size_t R, Integer *slots, Index64 *, Integer *d_intermediate,
                                           Integer degree, Integer *table1, size_t, Integer lastFeatureID) {
    colo_ctx *c = colo_compat_ctx();
    if (!c || R == 0 || degree < 3) return ST_CUDAERR;  // nothing to launch (cu:94-105)
    std::vector<Integer> cand(degree);
    if (hipMemcpy(cand.data(), table1, (degree - 1) * 4, hipMemcpyDeviceToHost) != hipSuccess) return ST_CUDAERR;
    for (Integer m = 0; m + 1 < degree; m++) cand[m] = feature_of_instance(cand[m]);
    cand[degree - 1] = lastFeatureID;
    uint64_t *offs = nullptr;
    if (hipMalloc(&offs, (R + 1) * 8) != hipSuccess) return ST_CUDAERR;
    int rc = colo_scan_slots(c, slots, R, offs, nullptr, nullptr);
    if (!rc) rc = colo_joink_write(c, table1, R, degree, cand.data(), slots, offs, d_intermediate, nullptr);
    (void)hipStreamSynchronize(nullptr);
    (void)hipFree(offs);
    return st(rc);
}

// getParticipationIndex (DeviceProcess.h:12): PI with the reference's
// numerators (cu:108-132).
inline Real getParticipationIndex(Integer *&d_bitmap, Integer degree, size_t i, std::vector<Integer> cand, Integer *&,
                                  Integer *&, Integer *&) {
    float pi = 0.f;
    colo_pi(colo_compat_ctx(), d_bitmap, degree, cand.data() + i * degree, nullptr, nullptr, &pi, nullptr, nullptr);
    return pi;
}

// getTotalInstances (DeviceProcess.h:13): host sum of slot counts (cu:154-159).
inline Integer getTotalInstances(Integer *slots, size_t n) {
    Integer t = 0;
    for (size_t i = 0; i < n; i++) t += slots[i];
    return t;
}

// scalebyConstant (DeviceProcess.h:15): the write passes scale by the degree.
inline StatusType scalebyConstant(Integer, Index64 *, size_t) { return ST_NOERR; }

}  // namespace colo_compat

#endif
