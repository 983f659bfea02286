// host_setup.hpp -- host-side setup of the solve path: device layout of an LDU
// matrix (cell ordering, SELL-32 slices, level schedules), pairwise
// agglomeration, coarse addressing and Galerkin coarse coefficients.
//
// Numbering: "external" ids are the caller's (for coarse levels: the coarse ids
// the reference's agglomeration would assign); "internal" ids are the device
// storage order chosen here for HBM locality.  Every per-row arithmetic
// sequence of the reference is kept: a row lists its faces in ascending face
// id, so for owner-sorted faces the neighbour-side (losort) terms come first,
// then the owner-side terms (src/matrix/lduMatrix/lduMatrixSpMV.cpp:87-91).
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <memory>
#include <string>
#include <thread>
#include <vector>

namespace unap_hip {

// f(begin, end) over [0, n) split among up to 16 host threads: for loops
// whose iterations write disjoint elements, so the result does not depend
// on the split.  Small ranges run on the calling thread.
template <class F>
void parallelFor(long long n, F&& f) {
    const int T = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    if (T == 1 || n < (1 << 16)) {
        f(0LL, n);
        return;
    }
    const long long chunk = (n + T - 1) / T;
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) {
        const long long a = t * chunk, b = std::min(n, a + chunk);
        if (a < b) th.emplace_back([&f, a, b] { f(a, b); });
    }
    for (auto& x : th) x.join();
}

// Items 0..m-1 grouped by key(i) in [0, nb) (key -1: left out), ascending
// inside each group: the result of a sequential counting sort over ascending
// i, built on host threads (counts and placements by atomics, then each
// group sorted).  start gets nb + 1 group boundaries.
template <class Key>
void stableBuckets(long long m, int nb, Key key, std::vector<int>& start, std::vector<int>& items) {
    std::vector<std::atomic<int>> cnt(nb + 1);
    parallelFor(nb + 1, [&](long long a, long long b) {
        for (long long i = a; i < b; ++i) cnt[i].store(0, std::memory_order_relaxed);
    });
    parallelFor(m, [&](long long a, long long b) {
        for (long long i = a; i < b; ++i) {
            const int k = key(i);
            if (k >= 0) cnt[k + 1].fetch_add(1, std::memory_order_relaxed);
        }
    });
    start.assign(nb + 1, 0);
    for (int k = 0; k < nb; ++k) start[k + 1] = start[k] + cnt[k + 1].load(std::memory_order_relaxed);
    parallelFor(nb, [&](long long a, long long b) {
        for (long long k = a; k < b; ++k) cnt[k].store(start[k], std::memory_order_relaxed);
    });
    items.assign(start[nb], 0);
    parallelFor(m, [&](long long a, long long b) {
        for (long long i = a; i < b; ++i) {
            const int k = key(i);
            if (k >= 0) items[cnt[k].fetch_add(1, std::memory_order_relaxed)] = (int)i;
        }
    });
    parallelFor(nb, [&](long long a, long long b) {
        for (long long k = a; k < b; ++k)
            if (start[k + 1] - start[k] > 1) std::sort(items.begin() + start[k], items.begin() + start[k + 1]);
    });
}

// Setup stage times on stderr ("unap_hip setup: <stage> <ms>") when the
// environment variable UNAP_HIP_SETUP_TIMES is set: a diagnostic for where
// host and device setup time goes.
inline bool setupTimesOn() {
    static const bool on = std::getenv("UNAP_HIP_SETUP_TIMES") != nullptr;
    return on;
}
struct StageTimer {
    const char* what;
    long long size;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    StageTimer(const char* w, long long n) : what(w), size(n) {}
    ~StageTimer() {
        if (!setupTimesOn()) return;
        const double ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        std::fprintf(stderr, "unap_hip setup: %s n=%lld %.1f ms\n", what, size, ms);
    }
};

// SELL slice height: slice s holds internal rows 32s .. 32s+31, entry k of a
// row at sliceStart[s] + 32k + (row & 31).  A 64-lane wavefront covers two
// slices; its rows are sorted by length (stably, inside each aligned group of
// 64 rows), so the two slices pad less.  A wavefront's k-th load still
// touches two 128 B lines (FP32) or four (FP64), as with 64-row slices,
// while on agglomerated coarse levels (6..40 entries a row) padded entries
// fall from 1.46-1.51x to 1.22-1.25x; the rows a wavefront gathers for are
// unchanged.  (16-row slices pad 1.10x but touch twice the lines per load:
// the finest FP32 Amul fell from 4.6 to 4.2 TB/s.)
constexpr int kSliceShift = 5;
constexpr int kSliceRows = 1 << kSliceShift;
constexpr int kSortWindow = 64;

// An LDU matrix in external numbering (the lduMatrix data of the reference,
// src/matrix/lduMatrix/lduMatrix.hpp:21-42, with 32-bit addressing).
struct HostLdu {
    int n = 0;
    std::vector<int> l, u;         // lowerAddr (owner), upperAddr (neighbour)
    std::vector<double> lo, up, d; // lower (empty when symmetric), upper, diag
    bool symm = true;
    int nFaces() const { return (int)u.size(); }
    const double* lower() const { return symm ? up.data() : lo.data(); }
    // faces sorted by owner with owner < neighbour: the order the reference's
    // sequential sweeps (GS, DIC, DILU) are defined on
    bool ownerSorted() const;
};

struct LayoutDev;  // device.hpp

// Device storage plan of one matrix.
struct Layout {
    int n = 0, nf = 0, nslices = 0;
    long long nnzPadded = 0;
    std::vector<int> perm;        // internal -> external
    std::vector<int> iperm;       // external -> internal
    std::vector<long long> sliceStart;  // nslices + 1, in entries
    std::vector<uint32_t> rowInfo;      // len | nLowerSide << 16, internal order
    std::vector<int> col;               // padded, internal ids (padding -> own row)
    std::vector<int> src;               // padded: 2*face + side (0 lower, 1 upper), -1 pad
    std::vector<int> fwdRows;           // internal rows by forward level (empty: level-major order)
    std::vector<int> fwdStart;          // forward levels: fwdRows[fwdStart[k] .. fwdStart[k+1]) or,
                                        // with fwdRows empty, internal rows [fwdStart[k], fwdStart[k+1])
    std::vector<int> bwdRows;           // internal rows sorted by backward level
    std::vector<int> bwdStart;          // backward levels: bwdRows[bwdStart[k] .. bwdStart[k+1])
    int maxRowLen = 0;
    bool sweepable = false;             // ownerSorted(): sweeps reproduce the reference
    // layouts built on the device keep col and src in HBM only (the host
    // vectors above then stay empty)
    std::shared_ptr<LayoutDev> dev;
};

// orderMode 0: forward-level major, breadth-first within a level (sweeps
// read contiguous rows); 1: breadth-first only (Amul and Chebyshev stream at
// twice the level-major rate; sweeps gather their rows through fwdRows);
// 2: the caller's order.
// sigma: rows are sorted by length inside windows of sigma rows (default
// kSortWindow, one wavefront; 0: off).
Layout buildLayout(const HostLdu& A, int orderMode, int sigma = -1);
// Order forced for every layout by the environment variable
// UNAP_HIP_CELL_ORDER ("level" 0, "bfs" 1, "natural" 2), or -1 when unset.
int forcedOrderMode();
// Forward / backward sweep level counts (0 when the faces are not owner-sorted).
void levelCounts(const HostLdu& A, int& nFwd, int& nBwd);
// An explicit internal order (perm: internal -> external).
Layout buildLayoutWithOrder(const HostLdu& A, const std::vector<int>& perm);

// Values in the device layout.  useUpperBothSides: DIC reads upper on both
// sweep directions (lduDICPrecond.cpp:68-76) even for asymmetric storage.
template <typename T>
void fillSellValues(const Layout& L, const HostLdu& A, bool useUpperBothSides, std::vector<T>& val);
template <typename T>
void fillDiag(const Layout& L, const HostLdu& A, std::vector<T>& d);

// rD of DIC / DILU in external numbering (lduDICPrecond.cpp:15-41,
// lduDILUPrecond.cpp:22-41): sequential face loop, then reciprocal.
std::vector<double> dicReciprocalD(const HostLdu& A);
std::vector<double> diluReciprocalD(const HostLdu& A);

// One coarse level of the agglomeration hierarchy.
struct CoarseLevel {
    int nCoarse = 0;
    std::vector<int> restrictMap;      // fine external -> coarse external
    std::vector<int> faceRestrictMap;  // fine face -> coarse face, or -(coarse cell + 1)
    HostLdu A;                         // coarse matrix: addressing here; the Galerkin
                                       // values are formed on the device and copied
                                       // back when host code needs them
};

// lduAgglomeration::agglomerate (lduAgglomerate.cpp:4-91) with the pairwise
// step (:94-271) and coarse addressing (lduAgglomerationAddressing.cpp:8-178).
// The Galerkin coefficients (lduAgglomerateMatrix.cpp:3-104) are not formed
// here (galerkin.hpp).
// lists (optional): builds a level's pairing lists -- per cell (row starts
// rs), the cell across each face and the face weight, neighbour side first,
// then owner side, each by face id -- and returns true, or returns false to
// leave them to the host.
using PairLists = std::function<bool(const HostLdu&, const std::vector<double>&, std::vector<int>& rs,
                                     std::vector<int>& nbr, std::vector<double>& wk)>;
// addressing (optional): a level's coarse faces, face map and next-level
// face weights (coarseAddressing), returning true, or false to leave them
// to the host.
using CoarseAddr =
    std::function<bool(const HostLdu&, CoarseLevel&, const std::vector<double>& w, std::vector<double>& cw)>;
struct AgglHooks {
    PairLists lists;
    CoarseAddr addressing;
};
std::vector<CoarseLevel> agglomerate(const HostLdu& fine, const std::vector<double>& weights,
                                     int maxLevels, int nCellsInCoarsestLevel, const AgglHooks* hooks = nullptr);

// Galerkin segment lists of level C on the host: children of each coarse
// cell ascending by fine cell, inner faces ascending, fine faces of each
// coarse face ascending (~f where the sides swap).  fineL: the fine level's
// lowerAddr.
void galerkinListsHost(const CoarseLevel& C, const std::vector<int>& fineL, std::vector<int>& cst,
                       std::vector<int>& cl, std::vector<int>& ist, std::vector<int>& il, std::vector<int>& fst,
                       std::vector<int>& ff);

// Restriction in device order: children of each coarse internal cell as fine
// internal ids, ascending in fine external id (the summation order of
// unapMatrix.hpp:403-406), and the fine-internal -> coarse-internal map.
struct Transfer {
    std::vector<int> childStart, child;  // CSR over coarse internal cells
    std::vector<int> fineToCoarse;       // fine internal -> coarse internal
};
Transfer buildTransfer(const Layout& fineL, const Layout& coarseL, const std::vector<int>& restrictMap);

// The same structures built on the device (setup_dev.hip), element for
// element equal to the host builder's.  buildLayoutDevice returns false
// (leaving the work to buildLayout) for graphs of more than 256 connected
// components.  Matrices of at least deviceSetupMin() cells are built there
// (UNAP_HIP_DEVICE_SETUP_MIN: 0 every matrix, negative none; default 4096).
bool buildLayoutDevice(const HostLdu& A, int orderMode, int sigma, Layout& L);
void pairingListsDevice(const HostLdu& A, const std::vector<double>& w, std::vector<int>& rs, std::vector<int>& nbr,
                        std::vector<double>& wk);
void coarseAddressingDevice(const HostLdu& F, CoarseLevel& C, const std::vector<double>& w, std::vector<double>& cw);
int deviceSetupMin();
// COO -> LDU (formats.cpp) on the device: false where no GPU is present; err
// set when the input is refused.
bool cooToLduDevice(int64_t n, int64_t nnz, const int64_t* rows, const int64_t* cols, const double* vals,
                    bool symmetric, int64_t base, std::vector<int64_t>& l, std::vector<int64_t>& u,
                    std::vector<double>& lo, std::vector<double>& d, std::vector<double>& up, std::string& err);

}  // namespace unap_hip
