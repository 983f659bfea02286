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

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace unap_hip {

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
std::vector<CoarseLevel> agglomerate(const HostLdu& fine, const std::vector<double>& weights,
                                     int maxLevels, int nCellsInCoarsestLevel);

// Restriction in device order: children of each coarse internal cell as fine
// internal ids, ascending in fine external id (the summation order of
// unapMatrix.hpp:403-406), and the fine-internal -> coarse-internal map.
struct Transfer {
    std::vector<int> childStart, child;  // CSR over coarse internal cells
    std::vector<int> fineToCoarse;       // fine internal -> coarse internal
};
Transfer buildTransfer(const Layout& fineL, const Layout& coarseL, const std::vector<int>& restrictMap);

}  // namespace unap_hip
