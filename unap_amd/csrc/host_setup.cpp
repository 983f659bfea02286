// host_setup.cpp -- see host_setup.hpp.
#include "host_setup.hpp"

#include <cstdlib>
#include <cstring>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace unap_hip {

static const double kSmall = 1.0e-37;  // SMALL, src/base/unap.hpp:88

bool HostLdu::ownerSorted() const {
    const int nf = nFaces();
    for (int f = 0; f < nf; ++f) {
        if (!(l[f] < u[f])) return false;
        if (f > 0 && l[f] < l[f - 1]) return false;
    }
    return true;
}

namespace {

// Per-row face lists in ascending face id (a single pass over faces keeps
// that order): side 0 = the row is the neighbour (coefficient lower[f],
// column owner), side 1 = the row is the owner (upper[f], column neighbour).
struct RowFaces {
    std::vector<int> rs, ent, nLowerSide;
};
RowFaces rowFaces(const HostLdu& A) {
    const int n = A.n, nf = A.nFaces();
    RowFaces R;
    R.rs.assign(n + 1, 0);
    R.ent.resize(2 * (size_t)nf);
    R.nLowerSide.assign(n, 0);
    if (nf < (1 << 16)) {
        for (int f = 0; f < nf; ++f) {
            R.rs[A.u[f] + 1]++;
            R.rs[A.l[f] + 1]++;
        }
        for (int c = 0; c < n; ++c) R.rs[c + 1] += R.rs[c];
        std::vector<int> fillp(R.rs.begin(), R.rs.end() - 1);
        for (int f = 0; f < nf; ++f) {
            R.ent[fillp[A.u[f]]++] = 2 * f;
            R.ent[fillp[A.l[f]]++] = 2 * f + 1;
            R.nLowerSide[A.u[f]]++;
        }
        return R;
    }
    // Large matrices on host threads: counts by atomic increments, entries
    // placed through atomic cursors, then each row's short list sorted by
    // face id -- the same lists as the sequential pass above.
    std::vector<std::atomic<int>> cnt(n + 1), low(n);
    parallelFor(n + 1, [&](long long a, long long b) {
        for (long long i = a; i < b; ++i) cnt[i].store(0, std::memory_order_relaxed);
    });
    parallelFor(n, [&](long long a, long long b) {
        for (long long i = a; i < b; ++i) low[i].store(0, std::memory_order_relaxed);
    });
    parallelFor(nf, [&](long long a, long long b) {
        for (long long f = a; f < b; ++f) {
            cnt[A.u[f] + 1].fetch_add(1, std::memory_order_relaxed);
            cnt[A.l[f] + 1].fetch_add(1, std::memory_order_relaxed);
            low[A.u[f]].fetch_add(1, std::memory_order_relaxed);
        }
    });
    for (int c = 0; c < n; ++c) R.rs[c + 1] = R.rs[c] + cnt[c + 1].load(std::memory_order_relaxed);
    parallelFor(n, [&](long long a, long long b) {
        for (long long c = a; c < b; ++c) {
            R.nLowerSide[c] = low[c].load(std::memory_order_relaxed);
            cnt[c].store(R.rs[c], std::memory_order_relaxed);  // cursor
        }
    });
    parallelFor(nf, [&](long long a, long long b) {
        for (long long f = a; f < b; ++f) {
            R.ent[cnt[A.u[f]].fetch_add(1, std::memory_order_relaxed)] = (int)(2 * f);
            R.ent[cnt[A.l[f]].fetch_add(1, std::memory_order_relaxed)] = (int)(2 * f + 1);
        }
    });
    parallelFor(n, [&](long long a, long long b) {
        for (long long c = a; c < b; ++c) std::sort(R.ent.begin() + R.rs[c], R.ent.begin() + R.rs[c + 1]);
    });
    return R;
}

// Breadth-first order over the cell graph: recovers the spatial locality
// that the caller's numbering may have scrambled.  Seeds in ascending cell
// id; a cell's neighbours in its row-face order.  Wide frontiers are expanded
// on host threads in the same order: every unvisited neighbour keeps only its
// first discovery (smallest frontier position, then row-face index), and the
// threads' lists are joined in frontier order.
std::vector<int> bfsOrder(const HostLdu& A, const RowFaces& R) {
    const int n = A.n;
    std::vector<int> bfs;
    bfs.reserve(n);
    std::vector<char> seen(n, 0);
    auto nbr = [&](int k) {
        const int f = R.ent[k] >> 1;
        return (R.ent[k] & 1) ? A.u[f] : A.l[f];
    };
    std::vector<std::atomic<unsigned long long>> claim;
    const unsigned long long kFree = ~0ull;
    constexpr int kWide = 1 << 14;  // frontier size worth the threads
    const int T = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    for (int s = 0; s < n; ++s) {
        if (seen[s]) continue;
        seen[s] = 1;
        bfs.push_back(s);
        size_t fb = bfs.size() - 1, fe = bfs.size();
        while (fb < fe) {
            if (fe - fb < (size_t)kWide || T == 1) {
                for (size_t p = fb; p < fe; ++p) {
                    const int c = bfs[p];
                    for (int k = R.rs[c]; k < R.rs[c + 1]; ++k) {
                        const int nb = nbr(k);
                        if (!seen[nb]) {
                            seen[nb] = 1;
                            bfs.push_back(nb);
                        }
                    }
                }
            } else {
                if (claim.empty()) {
                    claim = std::vector<std::atomic<unsigned long long>>(n);
                    parallelFor(n, [&](long long a, long long b) {
                        for (long long i = a; i < b; ++i) claim[i].store(kFree, std::memory_order_relaxed);
                    });
                }
                const long long m = (long long)(fe - fb);
                const long long chunk = (m + T - 1) / T;
                auto key = [&](size_t p, int k) { return ((unsigned long long)p << 24) | (unsigned)(k - R.rs[bfs[p]]); };
                std::vector<std::thread> th;
                for (int t = 0; t < T; ++t)  // first discoveries
                    th.emplace_back([&, t] {
                        const size_t pa = fb + t * chunk, pb = std::min(fe, (size_t)(fb + (t + 1) * chunk));
                        for (size_t p = pa; p < pb; ++p)
                            for (int k = R.rs[bfs[p]]; k < R.rs[bfs[p] + 1]; ++k) {
                                const int nb = nbr(k);
                                if (seen[nb]) continue;
                                const unsigned long long kk = key(p, k);
                                unsigned long long cur = claim[nb].load(std::memory_order_relaxed);
                                while (kk < cur && !claim[nb].compare_exchange_weak(cur, kk, std::memory_order_relaxed)) {
                                }
                            }
                    });
                for (auto& x : th) x.join();
                th.clear();
                std::vector<std::vector<int>> part(T);
                for (int t = 0; t < T; ++t)  // each thread's discoveries, in order
                    th.emplace_back([&, t] {
                        const size_t pa = fb + t * chunk, pb = std::min(fe, (size_t)(fb + (t + 1) * chunk));
                        for (size_t p = pa; p < pb; ++p)
                            for (int k = R.rs[bfs[p]]; k < R.rs[bfs[p] + 1]; ++k) {
                                const int nb = nbr(k);
                                if (!seen[nb] && claim[nb].load(std::memory_order_relaxed) == key(p, k))
                                    part[t].push_back(nb);
                            }
                    });
                for (auto& x : th) x.join();
                for (int t = 0; t < T; ++t)
                    for (int v : part[t]) {
                        seen[v] = 1;
                        claim[v].store(kFree, std::memory_order_relaxed);
                        bfs.push_back(v);
                    }
            }
            fb = fe;
            fe = bfs.size();
        }
    }
    return bfs;
}

// Rows of one sweep direction grouped by level; `start` gets the boundaries.
std::vector<int> groupByLevel(const std::vector<int>& levOfInternal, int nlev, std::vector<int>& start) {
    const int n = (int)levOfInternal.size();
    start.assign(nlev + 1, 0);
    for (int i = 0; i < n; ++i) start[levOfInternal[i] + 1]++;
    for (int k = 0; k < nlev; ++k) start[k + 1] += start[k];
    std::vector<int> pos(start.begin(), start.end() - 1), rows(n);
    for (int i = 0; i < n; ++i) rows[pos[levOfInternal[i]]++] = i;
    return rows;
}

Layout finishLayout(const HostLdu& A, const RowFaces& R, std::vector<int> perm, const std::vector<int>& lev,
                    int nlev, const std::vector<int>& blev, int nblev, bool sweepable) {
    Layout L;
    const int n = A.n;
    L.n = n;
    L.nf = A.nFaces();
    L.sweepable = sweepable;
    L.perm = std::move(perm);
    L.iperm.assign(n, 0);
    parallelFor(n, [&](long long a, long long b) {
        for (long long i = a; i < b; ++i) L.iperm[L.perm[i]] = (int)i;
    });

    // Sweep schedules over internal rows.  When the internal order is level
    // major the forward rows are contiguous and fwdRows stays empty.
    std::vector<int> li(n), bi(n);
    parallelFor(n, [&](long long a, long long b) {
        for (long long i = a; i < b; ++i) {
            li[i] = lev[L.perm[i]];
            bi[i] = blev[L.perm[i]];
        }
    });
    std::vector<int> fr = groupByLevel(li, nlev, L.fwdStart);
    bool contiguous = true;
    for (int i = 0; i < n && contiguous; ++i) contiguous = (fr[i] == i);
    if (!contiguous) L.fwdRows = std::move(fr);
    L.bwdRows = groupByLevel(bi, nblev, L.bwdStart);

    // SELL-32 (kSliceRows): slice s holds internal rows 32s..32s+31, entry k
    // of a row at sliceStart[s] + 32k + (row & 31).
    const int C = kSliceRows;
    L.nslices = (n + C - 1) / C;
    L.sliceStart.assign(L.nslices + 1, 0);
    L.rowInfo.assign(n, 0);
    std::vector<int> width(L.nslices, 0);
    std::atomic<bool> tooLong{false};
    parallelFor(L.nslices, [&](long long a, long long b) {
        for (long long s = a; s < b; ++s) {
            int mx = 0;
            for (int r = C * (int)s; r < std::min(n, C * (int)s + C); ++r) {
                const int c = L.perm[r];
                const int len = R.rs[c + 1] - R.rs[c];
                if (len > 0xffff) tooLong = true;
                mx = std::max(mx, len);
                L.rowInfo[r] = (uint32_t)len | ((uint32_t)R.nLowerSide[c] << 16);
            }
            width[s] = mx;
        }
    });
    if (tooLong) throw std::runtime_error("row longer than 65535 entries");
    for (int s = 0; s < L.nslices; ++s) {
        L.maxRowLen = std::max(L.maxRowLen, width[s]);
        L.sliceStart[s + 1] = L.sliceStart[s] + (long long)C * width[s];
    }
    L.nnzPadded = L.sliceStart[L.nslices];
    L.col.assign(L.nnzPadded, 0);
    L.src.assign(L.nnzPadded, -1);
    parallelFor(L.nslices, [&](long long sa, long long sb) {
        for (int s = (int)sa; s < (int)sb; ++s) {
            const int mx = width[s];
            for (int lane = 0; lane < C; ++lane) {
                const int r = C * s + lane;
                const int c = r < n ? L.perm[r] : -1;
                const int len = r < n ? R.rs[c + 1] - R.rs[c] : 0;
                for (int k = 0; k < mx; ++k) {
                    const long long p = L.sliceStart[s] + (long long)C * k + lane;
                    if (k < len) {
                        const int e = R.ent[R.rs[c] + k];
                        const int f = e >> 1;
                        L.col[p] = L.iperm[(e & 1) ? A.u[f] : A.l[f]];
                        L.src[p] = e;
                    } else {
                        L.col[p] = r < n ? r : 0;
                        L.src[p] = -1;
                    }
                }
            }
        }
    });
    return L;
}

// Forward level: 1 + max over lower-side neighbours (owners < row); backward
// level: 1 + max over upper-side neighbours.  Only meaningful for owner-sorted
// faces; otherwise every row is level 0.
void levels(const HostLdu& A, const RowFaces& R, bool sweepable, std::vector<int>& lev, int& nlev,
            std::vector<int>& blev, int& nblev) {
    const int n = A.n;
    lev.assign(n, 0);
    blev.assign(n, 0);
    nlev = nblev = 1;
    if (!sweepable) return;
    for (int c = 0; c < n; ++c) {
        int m = 0;
        for (int k = R.rs[c]; k < R.rs[c + 1]; ++k)
            if ((R.ent[k] & 1) == 0) m = std::max(m, lev[A.l[R.ent[k] >> 1]] + 1);
        lev[c] = m;
        nlev = std::max(nlev, m + 1);
    }
    for (int c = n - 1; c >= 0; --c) {
        int m = 0;
        for (int k = R.rs[c]; k < R.rs[c + 1]; ++k)
            if (R.ent[k] & 1) m = std::max(m, blev[A.u[R.ent[k] >> 1]] + 1);
        blev[c] = m;
        nblev = std::max(nblev, m + 1);
    }
}

// SELL-C-sigma: inside consecutive windows of sigma rows of perm[b, e) the
// rows are stably sorted by descending length, so each slice holds rows of
// similar length.  The window is one wavefront (64 rows): wider windows cut
// padding further but move rows between wavefronts and scatter their x
// gathers (on the agglomerated levels of a 256^3 Poisson matrix, 64-row
// slices sorted over 1024-row windows padded 1.04x yet ran the FP32 Amul over
// levels 1-8 in 0.80 ms against 0.55 ms unsorted).
constexpr int kSigma = kSortWindow;
void sortWindows(std::vector<int>& perm, const RowFaces& R, int b, int e, int sigma) {
    if (sigma <= 1) return;
    auto len = [&](int c) { return R.rs[c + 1] - R.rs[c]; };
    for (int s = b; s < e; s += sigma) {
        // stable insertion sort by descending length (the window is short and
        // its lengths mostly equal, so this is near linear)
        const int t = std::min(e, s + sigma);
        for (int i = s + 1; i < t; ++i) {
            const int v = perm[i], lv = len(v);
            int j = i;
            for (; j > s && len(perm[j - 1]) < lv; --j) perm[j] = perm[j - 1];
            perm[j] = v;
        }
    }
}

}  // namespace

int forcedOrderMode() {
    static const int mode = [] {
        const char* e = std::getenv("UNAP_HIP_CELL_ORDER");
        if (!e) return -1;
        if (std::strcmp(e, "level") == 0) return 0;
        if (std::strcmp(e, "bfs") == 0) return 1;
        if (std::strcmp(e, "natural") == 0) return 2;
        return -1;
    }();
    return mode;
}

void levelCounts(const HostLdu& A, int& nFwd, int& nBwd) {
    nFwd = nBwd = 0;
    if (!A.ownerSorted()) return;
    const RowFaces R = rowFaces(A);
    std::vector<int> lev, blev;
    levels(A, R, true, lev, nFwd, blev, nBwd);
}

Layout buildLayout(const HostLdu& A, int orderMode, int sigma) {
    if (sigma < 0) sigma = kSigma;
    const int forced = forcedOrderMode();
    if (forced >= 0) orderMode = forced;
    StageTimer all("layout (all)", A.n);
    RowFaces R;
    {
        StageTimer t("layout: row faces", A.n);
        R = rowFaces(A);
    }
    const bool sweepable = A.ownerSorted();
    std::vector<int> lev, blev;
    int nlev, nblev;
    {
        StageTimer t("layout: sweep levels", A.n);
        levels(A, R, sweepable, lev, nlev, blev, nblev);
    }
    std::vector<int> perm(A.n);
    if (orderMode == 2) {
        std::iota(perm.begin(), perm.end(), 0);  // the caller's order
    } else {
        StageTimer t("layout: breadth-first order + windows", A.n);
        const std::vector<int> bfs = bfsOrder(A, R);
        if (orderMode == 1) {
            perm = bfs;
            sortWindows(perm, R, 0, A.n, sigma);
        } else {  // forward level major, breadth-first minor
            std::vector<int> start(nlev + 1, 0);
            for (int c = 0; c < A.n; ++c) start[lev[c] + 1]++;
            for (int k = 0; k < nlev; ++k) start[k + 1] += start[k];
            std::vector<int> at(start.begin(), start.end() - 1);
            for (int i = 0; i < A.n; ++i) perm[at[lev[bfs[i]]]++] = bfs[i];
            for (int k = 0; k < nlev; ++k) sortWindows(perm, R, start[k], start[k + 1], sigma);
        }
    }
    StageTimer t("layout: schedules + SELL assembly", A.n);
    return finishLayout(A, R, std::move(perm), lev, nlev, blev, nblev, sweepable);
}

Layout buildLayoutWithOrder(const HostLdu& A, const std::vector<int>& perm) {
    const RowFaces R = rowFaces(A);
    const bool sweepable = A.ownerSorted();
    std::vector<int> lev, blev;
    int nlev, nblev;
    levels(A, R, sweepable, lev, nlev, blev, nblev);
    return finishLayout(A, R, perm, lev, nlev, blev, nblev, sweepable);
}

template <typename T>
void fillSellValues(const Layout& L, const HostLdu& A, bool useUpperBothSides, std::vector<T>& val) {
    val.assign(L.nnzPadded, T(0));
    const double* lw = useUpperBothSides ? A.up.data() : A.lower();
    parallelFor(L.nnzPadded, [&](long long a, long long b) {
        for (long long p = a; p < b; ++p) {
            const int e = L.src[p];
            if (e < 0) continue;
            const int f = e >> 1;
            val[p] = (T)((e & 1) ? A.up[f] : lw[f]);
        }
    });
}

template <typename T>
void fillDiag(const Layout& L, const HostLdu& A, std::vector<T>& d) {
    d.resize(L.n);
    for (int i = 0; i < L.n; ++i) d[i] = (T)A.d[L.perm[i]];
}

template void fillSellValues<double>(const Layout&, const HostLdu&, bool, std::vector<double>&);
template void fillSellValues<float>(const Layout&, const HostLdu&, bool, std::vector<float>&);
template void fillDiag<double>(const Layout&, const HostLdu&, std::vector<double>&);
template void fillDiag<float>(const Layout&, const HostLdu&, std::vector<float>&);

std::vector<double> dicReciprocalD(const HostLdu& A) {
    std::vector<double> rD(A.d);
    const int nf = A.nFaces();
    for (int f = 0; f < nf; ++f) rD[A.u[f]] -= A.up[f] * A.up[f] / rD[A.l[f]];
    for (int i = 0; i < A.n; ++i) rD[i] = 1.0 / rD[i];
    return rD;
}

std::vector<double> diluReciprocalD(const HostLdu& A) {
    std::vector<double> rD(A.d);
    const double* lw = A.lower();
    const int nf = A.nFaces();
    for (int f = 0; f < nf; ++f) rD[A.u[f]] -= A.up[f] * lw[f] / rD[A.l[f]];
    for (int i = 0; i < A.n; ++i) rD[i] = 1.0 / rD[i];
    return rD;
}

namespace {

// Pairwise greedy clustering of one level (lduAgglomerate.cpp:94-271, cells
// visited in ascending order).  Each cell scans its faces as the neighbour
// first, then as the owner, each in ascending face id (the cellFaces order
// of lduAgglomerate.cpp:114-152).  The lists are built on host threads and
// hold, per entry, the cell across the face and the face weight, so the
// sequential greedy pass reads them contiguously and looks up only the
// cluster of each neighbour (prefetched a few cells ahead).  Of the face's
// two cells one is the visited cell itself, whose cluster is still unset:
// the reference's tests on both ends reduce to the test on the neighbour.
std::vector<int> pairCells(const HostLdu& A, const std::vector<double>& w, int& nCoarse, const PairLists& lists) {
    const int n = A.n;
    std::vector<int> rs, nbr;
    std::vector<double> wk;
    {
        StageTimer tl("agglomeration: pairing lists", n);
        if (!(lists && lists(A, w, rs, nbr, wk))) {
            RowFaces R = rowFaces(A);  // per cell, by face id: 2f (neighbour side), 2f + 1 (owner side)
            rs = std::move(R.rs);
            nbr.resize(R.ent.size());
            wk.resize(R.ent.size());
            parallelFor(n, [&](long long a, long long b) {
                for (long long c = a; c < b; ++c) {
                    int k = rs[c];
                    for (int side = 0; side < 2; ++side)
                        for (int j = rs[c]; j < rs[c + 1]; ++j) {
                            const int e = R.ent[j];
                            if ((e & 1) != side) continue;
                            const int f = e >> 1;
                            nbr[k] = side ? A.u[f] : A.l[f];
                            wk[k] = w[f];
                            ++k;
                        }
                }
            });
        }
    }
    StageTimer tg("agglomeration: pairing greedy pass", n);
    std::vector<int> map(n, -1);
    nCoarse = 0;
    constexpr int kAhead = 8;
    for (int c = 0; c < n; ++c) {
        if (c + kAhead < n)
            for (int k = rs[c + kAhead]; k < rs[c + kAhead + 1]; ++k) __builtin_prefetch(&map[nbr[k]]);
        if (map[c] >= 0) continue;
        int best = -1;
        double bestW = kSmall;
        for (int k = rs[c]; k < rs[c + 1]; ++k)
            if (map[nbr[k]] < 0 && wk[k] > bestW) {
                best = k;
                bestW = wk[k];
            }
        if (best >= 0) {
            map[nbr[best]] = nCoarse;
            map[c] = nCoarse;
            ++nCoarse;
            continue;
        }
        // no free neighbour: join the cluster across the heaviest face
        int join = -1;
        double joinW = kSmall;
        for (int k = rs[c]; k < rs[c + 1]; ++k)
            if (wk[k] > joinW) {
                join = k;
                joinW = wk[k];
            }
        if (join >= 0) map[c] = std::max(map[c], map[nbr[join]]);
    }
    for (int c = 0; c < n; ++c)
        if (map[c] < 0) map[c] = nCoarse++;
    return map;
}

// Coarse faces from fine faces joining different clusters, numbered owner-major
// with neighbours in first-appearance order (lduAgglomerationAddressing.cpp:57-178).
// The fine faces are grouped by coarse owner, ascending inside each group;
// within a group a coarse neighbour met for the first time gets the group's
// next coarse face id, which is its owner-major, first-appearance position.
// Groups are independent, so they are numbered on host threads once their
// sizes are known.  The next level's face weights (restrictFaceField,
// unapMatrix.hpp:441-475) are summed per group in ascending fine face order,
// the order of the reference's loop over the fine faces.
void coarseAddressing(const HostLdu& F, CoarseLevel& C, const std::vector<double>& w, std::vector<double>& cw) {
    const int nf = F.nFaces();
    const int nc = C.nCoarse;
    C.faceRestrictMap.assign(nf, 0);
    std::vector<int> ownOf(nf), neiOf(nf);
    parallelFor(nf, [&](long long fa, long long fb) {
        for (long long f = fa; f < fb; ++f) {
            const int a = C.restrictMap[F.u[f]], b = C.restrictMap[F.l[f]];
            if (a == b) {
                C.faceRestrictMap[f] = -(a + 1);
                ownOf[f] = -1;
            } else {
                ownOf[f] = std::min(a, b);
                neiOf[f] = std::max(a, b);
            }
        }
    });
    std::vector<int> start, bucket;
    stableBuckets(nf, nc, [&](long long f) { return ownOf[f]; }, start, bucket);
    // first appearance of each coarse neighbour inside a group
    auto firstSeen = [&](int k, int own) {
        const int nei = neiOf[bucket[k]];
        for (int j = start[own]; j < k; ++j)
            if (neiOf[bucket[j]] == nei) return j;
        return k;
    };
    std::vector<int> first(bucket.size()), nNew(nc);
    parallelFor(nc, [&](long long a, long long b) {
        for (long long own = a; own < b; ++own) {
            int m = 0;
            for (int k = start[own]; k < start[own + 1]; ++k) {
                first[k] = firstSeen(k, (int)own);
                m += first[k] == k;
            }
            nNew[own] = m;
        }
    });
    std::vector<int> faceStart(nc + 1, 0);
    for (int own = 0; own < nc; ++own) faceStart[own + 1] = faceStart[own] + nNew[own];
    const int ncf = faceStart[nc];
    C.A.n = nc;
    C.A.l.assign(ncf, 0);
    C.A.u.assign(ncf, 0);
    cw.assign(ncf, 0.0);
    std::vector<int> slotAt(bucket.size());
    parallelFor(nc, [&](long long a, long long b) {
        for (long long own = a; own < b; ++own) {
            int next = faceStart[own];
            for (int k = start[own]; k < start[own + 1]; ++k) {
                const int f = bucket[k];
                int slot;
                if (first[k] == k) {
                    slot = next++;
                    C.A.l[slot] = (int)own;
                    C.A.u[slot] = neiOf[f];
                } else {
                    slot = slotAt[first[k]];
                }
                slotAt[k] = slot;
                C.faceRestrictMap[f] = slot;
                cw[slot] += w[f];
            }
        }
    });
}

}  // namespace

std::vector<CoarseLevel> agglomerate(const HostLdu& fine, const std::vector<double>& weights,
                                     int maxLevels, int nCellsInCoarsestLevel, const AgglHooks* hooks) {
    std::vector<CoarseLevel> levels;
    std::vector<double> w(weights);
    while ((int)levels.size() < maxLevels - 1) {
        const HostLdu& F = levels.empty() ? fine : levels.back().A;
        CoarseLevel C;
        {
            StageTimer t("agglomeration: pairing", F.n);
            C.restrictMap = pairCells(F, w, C.nCoarse, hooks ? hooks->lists : PairLists());
        }
        if (!(C.nCoarse >= nCellsInCoarsestLevel)) break;
        std::vector<double> cw;
        {
            StageTimer t("agglomeration: coarse addressing, face weights", F.n);
            if (!(hooks && hooks->addressing && hooks->addressing(F, C, w, cw))) coarseAddressing(F, C, w, cw);
        }
        w.swap(cw);
        levels.push_back(std::move(C));
    }
    // coefficients: on the device (galerkin.hpp), from the fine values of the time
    for (auto& C : levels) C.A.symm = fine.symm;
    return levels;
}

void galerkinListsHost(const CoarseLevel& C, const std::vector<int>& fl, std::vector<int>& cst, std::vector<int>& cl,
                       std::vector<int>& ist, std::vector<int>& il, std::vector<int>& fst, std::vector<int>& ff) {
    const int nfc = (int)C.restrictMap.size(), nff = (int)C.faceRestrictMap.size();
    const int nc = C.nCoarse, ncf = C.A.nFaces();
    stableBuckets(nfc, nc, [&](long long i) { return C.restrictMap[i]; }, cst, cl);
    stableBuckets(nff, nc, [&](long long f) {
        const int cf = C.faceRestrictMap[f];
        return cf < 0 ? -1 - cf : -1;
    }, ist, il);
    stableBuckets(nff, ncf, [&](long long f) { return C.faceRestrictMap[f]; }, fst, ff);
    // orientation (lduAgglomerateMatrix.cpp:50-66): the fine owner lies in
    // the coarse owner, or the sides swap (~f)
    std::atomic<bool> bad{false};
    parallelFor(ncf, [&](long long a, long long b) {
        for (long long cf = a; cf < b; ++cf)
            for (int k = fst[cf]; k < fst[cf + 1]; ++k) {
                const int f = ff[k];
                const int ownerCell = C.restrictMap[fl[f]];
                if (C.A.l[cf] == ownerCell) continue;
                if (C.A.u[cf] == ownerCell)
                    ff[k] = ~f;
                else
                    bad = true;
            }
    });
    if (bad) throw std::runtime_error("agglomeration: inconsistent coarse addressing");
}

Transfer buildTransfer(const Layout& fineL, const Layout& coarseL, const std::vector<int>& rm) {
    Transfer T;
    const int nf = fineL.n, nc = coarseL.n;
    // children of each coarse internal cell, ascending by fine external id,
    // as fine internal ids
    stableBuckets(nf, nc, [&](long long e) { return coarseL.iperm[rm[e]]; }, T.childStart, T.child);
    parallelFor(nf, [&](long long a, long long b) {
        for (long long k = a; k < b; ++k) T.child[k] = fineL.iperm[T.child[k]];
    });
    T.fineToCoarse.assign(nf, 0);
    parallelFor(nf, [&](long long a, long long b) {
        for (long long i = a; i < b; ++i) T.fineToCoarse[i] = coarseL.iperm[rm[fineL.perm[i]]];
    });
    return T;
}

}  // namespace unap_hip
