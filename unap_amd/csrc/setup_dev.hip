// setup_dev.hip -- setup structures built on the device: the per-cell face
// lists (the owner-start / losort views of the reference's lduMatrix,
// lduMatrix.cpp:245-340), the pairing lists of the agglomeration, the
// breadth-first cell order, the sweep level schedules and the SELL-32
// assembly of a matrix's device layout.
//
// Every result equals, element for element, the host builder's
// (host_setup.cpp) for the same matrix: the same per-row face order, the same
// breadth-first visit (a cell joins the frontier of its first discovery:
// smallest frontier position, then face-list position), the same levels (the
// longest-path fixed point), the same stable groupings and window sorts.
// Sorting and scanning run through rocPRIM; the remaining work is one thread
// per cell, face or frontier cell.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_scan.hpp>

#include "device.hpp"
#include "host_setup.hpp"
#include "unap_hip.h"

namespace unap_hip {
namespace {

constexpr int kT = 256;
inline unsigned blocksFor(long long n) { return (unsigned)std::max<long long>(1, (n + kT - 1) / kT); }
inline unsigned bitsFor(long long n) {
    unsigned b = 1;
    while ((1LL << b) < n) ++b;
    return b;
}

// rocPRIM temporary storage, grown on demand
struct Temp {
    DBuf<char> b;
    void* get(size_t need) {
        if (b.n < need) b.alloc(std::max<size_t>(need, 1));
        return b.get();
    }
};

template <class T>
std::vector<T> download(const T* d, size_t n, hipStream_t s) {
    std::vector<T> h(n);
    if (n) UH_CHECK(hipMemcpyAsync(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost, s));
    UH_CHECK(hipStreamSynchronize(s));
    return h;
}
template <class T>
T readOne(const T* d, hipStream_t s) {
    T v;
    UH_CHECK(hipMemcpyAsync(&v, d, sizeof(T), hipMemcpyDeviceToHost, s));
    UH_CHECK(hipStreamSynchronize(s));
    return v;
}

void sortKeys(Temp& t, uint64_t* in, uint64_t* out, long long m, unsigned endBit, hipStream_t s) {
    size_t bytes = 0;
    UH_CHECK(rocprim::radix_sort_keys(nullptr, bytes, in, out, (size_t)m, 0u, endBit, s));
    UH_CHECK(rocprim::radix_sort_keys(t.get(bytes), bytes, in, out, (size_t)m, 0u, endBit, s));
}
// stable: equal keys keep their input order
void sortPairs(Temp& t, int* kin, int* kout, int* vin, int* vout, long long m, unsigned endBit, hipStream_t s) {
    size_t bytes = 0;
    UH_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, (size_t)m, 0u, endBit, s));
    UH_CHECK(rocprim::radix_sort_pairs(t.get(bytes), bytes, kin, kout, vin, vout, (size_t)m, 0u, endBit, s));
}
template <class T>
void exclusiveSum(Temp& t, const T* in, T* out, long long m, hipStream_t s) {
    size_t bytes = 0;
    UH_CHECK(rocprim::exclusive_scan(nullptr, bytes, in, out, T(0), (size_t)m, rocprim::plus<T>(), s));
    UH_CHECK(rocprim::exclusive_scan(t.get(bytes), bytes, in, out, T(0), (size_t)m, rocprim::plus<T>(), s));
}
int maxOf(Temp& t, const int* in, long long m, hipStream_t s) {
    DBuf<int> r(1);
    size_t bytes = 0;
    UH_CHECK(rocprim::reduce(nullptr, bytes, in, r.get(), 0, (size_t)m, rocprim::maximum<int>(), s));
    UH_CHECK(rocprim::reduce(t.get(bytes), bytes, in, r.get(), 0, (size_t)m, rocprim::maximum<int>(), s));
    return readOne(r.get(), s);
}

// ------------------------------------------------------------ row faces
// Two keys a face, (cell << 32) | code.  Layout order: code 2f (the cell is
// the neighbour) or 2f + 1 (the owner), so a row's entries come by face id.
// Pairing order: code (side << 31) | f, the neighbour side first, then the
// owner side, each by face id (lduAgglomerate.cpp:114-152).
__global__ void kFaceKeys(int nf, const int* l, const int* u, int pairing, uint64_t* keys) {
    const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const uint64_t cu = (uint64_t)(uint32_t)u[f], cl = (uint64_t)(uint32_t)l[f];
    if (pairing) {
        keys[2 * f] = (cu << 32) | (uint64_t)f;
        keys[2 * f + 1] = (cl << 32) | (1ull << 31) | (uint64_t)f;
    } else {
        keys[2 * f] = (cu << 32) | (uint64_t)(2 * f);
        keys[2 * f + 1] = (cl << 32) | (uint64_t)(2 * f + 1);
    }
}
// row starts from the sorted keys (empty rows included) and the codes
__global__ void kRowStarts(long long m, int n, const uint64_t* keys, int* rs, int* code) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > m) return;
    const int cell = k < m ? (int)(keys[k] >> 32) : n;
    const int prev = k > 0 ? (int)(keys[k - 1] >> 32) : -1;
    for (int c = prev + 1; c <= cell; ++c) rs[c] = (int)k;
    if (k < m) code[k] = (int)(uint32_t)keys[k];
}

struct DevFaces {
    int n = 0, nf = 0;
    DBuf<int> l, u, rs, ent;
};
void rowFacesDev(const HostLdu& A, bool pairing, DevFaces& D, Temp& t, hipStream_t s) {
    D.n = A.n;
    D.nf = A.nFaces();
    D.l.upload(A.l, s);
    D.u.upload(A.u, s);
    const long long m = 2LL * D.nf;
    DBuf<uint64_t> k0(std::max<long long>(m, 1)), k1(std::max<long long>(m, 1));
    if (D.nf > 0) {
        kFaceKeys<<<blocksFor(D.nf), kT, 0, s>>>(D.nf, D.l.get(), D.u.get(), pairing ? 1 : 0, k0.get());
        UH_LAUNCHED();
        sortKeys(t, k0.get(), k1.get(), m, 32 + bitsFor(D.n), s);
    }
    D.rs.alloc(D.n + 1);
    D.ent.alloc(std::max<long long>(m, 1));
    kRowStarts<<<blocksFor(m + 1), kT, 0, s>>>(m, D.n, k1.get(), D.rs.get(), D.ent.get());
    UH_LAUNCHED();
}

// ------------------------------------------------------- pairing lists
__global__ void kPairLists(long long m, const int* code, const int* l, const int* u, const double* w, int* nbr,
                           double* wk) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const uint32_t c = (uint32_t)code[k];
    const int f = (int)(c & 0x7fffffffu);
    nbr[k] = (c >> 31) ? u[f] : l[f];
    wk[k] = w[f];
}

// ------------------------------------------------------------ levels
// One relaxation pass towards the longest-path levels: forward over the
// neighbour-side entries (the owner precedes), backward over the owner-side
// ones.  Values only grow and never pass the true level, so repeating until
// nothing changes ends at the host's sequential result.
__global__ void kLevelPass(int n, const int* rs, const int* ent, const int* l, const int* u, int forward, int* lev,
                           int* changed) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    int m = 0;
    for (int k = rs[c]; k < rs[c + 1]; ++k) {
        const int e = ent[k];
        if (forward && (e & 1) == 0) m = max(m, lev[l[e >> 1]] + 1);
        if (!forward && (e & 1)) m = max(m, lev[u[e >> 1]] + 1);
    }
    if (m > lev[c]) {
        lev[c] = m;
        *changed = 1;
    }
}
int levelsDev(const DevFaces& D, bool forward, DBuf<int>& lev, Temp& t, hipStream_t s) {
    lev.alloc(std::max(D.n, 1));
    UH_CHECK(hipMemsetAsync(lev.get(), 0, sizeof(int) * D.n, s));
    DBuf<int> changed(1);
    for (;;) {
        UH_CHECK(hipMemsetAsync(changed.get(), 0, sizeof(int), s));
        kLevelPass<<<blocksFor(D.n), kT, 0, s>>>(D.n, D.rs.get(), D.ent.get(), D.l.get(), D.u.get(), forward ? 1 : 0,
                                                 lev.get(), changed.get());
        UH_LAUNCHED();
        if (!readOne(changed.get(), s)) break;
    }
    return maxOf(t, lev.get(), D.n, s) + 1;
}

// ----------------------------------------------------- breadth-first order
// Frontier cells p in [fb, fe) of bfs claim their unseen neighbours with the
// key (p << 24) | (position in the row); the smallest key wins, which is the
// first discovery of the sequential visit.  Claims left on cells that were
// placed are never read again (placed cells are seen).
__device__ inline int nbrOf(int e, const int* l, const int* u) { return (e & 1) ? u[e >> 1] : l[e >> 1]; }
__global__ void kClaim(int fb, int fe, const int* bfs, const int* rs, const int* ent, const int* l, const int* u,
                       const char* seen, unsigned long long* claim) {
    const int p = fb + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= fe) return;
    const int c = bfs[p];
    for (int k = rs[c]; k < rs[c + 1]; ++k) {
        const int nb = nbrOf(ent[k], l, u);
        if (!seen[nb]) atomicMin(&claim[nb], ((unsigned long long)p << 24) | (unsigned)(k - rs[c]));
    }
}
__global__ void kWins(int fb, int fe, const int* bfs, const int* rs, const int* ent, const int* l, const int* u,
                      const char* seen, const unsigned long long* claim, int* cnt) {
    const int p = fb + blockIdx.x * blockDim.x + threadIdx.x;
    if (p > fe) return;
    if (p == fe) {
        cnt[p - fb] = 0;
        return;
    }
    const int c = bfs[p];
    int w = 0;
    for (int k = rs[c]; k < rs[c + 1]; ++k) {
        const int nb = nbrOf(ent[k], l, u);
        w += (!seen[nb] && claim[nb] == (((unsigned long long)p << 24) | (unsigned)(k - rs[c])));
    }
    cnt[p - fb] = w;
}
__global__ void kPlace(int fb, int fe, int* bfs, const int* rs, const int* ent, const int* l, const int* u,
                       const char* seen, const unsigned long long* claim, const int* pos) {
    const int p = fb + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= fe) return;
    const int c = bfs[p];
    int j = fe + pos[p - fb];
    for (int k = rs[c]; k < rs[c + 1]; ++k) {
        const int nb = nbrOf(ent[k], l, u);
        if (!seen[nb] && claim[nb] == (((unsigned long long)p << 24) | (unsigned)(k - rs[c]))) bfs[j++] = nb;
    }
}
__global__ void kMarkSeen(int a, int b, const int* bfs, char* seen) {
    const int q = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (q < b) seen[bfs[q]] = 1;
}
__global__ void kFirstUnseen(int from, int n, const char* seen, int* res) {
    const int i = from + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !seen[i]) atomicMin(res, i);
}
__global__ void kSetInt(int* a, int v) { *a = v; }
__global__ void kSeed(int at, int cell, int* bfs, char* seen) {
    bfs[at] = cell;
    seen[cell] = 1;
}
__global__ void kFillU64(long long n, unsigned long long v, unsigned long long* a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
// false when the graph has too many components for a launch sequence per
// component (the caller then builds on the host)
bool bfsDev(const DevFaces& D, DBuf<int>& bfs, Temp& t, hipStream_t s) {
    const int n = D.n;
    bfs.alloc(std::max(n, 1));
    DBuf<char> seen(std::max(n, 1));
    UH_CHECK(hipMemsetAsync(seen.get(), 0, (size_t)n, s));
    DBuf<unsigned long long> claim(std::max(n, 1));
    kFillU64<<<blocksFor(n), kT, 0, s>>>(n, ~0ull, claim.get());
    UH_LAUNCHED();
    DBuf<int> cnt(n + 1), pos(n + 1), first(1);
    int placed = 0, seedFrom = 0, components = 0;
    while (placed < n) {
        if (++components > 256) return false;
        kSetInt<<<1, 1, 0, s>>>(first.get(), n);
        UH_LAUNCHED();
        kFirstUnseen<<<blocksFor(n - seedFrom), kT, 0, s>>>(seedFrom, n, seen.get(), first.get());
        UH_LAUNCHED();
        const int seed = readOne(first.get(), s);
        kSeed<<<1, 1, 0, s>>>(placed, seed, bfs.get(), seen.get());
        UH_LAUNCHED();
        seedFrom = seed + 1;
        int fb = placed, fe = placed + 1;
        while (fb < fe) {
            const int m = fe - fb;
            kClaim<<<blocksFor(m), kT, 0, s>>>(fb, fe, bfs.get(), D.rs.get(), D.ent.get(), D.l.get(), D.u.get(),
                                               seen.get(), claim.get());
            UH_LAUNCHED();
            kWins<<<blocksFor(m + 1), kT, 0, s>>>(fb, fe, bfs.get(), D.rs.get(), D.ent.get(), D.l.get(), D.u.get(),
                                                  seen.get(), claim.get(), cnt.get());
            UH_LAUNCHED();
            exclusiveSum(t, cnt.get(), pos.get(), m + 1, s);
            const int added = readOne(pos.get() + m, s);
            kPlace<<<blocksFor(m), kT, 0, s>>>(fb, fe, bfs.get(), D.rs.get(), D.ent.get(), D.l.get(), D.u.get(),
                                               seen.get(), claim.get(), pos.get());
            UH_LAUNCHED();
            if (added > 0) {
                kMarkSeen<<<blocksFor(added), kT, 0, s>>>(fe, fe + added, bfs.get(), seen.get());
                UH_LAUNCHED();
            }
            fb = fe;
            fe += added;
        }
        placed = fe;
    }
    return true;
}

// ------------------------------------------------------------ permutation
__global__ void kGather(int n, const int* idx, const int* val, int* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = val[idx[i]];
}
__global__ void kIota(int n, int* a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}
__global__ void kScatterInverse(int n, const int* perm, int* iperm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) iperm[perm[i]] = i;
}
// Windows of at most 64 rows, one wavefront each: rows stably sorted by
// descending length (the host's insertion sort).
__global__ void kSortWindows(int nw, const int* wStart, const int* wEnd, const int* rs, const int* in, int* out) {
    __shared__ int len[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int w = blockIdx.x * 4 + wv;
    const int a = w < nw ? wStart[w] : 0, m = w < nw ? wEnd[w] - a : 0;
    int v = 0, lv = -1;
    if (lane < m) {
        v = in[a + lane];
        lv = rs[v + 1] - rs[v];
    }
    len[wv][lane] = lv;
    __syncthreads();
    if (lane >= m) return;
    int rank = 0;
    for (int j = 0; j < m; ++j) {
        const int lj = len[wv][j];
        rank += (lj > lv) || (lj == lv && j < lane);
    }
    out[a + rank] = v;
}
void sortWindowsDev(const std::vector<int>& segStart, int sigma, const DevFaces& D, DBuf<int>& perm, hipStream_t s) {
    if (sigma <= 1) return;
    if (sigma > 64) throw Error(UNAP_HIP_ERR_INTERNAL, "device window sort: windows of at most 64 rows");
    std::vector<int> ws, we;
    for (size_t k = 0; k + 1 < segStart.size(); ++k)
        for (int a = segStart[k]; a < segStart[k + 1]; a += sigma) {
            ws.push_back(a);
            we.push_back(std::min(segStart[k + 1], a + sigma));
        }
    if (ws.empty()) return;
    DBuf<int> dws, dwe, out(perm.n);
    dws.upload(ws, s);
    dwe.upload(we, s);
    const int nw = (int)ws.size();
    kSortWindows<<<(nw + 3) / 4, 256, 0, s>>>(nw, dws.get(), dwe.get(), D.rs.get(), perm.get(), out.get());
    UH_LAUNCHED();
    UH_CHECK(hipStreamSynchronize(s));
    perm = std::move(out);
}

// Rows grouped by level (stable: ascending internal row inside a level) and
// the level boundaries.
__global__ void kLevelStarts(int n, const int* sortedLev, int* start) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    const int cur = i < n ? sortedLev[i] : -1;
    const int prev = i > 0 ? sortedLev[i - 1] : -1;
    if (i < n && cur != prev)
        for (int k = prev + 1; k <= cur; ++k) start[k] = i;
    if (i == n && n > 0) start[prev + 1] = n;
}
void groupDev(const int* levOfRow, int n, int nlev, std::vector<int>& rows, std::vector<int>& start, Temp& t,
              hipStream_t s) {
    DBuf<int> kin(std::max(n, 1)), kout(std::max(n, 1)), vin(std::max(n, 1)), vout(std::max(n, 1)), st(nlev + 1);
    UH_CHECK(hipMemcpyAsync(kin.get(), levOfRow, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
    kIota<<<blocksFor(n), kT, 0, s>>>(n, vin.get());
    UH_LAUNCHED();
    sortPairs(t, kin.get(), kout.get(), vin.get(), vout.get(), n, bitsFor(nlev + 1), s);
    UH_CHECK(hipMemsetAsync(st.get(), 0, sizeof(int) * (nlev + 1), s));
    kLevelStarts<<<blocksFor(n + 1), kT, 0, s>>>(n, kout.get(), st.get());
    UH_LAUNCHED();
    rows = download(vout.get(), n, s);
    start = download(st.get(), nlev + 1, s);
}

// ---------------------------------------------------------------- SELL-32
__global__ void kRowInfo(int n, const int* perm, const int* rs, const int* ent, uint32_t* info, int* tooLong) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int c = perm[r];
    const int len = rs[c + 1] - rs[c];
    int low = 0;
    for (int k = rs[c]; k < rs[c + 1]; ++k) low += (ent[k] & 1) == 0;
    if (len > 0xffff) *tooLong = 1;
    info[r] = (uint32_t)len | ((uint32_t)low << 16);
}
__global__ void kSliceWidth(int n, int nslices, const uint32_t* info, int* width, long long* entries) {
    const int sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= nslices) return;
    int mx = 0;
    for (int r = sl * kSliceRows; r < min(n, sl * kSliceRows + kSliceRows); ++r) mx = max(mx, (int)(info[r] & 0xffff));
    width[sl] = mx;
    entries[sl] = (long long)kSliceRows * mx;
}
__global__ void kSellFill(int n, int nslices, const int* perm, const int* iperm, const int* rs, const int* ent,
                          const int* l, const int* u, const int* width, const long long* sliceStart, int* col,
                          int* src) {
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long long)nslices * kSliceRows) return;
    const int sl = (int)(tid >> kSliceShift), lane = (int)(tid & (kSliceRows - 1));
    const int r = sl * kSliceRows + lane;
    const int c = r < n ? perm[r] : -1;
    const int len = r < n ? rs[c + 1] - rs[c] : 0;
    for (int k = 0; k < width[sl]; ++k) {
        const long long p = sliceStart[sl] + (long long)kSliceRows * k + lane;
        if (k < len) {
            const int e = ent[rs[c] + k];
            col[p] = iperm[nbrOf(e, l, u)];
            src[p] = e;
        } else {
            col[p] = r < n ? r : 0;
            src[p] = -1;
        }
    }
}

// ------------------------------------------------- grouping by bucket
// start[k] = the first position whose sorted key is at least k, k in [0, nb]
__global__ void kLowerBounds(long long m, int nb, const int* sorted, int* start) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nb) return;
    long long lo = 0, hi = m;
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (sorted[mid] < k)
            lo = mid + 1;
        else
            hi = mid;
    }
    start[k] = (int)lo;
}
// Items 0..m-1 grouped by key in [0, nb] (key nb: left out, sorted last),
// ascending inside each group (a stable sort of the keys with the items as
// values): start gets nb + 1 boundaries, items the grouped items, sortedKey
// the keys in group order.
void groupByKeyDev(DBuf<int>& key, long long m, int nb, DBuf<int>& start, DBuf<int>& items, DBuf<int>& sortedKey,
                   Temp& t, hipStream_t s) {
    DBuf<int> iota(std::max<long long>(m, 1));
    items.alloc(std::max<long long>(m, 1));
    sortedKey.alloc(std::max<long long>(m, 1));
    if (m > 0) {
        kIota<<<blocksFor(m), kT, 0, s>>>((int)m, iota.get());
        UH_LAUNCHED();
        sortPairs(t, key.get(), sortedKey.get(), iota.get(), items.get(), m, bitsFor((long long)nb + 1), s);
    }
    start.alloc(nb + 1);
    kLowerBounds<<<blocksFor(nb + 1), kT, 0, s>>>(m, nb, sortedKey.get(), start.get());
    UH_LAUNCHED();
}

// ------------------------------------------------- Galerkin segment lists
__global__ void kKeyFromMap(int m, const int* map, int* key) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) key[i] = map[i];
}
__global__ void kInnerKey(int m, int nc, const int* frm, int* key) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < m) key[f] = frm[f] < 0 ? -1 - frm[f] : nc;
}
__global__ void kFaceKey(int m, int ncf, const int* frm, int* key) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < m) key[f] = frm[f] >= 0 ? frm[f] : ncf;
}
// orientation (lduAgglomerateMatrix.cpp:50-66): the fine owner lies in the
// coarse owner, or the sides swap (~f)
__global__ void kOrient(const int* count, const int* cfOf, int* ff, const int* rm, const int* fl, const int* cl,
                        const int* cu, int* bad) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *count) return;
    const int f = ff[k], cf = cfOf[k];
    const int owner = rm[fl[f]];
    if (cl[cf] == owner) return;
    if (cu[cf] == owner)
        ff[k] = ~f;
    else
        *bad = 1;
}

// ------------------------------------------------- coarse addressing
__global__ void kFaceOwners(int nf, const int* l, const int* u, const int* rm, int nc, int* frm, int* ownKey,
                            int* nei) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const int a = rm[u[f]], b = rm[l[f]];
    if (a == b) {
        frm[f] = -(a + 1);
        ownKey[f] = nc;
    } else {
        ownKey[f] = min(a, b);
        nei[f] = max(a, b);
    }
}
// first appearance of each element's coarse neighbour inside its owner group
__global__ void kFirstInGroup(const int* count, const int* own, const int* start, const int* bucket, const int* nei,
                              int* isFirst, int* firstOf) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > *count) return;
    if (k == *count) {
        isFirst[k] = 0;
        return;
    }
    const int v = nei[bucket[k]];
    int first = k;
    for (int j = start[own[k]]; j < k; ++j)
        if (nei[bucket[j]] == v) {
            first = j;
            break;
        }
    isFirst[k] = first == k;
    firstOf[k] = first;
}
__global__ void kCoarseFaces(const int* count, const int* own, const int* bucket, const int* nei, const int* firstOf,
                             const int* slotScan, int* frm, int* cl, int* cu) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *count) return;
    const int first = firstOf[k];
    const int slot = slotScan[first];
    frm[bucket[k]] = slot;
    if (first == k) {
        cl[slot] = own[k];
        cu[slot] = nei[bucket[k]];
    }
}
// next level's face weights: per owner group, in ascending fine face order
__global__ void kGroupWeights(int nc, const int* start, const int* bucket, const int* frm, const double* w,
                              double* cw) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nc) return;
    for (int k = start[o]; k < start[o + 1]; ++k) {
        const int f = bucket[k];
        cw[frm[f]] += w[f];
    }
}
__global__ void kZeroD(long long n, double* a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

// ------------------------------------------------- COO -> LDU
__global__ void kCooKeys(long long m, int n, const int64_t* rows, const int64_t* cols, int64_t base, uint64_t* key,
                         long long* idx, int* bad) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int64_t r = rows[k] - base, c = cols[k] - base;
    if (r < 0 || r >= n || c < 0 || c >= n) {
        *bad = 1;
        key[k] = 0;
    } else {
        key[k] = ((uint64_t)r << 32) | (uint64_t)c;
    }
    idx[k] = k;
}
__global__ void kCooClassify(long long m, const uint64_t* key, int* dup, int* isUpper) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > m) return;
    if (k == m) {
        isUpper[k] = 0;
        return;
    }
    if (k > 0 && key[k] == key[k - 1]) *dup = 1;
    isUpper[k] = (uint32_t)key[k] > (uint32_t)(key[k] >> 32);
}
__global__ void kCooPlace(long long m, const uint64_t* key, const long long* idx, const double* vals,
                          const int* faceAt, int64_t* l, int64_t* u, double* up, double* d) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int r = (int)(key[k] >> 32), c = (int)(uint32_t)key[k];
    const double v = vals[idx[k]];
    if (c == r) {
        d[r] = v;
    } else if (c > r) {
        const int f = faceAt[k];
        l[f] = r;
        u[f] = c;
        up[f] = v;
    }
}
// lower entry (r, c < r): its face is the upper entry (c, r), found by a
// binary search of the sorted keys
__global__ void kCooLower(long long m, const uint64_t* key, const long long* idx, const double* vals,
                          const int* faceAt, double* lo, int* noUpper, int* matched) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const uint64_t kk = key[k];
    const int r = (int)(kk >> 32), c = (int)(uint32_t)kk;
    if (c >= r) return;
    const uint64_t want = ((uint64_t)c << 32) | (uint64_t)r;
    long long lo_ = 0, hi = m;
    while (lo_ < hi) {
        const long long mid = (lo_ + hi) >> 1;
        if (key[mid] < want)
            lo_ = mid + 1;
        else
            hi = mid;
    }
    if (lo_ == m || key[lo_] != want) {
        *noUpper = 1;
        return;
    }
    lo[faceAt[lo_]] = vals[idx[k]];
    atomicAdd(matched, 1);
}

}  // namespace

// Galerkin segment lists of one level on the device: children ascending by
// fine cell, inner faces ascending by fine face (per coarse cell), and the
// fine faces of each coarse face ascending (~f where the sides swap).
void galerkinListsDevice(const CoarseLevel& C, const std::vector<int>& fineL, DBuf<int>& cst, DBuf<int>& cl,
                         DBuf<int>& ist, DBuf<int>& il, DBuf<int>& fst, DBuf<int>& ff) {
    Ctx& c = Ctx::get();
    const hipStream_t s = c.stream;
    Temp t;
    const int nfc = (int)C.restrictMap.size(), nff = (int)C.faceRestrictMap.size();
    const int nc = C.nCoarse, ncf = C.A.nFaces();
    DBuf<int> rm, frm, fl, cL, cU, key(std::max(std::max(nfc, nff), 1)), sorted;
    rm.upload(C.restrictMap, s);
    frm.upload(C.faceRestrictMap, s);
    fl.upload(fineL, s);
    cL.upload(C.A.l, s);
    cU.upload(C.A.u, s);
    if (nfc > 0) {
        kKeyFromMap<<<blocksFor(nfc), kT, 0, s>>>(nfc, rm.get(), key.get());
        UH_LAUNCHED();
    }
    groupByKeyDev(key, nfc, nc, cst, cl, sorted, t, s);
    if (nff > 0) {
        kInnerKey<<<blocksFor(nff), kT, 0, s>>>(nff, nc, frm.get(), key.get());
        UH_LAUNCHED();
    }
    groupByKeyDev(key, nff, nc, ist, il, sorted, t, s);
    if (nff > 0) {
        kFaceKey<<<blocksFor(nff), kT, 0, s>>>(nff, ncf, frm.get(), key.get());
        UH_LAUNCHED();
    }
    groupByKeyDev(key, nff, ncf, fst, ff, sorted, t, s);
    DBuf<int> bad(1);
    UH_CHECK(hipMemsetAsync(bad.get(), 0, sizeof(int), s));
    if (nff > 0) {
        kOrient<<<blocksFor(nff), kT, 0, s>>>(fst.get() + ncf, sorted.get(), ff.get(), rm.get(), fl.get(), cL.get(),
                                              cU.get(), bad.get());
        UH_LAUNCHED();
    }
    if (readOne(bad.get(), s)) throw Error(UNAP_HIP_ERR_INTERNAL, "agglomeration: inconsistent coarse addressing");
}

// coarseAddressing of host_setup.cpp on the device: the coarse faces
// (owner-major, neighbours in first-appearance order), the fine-to-coarse
// face map and the next level's face weights.
void coarseAddressingDevice(const HostLdu& F, CoarseLevel& C, const std::vector<double>& w, std::vector<double>& cw) {
    Ctx& c = Ctx::get();
    const hipStream_t s = c.stream;
    Temp t;
    const int nf = F.nFaces(), nc = C.nCoarse;
    DBuf<int> l, u, rm, wFrm(std::max(nf, 1)), key(std::max(nf, 1)), nei(std::max(nf, 1));
    DBuf<double> dw;
    l.upload(F.l, s);
    u.upload(F.u, s);
    rm.upload(C.restrictMap, s);
    dw.upload(w, s);
    if (nf > 0) {
        kFaceOwners<<<blocksFor(nf), kT, 0, s>>>(nf, l.get(), u.get(), rm.get(), nc, wFrm.get(), key.get(),
                                                 nei.get());
        UH_LAUNCHED();
    }
    DBuf<int> start, bucket, own;
    groupByKeyDev(key, nf, nc, start, bucket, own, t, s);
    const int m = readOne(start.get() + nc, s);  // faces between clusters
    DBuf<int> isFirst(m + 1), firstOf(std::max(m, 1)), slotScan(m + 1);
    kFirstInGroup<<<blocksFor(m + 1), kT, 0, s>>>(start.get() + nc, own.get(), start.get(), bucket.get(), nei.get(),
                                                  isFirst.get(), firstOf.get());
    UH_LAUNCHED();
    exclusiveSum(t, isFirst.get(), slotScan.get(), m + 1, s);
    const int ncf = readOne(slotScan.get() + m, s);
    DBuf<int> cl(std::max(ncf, 1)), cu(std::max(ncf, 1));
    DBuf<double> dcw(std::max(ncf, 1));
    if (m > 0) {
        kCoarseFaces<<<blocksFor(m), kT, 0, s>>>(start.get() + nc, own.get(), bucket.get(), nei.get(), firstOf.get(),
                                                 slotScan.get(), wFrm.get(), cl.get(), cu.get());
        UH_LAUNCHED();
    }
    kZeroD<<<blocksFor(ncf), kT, 0, s>>>(ncf, dcw.get());
    UH_LAUNCHED();
    if (nc > 0) {
        kGroupWeights<<<blocksFor(nc), kT, 0, s>>>(nc, start.get(), bucket.get(), wFrm.get(), dw.get(), dcw.get());
        UH_LAUNCHED();
    }
    C.faceRestrictMap = download(wFrm.get(), nf, s);
    C.A.n = nc;
    C.A.l = download(cl.get(), ncf, s);
    C.A.u = download(cu.get(), ncf, s);
    cw = download(dcw.get(), ncf, s);
}

// The host builder's buildLayout, on the device.  Returns false (L untouched)
// where it hands the work back: a graph of more than 256 components.
bool buildLayoutDevice(const HostLdu& A, int orderMode, int sigma, Layout& L) {
    Ctx& c = Ctx::get();
    const hipStream_t s = c.stream;
    Temp t;
    DevFaces D;
    rowFacesDev(A, false, D, t, s);
    const int n = A.n;
    const bool sweepable = A.ownerSorted();
    DBuf<int> lev, blev;
    int nlev = 1, nblev = 1;
    if (sweepable) {
        nlev = levelsDev(D, true, lev, t, s);
        nblev = levelsDev(D, false, blev, t, s);
    } else {
        lev.alloc(std::max(n, 1));
        blev.alloc(std::max(n, 1));
        UH_CHECK(hipMemsetAsync(lev.get(), 0, sizeof(int) * n, s));
        UH_CHECK(hipMemsetAsync(blev.get(), 0, sizeof(int) * n, s));
    }
    DBuf<int> perm;
    if (orderMode == 2) {
        perm.alloc(std::max(n, 1));
        kIota<<<blocksFor(n), kT, 0, s>>>(n, perm.get());
        UH_LAUNCHED();
    } else {
        DBuf<int> bfs;
        if (!bfsDev(D, bfs, t, s)) return false;
        if (orderMode == 1) {
            perm = std::move(bfs);
            sortWindowsDev({0, n}, sigma, D, perm, s);
        } else {  // forward level major, breadth-first minor
            DBuf<int> key(std::max(n, 1)), kout(std::max(n, 1));
            perm.alloc(std::max(n, 1));
            kGather<<<blocksFor(n), kT, 0, s>>>(n, bfs.get(), lev.get(), key.get());
            UH_LAUNCHED();
            sortPairs(t, key.get(), kout.get(), bfs.get(), perm.get(), n, bitsFor(nlev + 1), s);
            DBuf<int> st(nlev + 1);
            UH_CHECK(hipMemsetAsync(st.get(), 0, sizeof(int) * (nlev + 1), s));
            kLevelStarts<<<blocksFor(n + 1), kT, 0, s>>>(n, kout.get(), st.get());
            UH_LAUNCHED();
            sortWindowsDev(download(st.get(), nlev + 1, s), sigma, D, perm, s);
        }
    }
    L = Layout();
    L.n = n;
    L.nf = A.nFaces();
    L.sweepable = sweepable;
    DBuf<int> iperm(std::max(n, 1));
    kScatterInverse<<<blocksFor(n), kT, 0, s>>>(n, perm.get(), iperm.get());
    UH_LAUNCHED();
    // sweep schedules over internal rows
    {
        DBuf<int> li(std::max(n, 1)), bi(std::max(n, 1));
        kGather<<<blocksFor(n), kT, 0, s>>>(n, perm.get(), lev.get(), li.get());
        UH_LAUNCHED();
        kGather<<<blocksFor(n), kT, 0, s>>>(n, perm.get(), blev.get(), bi.get());
        UH_LAUNCHED();
        std::vector<int> fr;
        groupDev(li.get(), n, nlev, fr, L.fwdStart, t, s);
        bool contiguous = true;
        for (int i = 0; i < n && contiguous; ++i) contiguous = (fr[i] == i);
        if (!contiguous) L.fwdRows = std::move(fr);
        groupDev(bi.get(), n, nblev, L.bwdRows, L.bwdStart, t, s);
    }
    // SELL-32
    L.nslices = (n + kSliceRows - 1) / kSliceRows;
    DBuf<uint32_t> info(std::max(n, 1));
    DBuf<int> tooLong(1), width(std::max(L.nslices, 1));
    DBuf<long long> entries(L.nslices + 1), sliceStart(L.nslices + 1);
    UH_CHECK(hipMemsetAsync(tooLong.get(), 0, sizeof(int), s));
    kRowInfo<<<blocksFor(n), kT, 0, s>>>(n, perm.get(), D.rs.get(), D.ent.get(), info.get(), tooLong.get());
    UH_LAUNCHED();
    UH_CHECK(hipMemsetAsync(entries.get() + L.nslices, 0, sizeof(long long), s));
    kSliceWidth<<<blocksFor(L.nslices), kT, 0, s>>>(n, L.nslices, info.get(), width.get(), entries.get());
    UH_LAUNCHED();
    exclusiveSum(t, entries.get(), sliceStart.get(), L.nslices + 1, s);
    if (readOne(tooLong.get(), s)) throw std::runtime_error("row longer than 65535 entries");
    L.maxRowLen = L.nslices > 0 ? maxOf(t, width.get(), L.nslices, s) : 0;
    L.sliceStart = download(sliceStart.get(), L.nslices + 1, s);
    L.nnzPadded = L.sliceStart[L.nslices];
    DBuf<int> col(std::max<long long>(L.nnzPadded, 1)), src(std::max<long long>(L.nnzPadded, 1));
    if (L.nnzPadded > 0) {
        kSellFill<<<blocksFor((long long)L.nslices * kSliceRows), kT, 0, s>>>(
            n, L.nslices, perm.get(), iperm.get(), D.rs.get(), D.ent.get(), D.l.get(), D.u.get(), width.get(),
            sliceStart.get(), col.get(), src.get());
        UH_LAUNCHED();
    }
    L.perm = download(perm.get(), n, s);
    L.iperm = download(iperm.get(), n, s);
    L.rowInfo = download(info.get(), n, s);
    L.dev = std::make_shared<LayoutDev>();
    L.dev->col = std::move(col);
    L.dev->src = std::move(src);
    return true;
}

// The pairing lists of pairCells: per cell the cell across each face and the
// face weight, neighbour side first, then owner side, each by face id.
void pairingListsDevice(const HostLdu& A, const std::vector<double>& w, std::vector<int>& rs, std::vector<int>& nbr,
                        std::vector<double>& wk) {
    Ctx& c = Ctx::get();
    const hipStream_t s = c.stream;
    Temp t;
    DevFaces D;
    rowFacesDev(A, true, D, t, s);
    const long long m = 2LL * D.nf;
    DBuf<double> dw;
    dw.upload(w, s);
    DBuf<int> dn(std::max<long long>(m, 1));
    DBuf<double> dwk(std::max<long long>(m, 1));
    if (m > 0) {
        kPairLists<<<blocksFor(m), kT, 0, s>>>(m, D.ent.get(), D.l.get(), D.u.get(), dw.get(), dn.get(), dwk.get());
        UH_LAUNCHED();
    }
    rs = download(D.rs.get(), D.n + 1, s);
    nbr = download(dn.get(), m, s);
    wk = download(dwk.get(), m, s);
}

// cooToLdu of formats.cpp on the device: the same faces, values, checks and
// messages.  Returns false where no GPU is present (the caller converts on
// the host); err is set when the input is refused.
bool cooToLduDevice(int64_t n, int64_t nnz, const int64_t* rows, const int64_t* cols, const double* vals,
                    bool symmetric, int64_t base, std::vector<int64_t>& l, std::vector<int64_t>& u,
                    std::vector<double>& lo, std::vector<double>& d, std::vector<double>& up, std::string& err) {
    int devices = 0;
    if (!Ctx::peek() && (hipGetDeviceCount(&devices) != hipSuccess || devices == 0)) {
        (void)hipGetLastError();  // no device: the host converts
        return false;
    }
    Ctx& c = Ctx::get();
    const hipStream_t s = c.stream;
    Temp t;
    const long long m = nnz;
    DBuf<int64_t> dr, dc;
    DBuf<double> dv;
    dr.alloc(std::max<long long>(m, 1));
    dc.alloc(std::max<long long>(m, 1));
    dv.alloc(std::max<long long>(m, 1));
    dr.upload(rows, (size_t)m, s);
    dc.upload(cols, (size_t)m, s);
    dv.upload(vals, (size_t)m, s);
    DBuf<uint64_t> k0(std::max<long long>(m, 1)), k1(std::max<long long>(m, 1));
    DBuf<long long> i0(std::max<long long>(m, 1)), i1(std::max<long long>(m, 1));
    DBuf<int> flags(4);  // out of range, duplicate, lower without upper, matched lowers
    UH_CHECK(hipMemsetAsync(flags.get(), 0, sizeof(int) * 4, s));
    if (m > 0) {
        kCooKeys<<<blocksFor(m), kT, 0, s>>>(m, (int)n, dr.get(), dc.get(), base, k0.get(), i0.get(), flags.get());
        UH_LAUNCHED();
    }
    if (readOne(flags.get(), s)) {
        err = "COO index out of range";
        return true;
    }
    if (m > 0) {
        size_t bytes = 0;
        const unsigned endBit = 32 + bitsFor(n);
        UH_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, k0.get(), k1.get(), i0.get(), i1.get(), (size_t)m, 0u, endBit,
                                           s));
        UH_CHECK(rocprim::radix_sort_pairs(t.get(bytes), bytes, k0.get(), k1.get(), i0.get(), i1.get(), (size_t)m, 0u,
                                           endBit, s));
    }
    DBuf<int> isUpper(m + 1), faceAt(m + 1);
    kCooClassify<<<blocksFor(m + 1), kT, 0, s>>>(m, k1.get(), flags.get() + 1, isUpper.get());
    UH_LAUNCHED();
    if (readOne(flags.get() + 1, s)) {
        err = "duplicate COO entry";
        return true;
    }
    exclusiveSum(t, isUpper.get(), faceAt.get(), m + 1, s);
    const int nf = readOne(faceAt.get() + m, s);
    DBuf<int64_t> dl(std::max(nf, 1)), du(std::max(nf, 1));
    DBuf<double> dup(std::max(nf, 1)), dlo(std::max(nf, 1)), dd(std::max<long long>(n, 1));
    kZeroD<<<blocksFor(n), kT, 0, s>>>(n, dd.get());
    UH_LAUNCHED();
    if (m > 0) {
        kCooPlace<<<blocksFor(m), kT, 0, s>>>(m, k1.get(), i1.get(), dv.get(), faceAt.get(), dl.get(), du.get(),
                                              dup.get(), dd.get());
        UH_LAUNCHED();
    }
    if (!symmetric) {
        kZeroD<<<blocksFor(nf), kT, 0, s>>>(nf, dlo.get());
        UH_LAUNCHED();
        if (m > 0) {
            kCooLower<<<blocksFor(m), kT, 0, s>>>(m, k1.get(), i1.get(), dv.get(), faceAt.get(), dlo.get(),
                                                  flags.get() + 2, flags.get() + 3);
            UH_LAUNCHED();
        }
        const std::vector<int> f = download(flags.get(), 4, s);
        if (f[2]) {
            err = "COO structure is not symmetric (a lower entry has no upper one)";
            return true;
        }
        if (f[3] != nf) {
            err = "COO structure is not symmetric (an upper entry has no lower one)";
            return true;
        }
        lo = download(dlo.get(), nf, s);
    } else {
        lo.clear();
    }
    l = download(dl.get(), nf, s);
    u = download(du.get(), nf, s);
    up = download(dup.get(), nf, s);
    d = download(dd.get(), n, s);
    return true;
}

// Smallest matrix built on the device (UNAP_HIP_DEVICE_SETUP_MIN: cells;
// 0 every matrix, a negative value none).
int deviceSetupMin() {
    static const int v = [] {
        const char* e = std::getenv("UNAP_HIP_DEVICE_SETUP_MIN");
        return e ? std::atoi(e) : 4096;
    }();
    return v;
}

}  // namespace unap_hip
