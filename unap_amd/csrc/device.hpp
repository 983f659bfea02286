// device.hpp -- host-side plumbing of the HIP solve path: errors, device
// buffers, the process-wide stream and reduction scratch, and the device copy
// of one matrix (DevMatrix) in the layout planned by host_setup.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "galerkin.hpp"
#include "host_setup.hpp"
#include "kernels.hpp"

namespace unap_hip {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define UH_CHECK(expr)                                                                       \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw ::unap_hip::Error(-2, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define UH_LAUNCHED() UH_CHECK(hipGetLastError())

// Message returned by unap_hip_last_error() on this thread.
void setLastError(const std::string& m);

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) UH_CHECK(hipMalloc((void**)&p, sizeof(T) * count));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        if (count > n) throw Error(-5, "upload larger than buffer");
        if (count) UH_CHECK(hipMemcpyAsync(p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s) {
        if (v.size() != n) alloc(v.size());
        upload(v.data(), v.size(), s);
    }
    T* get() const { return p; }
};

// Process-wide execution context: one stream, reduction partials.
struct Ctx {
    hipStream_t stream = nullptr;
    int serialReductions = 0;  // 1: reductions in the reference's serial order
    DBuf<double> partials;
    DBuf<KState> scratchState;  // for standalone calls (eigen estimate, precond apply)
    static Ctx& get();
    static Ctx* peek();  // the context if it exists (creating it allocates)
    void sync() { UH_CHECK(hipStreamSynchronize(stream)); }
};

struct Launch {
    int grid;
};
// One block per 256 rows up to kMaxBlocks (2048 = 8 per CU); larger row
// counts loop grid-stride inside the kernel.
inline Launch rowLaunch(int n) {
    long long g = (n + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > kMaxBlocks) g = kMaxBlocks;
    return {(int)g};
}

template <class Body>
void launchRows(const Body& b, int n, Skip sk = {nullptr, nullptr}) {
    if (n <= 0) return;
    Ctx& c = Ctx::get();
    const Launch L = rowLaunch(n);
    kRows<Body><<<L.grid, kBlock, 0, c.stream>>>(b, n, sk);
    UH_LAUNCHED();
}
// launchRows with the rows in XCD slabs (kRowsXcd), for gathers whose cache
// lines neighbouring rows share; small row counts take the plain walk.
template <class Body>
void launchRowsXcd(const Body& b, int n, Skip sk = {nullptr, nullptr}) {
    if (n <= 0) return;
    const int g = rowLaunch(n).grid;
    if (g < 64) {
        launchRows(b, n, sk);
        return;
    }
    Ctx& c = Ctx::get();
    kRowsXcd<Body><<<(g + 7) / 8 * 8, kBlock, 0, c.stream>>>(b, n, sk);
    UH_LAUNCHED();
}

// Row kernel with NS sums followed by the one-block finalize.  `order` is the
// level's external->internal map, walked by the reference-order mode.
template <int NS, class Body, class Fin>
void launchReduce(const Body& b, int n, const int* order, const Fin& fin, Skip sk = {nullptr, nullptr}) {
    Ctx& c = Ctx::get();
    const Launch L = rowLaunch(n);
    kRowsReduce<Body, NS><<<L.grid, kBlock, 0, c.stream>>>(b, n, c.partials.get(), sk);
    UH_LAUNCHED();
    int nb = L.grid;
    if (c.serialReductions) {
        // reference order: one thread sums the terms serially
        kSerialSum<Body, NS><<<1, kSerialChunk, 0, c.stream>>>(b, order, n, c.partials.get(), sk);
        UH_LAUNCHED();
        nb = 1;
    }
    // A separate one-block finalize: folding it into the row kernel (the last
    // block to finish sums the partials) needs a device-scope fence in every
    // block, which waits for the block's stores and made the large reductions
    // 60-90% slower.
    kFinalize<Fin, NS><<<1, kBlock, 0, c.stream>>>(fin, c.partials.get(), nb, sk);
    UH_LAUNCHED();
}

// A device-built layout's column and slot-source arrays, kept in HBM.
struct LayoutDev {
    DBuf<int> col, src;
};
// The layout with col and src on the host too (downloaded where they live in
// HBM only), for host code that walks a small matrix's slots.
inline Layout hostArrays(const Layout& L, hipStream_t s) {
    Layout H = L;
    if (!L.dev) return H;
    H.col.resize(L.nnzPadded);
    H.src.resize(L.nnzPadded);
    if (L.nnzPadded > 0) {
        UH_CHECK(hipMemcpyAsync(H.col.data(), L.dev->col.get(), sizeof(int) * L.nnzPadded, hipMemcpyDeviceToHost, s));
        UH_CHECK(hipMemcpyAsync(H.src.data(), L.dev->src.get(), sizeof(int) * L.nnzPadded, hipMemcpyDeviceToHost, s));
    }
    UH_CHECK(hipStreamSynchronize(s));
    H.dev.reset();
    return H;
}

// The backward sweep's entries in backward-level order: position q (row
// bwdRows[q]) holds that row's owner-side entries from the last down to the
// first, the order the sweep consumes them, in SELL slices over positions.
// A backward level is then a contiguous run of positions, so its entry loads
// coalesce although its rows are scattered in the forward-level-major layout.
template <typename T>
__global__ void kBuildBwdCopy(DMat<T> A, const int* __restrict__ bwdRows, const long long* __restrict__ bSlice,
                              int* __restrict__ bCol, T* __restrict__ bVal, int n) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int row = bwdRows[q];
    const long long src = A.sliceStart[row >> kSliceShift] + (row & (kSliceRows - 1));
    const uint32_t info = A.rowInfo[row];
    const int len = (int)(info & 0xffffu), mid = (int)(info >> 16);
    const long long dst = bSlice[q >> kSliceShift] + (q & (kSliceRows - 1));
    for (int j = 0; j < len - mid; ++j) {
        const long long from = src + ((long long)(len - 1 - j) << kSliceShift);
        bCol[dst + ((long long)j << kSliceShift)] = A.col[from];
        bVal[dst + ((long long)j << kSliceShift)] = A.val[from];
    }
}

// Device copy of one matrix in precision T.
template <typename T>
struct DevMatrix {
    int n = 0;
    DBuf<long long> sliceStart;
    DBuf<uint32_t> rowInfo;
    DBuf<int> col;
    DBuf<T> val, diag;
    DBuf<T> valU;  // upper on both sides (DIC on asymmetric storage), lazily
    DBuf<int> perm, iperm;
    DBuf<int> src;  // slot -> 2 face + side, kept for matrices filled on the device
    long long nnzPadded = 0;
    DBuf<int> fwdRows, fwdStart, bwdRows, bwdStart;  // fwdRows empty: level-major order
    std::vector<int> fwdStartH, bwdStartH;
    bool sweepable = false;
    int shape = 0;  // rowAmul load shape of the V-cycle kernels: kShapeBatch4 above 14 entries a row
    double entriesPerRow = 0;  // off-diagonal entries a row, on average
    int maxRowLen = 0;
    // backward-sweep copy (kBuildBwdCopy): structure from the layout, entries
    // gathered on the device on first use and dropped when the values change
    std::vector<long long> bwdSliceH;
    std::vector<uint32_t> bwdInfoH;
    DBuf<long long> bwdSlice;
    DBuf<uint32_t> bwdInfo;
    DBuf<int> bwdCol;
    DBuf<T> bwdVal;

    DMat<T> view() const { return {n, sliceStart.get(), rowInfo.get(), col.get(), val.get(), diag.get()}; }
    DMat<T> viewU() const {
        return {n, sliceStart.get(), rowInfo.get(), col.get(), valU.p ? valU.get() : val.get(), diag.get()};
    }
    // The backward copy as a matrix over positions (row length = owner-side
    // entries, no lower side); built on first call.  Owner-side coefficients
    // are the upper ones in both val and valU.
    DMat<T> bwdView(hipStream_t s) {
        if (!bwdVal.p) {
            if (!bwdSlice.p) {
                bwdSlice.upload(bwdSliceH, s);
                bwdInfo.upload(bwdInfoH, s);
            }
            const long long nnz = bwdSliceH.empty() ? 0 : bwdSliceH.back();
            bwdCol.alloc(std::max<long long>(1, nnz));
            bwdVal.alloc(std::max<long long>(1, nnz));
            if (n > 0) {
                kBuildBwdCopy<T><<<(n + 255) / 256, 256, 0, s>>>(view(), bwdRows.get(), bwdSlice.get(), bwdCol.get(),
                                                                  bwdVal.get(), n);
                UH_LAUNCHED();
            }
        }
        return {n, bwdSlice.get(), bwdInfo.get(), bwdCol.get(), bwdVal.get(), nullptr};
    }

    // Values from host coefficients (external order).
    void build(const HostLdu& A, const Layout& L, hipStream_t s) {
        StageTimer t("device matrix: upload + values from the host", L.n);
        buildStructure(L, s);
        fillValues(A, L, false, val, s);
        fillDiagFrom(A, s);
        UH_CHECK(hipStreamSynchronize(s));  // host staging vectors die here
    }
    // SELL values from host coefficients: on the device from the face arrays
    // where the layout's slot sources are in HBM, else assembled on the host
    // (the same values either way).  upperBoth: upper on both sides.
    void fillValues(const HostLdu& A, const Layout& L, bool upperBoth, DBuf<T>& out, hipStream_t s) {
        if (!L.dev) {
            std::vector<T> v;
            fillSellValues<T>(L, A, upperBoth, v);
            out.upload(v, s);
            UH_CHECK(hipStreamSynchronize(s));
            return;
        }
        DBuf<double> up, lo;
        up.upload(A.up, s);
        if (!A.symm && !upperBoth) lo.upload(A.lo, s);
        out.alloc(std::max<long long>(L.nnzPadded, 1));
        if (L.nnzPadded > 0) {
            const long long g = std::min<long long>((L.nnzPadded + 255) / 256, 8192);
            kFillSell<T><<<(unsigned)g, 256, 0, s>>>(L.nnzPadded, L.dev->src.get(), up.get(), lo.p ? lo.get() : nullptr,
                                                     upperBoth ? 1 : 0, out.get());
            UH_LAUNCHED();
        }
        UH_CHECK(hipStreamSynchronize(s));
    }
    // diag in internal order, gathered on the device through perm
    void fillDiagFrom(const HostLdu& A, hipStream_t s) {
        DBuf<double> d;
        d.upload(A.d, s);
        diag.alloc(std::max(n, 1));
        if (n > 0) {
            kFillDiag<T><<<(n + 255) / 256, 256, 0, s>>>(n, perm.get(), d.get(), diag.get());
            UH_LAUNCHED();
        }
        UH_CHECK(hipStreamSynchronize(s));
    }
    // Values from device coefficients in external order (a coarse level's
    // Galerkin product); the slot map stays on the device for later refills.
    void buildOnDevice(const Layout& L, const double* d, const double* up, const double* lo, hipStream_t s) {
        StageTimer t("device matrix: upload + values on the device", L.n);
        buildStructure(L, s);
        if (L.dev) {
            src.alloc(std::max<long long>(L.nnzPadded, 1));
            if (L.nnzPadded > 0)
                UH_CHECK(hipMemcpyAsync(src.get(), L.dev->src.get(), sizeof(int) * L.nnzPadded,
                                        hipMemcpyDeviceToDevice, s));
        } else {
            src.upload(L.src, s);
        }
        nnzPadded = L.nnzPadded;
        fillOnDevice(d, up, lo, s);
        UH_CHECK(hipStreamSynchronize(s));
    }
    // New values (same structure) from device coefficients in external order.
    void fillOnDevice(const double* d, const double* up, const double* lo, hipStream_t s) {
        if (!src.p && nnzPadded > 0) throw Error(-5, "device refill of a matrix built from host values");
        if (val.n != (size_t)std::max<long long>(nnzPadded, 1)) val.alloc(std::max<long long>(nnzPadded, 1));
        if (diag.n != (size_t)n) diag.alloc(n);
        if (nnzPadded > 0) {
            const long long g = std::min<long long>((nnzPadded + 255) / 256, 8192);
            kFillSell<T><<<(unsigned)g, 256, 0, s>>>(nnzPadded, src.get(), up, lo, 0, val.get());
            UH_LAUNCHED();
        }
        if (n > 0) {
            kFillDiag<T><<<(n + 255) / 256, 256, 0, s>>>(n, perm.get(), d, diag.get());
            UH_LAUNCHED();
        }
        valU.release();
        bwdVal.release();  // rebuilt from the new values on the next backward sweep
    }
    // Addressing, schedules and the backward-copy structure; no values.
    void buildStructure(const Layout& L, hipStream_t s) {
        n = L.n;
        sweepable = L.sweepable;
        shape = (2.0 * L.nf > 14.0 * L.n) ? kShapeBatch4 : kShapeWide8;
        entriesPerRow = L.n > 0 ? 2.0 * L.nf / L.n : 0.0;
        maxRowLen = L.maxRowLen;
        sliceStart.upload(L.sliceStart, s);
        rowInfo.upload(L.rowInfo, s);
        if (L.dev) {
            col.alloc(std::max<long long>(L.nnzPadded, 1));
            if (L.nnzPadded > 0)
                UH_CHECK(hipMemcpyAsync(col.get(), L.dev->col.get(), sizeof(int) * L.nnzPadded,
                                        hipMemcpyDeviceToDevice, s));
        } else {
            col.upload(L.col, s);
        }
        perm.upload(L.perm, s);
        iperm.upload(L.iperm, s);
        fwdStartH = L.fwdStart;
        bwdStartH = L.bwdStart;
        fwdStart.upload(L.fwdStart, s);
        if (!L.fwdRows.empty()) fwdRows.upload(L.fwdRows, s);
        bwdRows.upload(L.bwdRows, s);
        bwdStart.upload(L.bwdStart, s);
        if (sweepable) {
            const int ns = (n + kSliceRows - 1) / kSliceRows;
            bwdSliceH.assign(ns + 1, 0);
            bwdInfoH.assign(n, 0);
            for (int q = 0; q < n; ++q) {
                const uint32_t info = L.rowInfo[L.bwdRows[q]];
                bwdInfoH[q] = (info & 0xffffu) - (info >> 16);
            }
            for (int sl = 0; sl < ns; ++sl) {
                uint32_t w = 0;
                for (int q = sl * kSliceRows; q < std::min(n, (sl + 1) * kSliceRows); ++q) w = std::max(w, bwdInfoH[q]);
                bwdSliceH[sl + 1] = bwdSliceH[sl] + (long long)w * kSliceRows;
            }
        }
    }
    void buildValU(const HostLdu& A, const Layout& L, hipStream_t s) {
        if (A.symm || valU.p) return;
        fillValues(A, L, true, valU, s);
    }
    void updateValues(const HostLdu& A, const Layout& L, hipStream_t s) {
        const bool hadU = valU.p != nullptr;
        fillValues(A, L, false, val, s);
        fillDiagFrom(A, s);
        valU.release();
        bwdVal.release();  // rebuilt from the new values on the next backward sweep
        if (hadU) buildValU(A, L, s);
    }
};

}  // namespace unap_hip
