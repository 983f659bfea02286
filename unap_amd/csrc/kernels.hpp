// kernels.hpp -- CDNA4 (gfx950) kernels of the LDU solve path.
//
// Storage: every matrix lives in SELL-32 slices in the internal cell order (see
// host_setup.hpp).  One thread owns one row; a wavefront owns two 32-row
// slices, so the k-th coefficient/column load of a wavefront is two
// contiguous 256 B (FP64) or 128 B (FP32) segments.  Rows accumulate strictly in the reference's
// order (diagonal first, then faces in ascending face id) with separately
// rounded multiplies and adds (built with -ffp-contract=off), which makes the
// FP64 Amul, sweeps and transfers bit-identical to the serial x86 loops.
//
// Reductions: every thread keeps FP64 partial sums over a fixed row set, the
// block reduces them with a fixed wave64 shuffle tree, block totals go to a
// partials array, and a one-block finalize kernel sums them in index order and
// runs the solver's scalar step.  Results are run-to-run deterministic.
//
// Control: scalar state (rho, alpha, convergence flags, ...) stays on the
// device; every kernel takes `skip` and returns at once when it is set, so the
// host can enqueue several iterations ahead and only read back the flags.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "host_setup.hpp"  // kSliceShift

namespace unap_hip {

constexpr int kBlock = 256;      // threads per block of the row kernels
constexpr int kMaxBlocks = 2048; // partial sums per reduction slot
constexpr int kMaxSums = 3;
constexpr double kSmallD = 1.0e-37;   // SMALL (src/base/unap.hpp:88)
constexpr double kVSmallD = 1.0e-128; // VSMALL (src/base/unap.hpp:89)

// ---------------------------------------------------------------- matrix view
template <typename T>
struct DMat {
    int n;
    const long long* sliceStart;
    const uint32_t* rowInfo;  // len | nLowerSide << 16
    const int* col;
    const T* val;
    const T* diag;
};

// Solver scalar state (one per running Krylov / eigen / MG context).
struct KState {
    double rho, rhoOld, alpha, beta, omega, pq;
    double initRes, finalRes, prevRes;
    double tol, relTol;
    long long maxIter, minIter, nIter;
    int converged, singular, done, stop;
    long long exitIt;  // the iteration that left on ||s|| (-1: none)
    int exitApplied;   // its x += alpha y has been applied
    // flexible BiCGStab: shadow-residual restart state
    double r0norm;
    long long kShadow;       // iterations since the shadow residual was (re)set
    int restart, noRestart;  // this iteration restarts / does not
    double* hist;
    long long histCap;
    double alphas[32], betas[32];
    double lambda;
    double sf[64];      // scaling factors per AMG level (+ finest at [63])
    double cheb[3 * 64];  // Chebyshev {1/theta, dtemp1, dtemp2} per sweep
};

// Up to two device flags; a kernel is skipped when either is set.
struct Skip {
    const int* a;
    const int* b;
    __device__ __forceinline__ bool on() const { return (a && *a) || (b && *b); }
};

// acc = 0 + d*x[r]; acc += v_k * x[c_k] over the row's faces in order.
// Entries are fetched in batches (the batch's columns and coefficients first,
// then its x gathers) so each lane keeps several loads in flight; the
// additions stay in entry order.  Two load shapes, chosen per matrix at
// compile time of the launching kernel:
//  * kShapeWide8 (default): batches of 8 whose positions past the slice width
//    load nothing.  Best where rows carry up to ~12 entries: at 256^3 the
//    finest FP32 Chebyshev sweep 263 -> 255 us, the FP64 Amul 406 -> 360 us,
//    the first coarse level (12 entries a row) 150 -> 145 us;
//  * kShapeBatch4: batches of 4 and a scalar tail, better for the longer rows
//    (17-23 entries) of the smaller coarse levels, whose launches are
//    latency-bound (e.g. 198k rows: 14.7 us against 20.7 us).
// The wide shape serves the large levels (the finest and the first coarse
// one), whose matrix arrays are read once a launch: those loads are
// nontemporal, so streaming them does not evict the gathered vector from the
// L2 (256^3 FP32 Amul 229 -> 210 us, FP64 358 -> 343 us).  The small levels'
// matrices, re-read by every sweep of a V-cycle, stay cacheable.
constexpr int kUnroll = 4;  // also the load batch of a sweep row (rowSub)
constexpr int kAmulBatch = 8;
constexpr int kShapeWide8 = 0, kShapeBatch4 = 1;

// Streaming accesses (nontemporal): data a kernel touches once.
template <typename T>
__device__ __forceinline__ T ldS(const T* p) {
    return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ __forceinline__ void stS(T* p, T v) {
    __builtin_nontemporal_store(v, p);
}
// Streaming when the rows belong to a large (wide-shape) level.
template <int Shape, typename T>
__device__ __forceinline__ T ldL(const T* p) {
    if constexpr (Shape == kShapeWide8) return ldS(p);
    return *p;
}
template <int Shape, typename T>
__device__ __forceinline__ void stL(T* p, T v) {
    if constexpr (Shape == kShapeWide8)
        stS(p, v);
    else
        *p = v;
}

template <typename T, typename TA, typename TX, int Shape = kShapeWide8>
__device__ __forceinline__ TA rowAmul(const DMat<T>& A, int r, const TX* __restrict__ x) {
    const int s = r >> kSliceShift;
    const long long s0 = A.sliceStart[s];
    const int slen = (int)((A.sliceStart[s + 1] - s0) >> kSliceShift);
    const int len = (int)(A.rowInfo[r] & 0xffffu);
    const long long base = s0 + (r & (kSliceRows - 1));
    TA acc = (TA)0 + (TA)ldL<Shape>(&A.diag[r]) * (TA)x[r];
    if constexpr (Shape == kShapeWide8) {
        for (int k = 0; k < slen; k += kAmulBatch) {
            int c[kAmulBatch];
            T v[kAmulBatch];
            TX xv[kAmulBatch];
#pragma unroll
            for (int j = 0; j < kAmulBatch; ++j) {
                const long long p = base + ((long long)(k + j) << kSliceShift);
                c[j] = (k + j < slen) ? ldS(&A.col[p]) : 0;
                v[j] = (k + j < slen) ? ldS(&A.val[p]) : (T)0;
            }
#pragma unroll
            for (int j = 0; j < kAmulBatch; ++j) xv[j] = (k + j < slen) ? x[c[j]] : (TX)0;
#pragma unroll
            for (int j = 0; j < kAmulBatch; ++j)
                if (k + j < len) acc = acc + (TA)v[j] * (TA)xv[j];
        }
    } else {
        int k = 0;
        for (; k + kUnroll <= slen; k += kUnroll) {
            int c[kUnroll];
            T v[kUnroll];
            TX xv[kUnroll];
#pragma unroll
            for (int j = 0; j < kUnroll; ++j) {
                const long long p = base + ((long long)(k + j) << kSliceShift);
                c[j] = A.col[p];
                v[j] = A.val[p];
            }
#pragma unroll
            for (int j = 0; j < kUnroll; ++j) xv[j] = x[c[j]];
#pragma unroll
            for (int j = 0; j < kUnroll; ++j)
                if (k + j < len) acc = acc + (TA)v[j] * (TA)xv[j];
        }
        for (; k < slen; ++k) {
            const long long p = base + ((long long)k << kSliceShift);
            const TA v = (TA)A.val[p];
            const TA xv = (TA)x[A.col[p]];
            if (k < len) acc = acc + v * xv;
        }
    }
    return acc;
}

// ------------------------------------------------------------ reductions
template <int NS>
__device__ __forceinline__ void blockReduce(double (&v)[NS], double* sm) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double a = v[s];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) a += __shfl_xor(a, o, 64);
        v[s] = a;
    }
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) sm[s * 16 + wave] = v[s];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            double a = 0.0;
            for (int w = 0; w < nw; ++w) a += sm[s * 16 + w];
            v[s] = a;
        }
    }
}

// Row kernels walk the rows grid-stride: block b takes rows b*256 + t,
// then b*256 + t + G*256, ...  At any moment the whole chip works on one
// window of G*256 consecutive rows, so with a bandwidth-reducing (breadth-
// first) cell order the x[] gathers of all blocks hit the same halo in L2 and
// the infinity cache.  On a 256^3 Poisson matrix this streamed Amul at
// 4.1-4.2 TB/s (FP64) against 3.5-3.8 TB/s for one contiguous chunk per
// block, with or without XCD-grouped chunks.
#define UH_ROWS(r, n) \
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < (n); r += gridDim.x * blockDim.x)

// Row kernel with NS FP64 sums.  Body: void operator()(int r, double (&acc)[NS]).
// Block b's partial goes to partials[s * kMaxBlocks + b]; the grid depends on
// n only, so the summation order is fixed for a given matrix.
template <class Body, int NS>
__global__ void __launch_bounds__(kBlock) kRowsReduce(Body body, int n, double* __restrict__ partials, Skip skip) {
    if (skip.on()) return;
    __shared__ double sm[kMaxSums * 16];
    double acc[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[s] = 0.0;
    UH_ROWS(r, n) body(r, acc);
    blockReduce<NS>(acc, sm);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) partials[s * kMaxBlocks + blockIdx.x] = acc[s];
    }
}

// Sum partials[s * kMaxBlocks + 0 .. nblocks) per slot, thread-strided then
// the block tree, and hand the totals to fin (thread 0).
template <class Fin, int NS>
__device__ __forceinline__ void finishPartials(const Fin& fin, const double* partials, int nblocks, double* sm) {
    double acc[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double a = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) a += partials[s * kMaxBlocks + b];
        acc[s] = a;
    }
    blockReduce<NS>(acc, sm);
    if (threadIdx.x == 0) fin(acc);
}

// Row kernel without sums.
template <class Body>
__global__ void __launch_bounds__(kBlock) kRows(Body body, int n, Skip skip) {
    if (skip.on()) return;
    UH_ROWS(r, n) body(r);
}

// Rows in XCD slabs.  Workgroups go to the eight XCDs in turn (workgroup b
// to XCD b mod 8), so the grid-stride walk above spreads consecutive rows
// over all eight L2 caches.  Here, in each step of the walk, XCD x takes the
// x-th eighth of the rows in flight as one contiguous slab, and its L2
// caches only the gathers of that slab.  The grid is a multiple of 8; every
// row is still visited exactly once.
template <class Body>
__global__ void __launch_bounds__(kBlock) kRowsXcd(Body body, int n, Skip skip) {
    if (skip.on()) return;
    const int per = gridDim.x >> 3;
    const long long lane = ((long long)(blockIdx.x & 7) * per + (blockIdx.x >> 3)) * kBlock + threadIdx.x;
    const long long step = (long long)gridDim.x * kBlock;
    for (long long r = lane; r < n; r += step) body((int)r);
}

// Reference-order reduction: body.terms(i) summed over the caller's cell
// order (internal id iperm[e] for e = 0, 1, ...), exactly the serial
// left-to-right loops of the reference (unapVectorOperations.cpp:21-30).
// Used when bit-identical residual histories are wanted.  One block of
// kSerialChunk threads: each fetches the terms of one element of a chunk
// into LDS, then thread 0 adds the chunk in element order while the block's
// loads of the next chunk are already in flight (each thread issues them
// before the barrier that ends the chunk).  The additions are the serial
// loop's; only the loads run in parallel.
constexpr int kSerialChunk = 256;
template <class Body, int NS>
__global__ void __launch_bounds__(kSerialChunk) kSerialSum(Body body, const int* __restrict__ iperm, int n,
                                                           double* __restrict__ partials, Skip skip) {
    if (skip.on()) return;
    __shared__ double sm[2][NS][kSerialChunk];
    const int t = threadIdx.x;
    double acc[NS], v[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        acc[s] = 0.0;
        v[s] = 0.0;
    }
    if (t < n) body.terms(iperm[t], v);
#pragma unroll
    for (int s = 0; s < NS; ++s) sm[0][s][t] = v[s];
    __syncthreads();
    for (int e0 = 0, buf = 0; e0 < n; e0 += kSerialChunk, buf ^= 1) {
        // next chunk's terms: loads issued now, stored after the serial adds
        const int e1 = e0 + kSerialChunk + t;
        if (e1 < n) body.terms(iperm[e1], v);
        if (t == 0) {
            const int m = min(kSerialChunk, n - e0);
            for (int j = 0; j < m; ++j) {
#pragma unroll
                for (int s = 0; s < NS; ++s) acc[s] = acc[s] + sm[buf][s][j];
            }
        }
        if (e1 < n) {
#pragma unroll
            for (int s = 0; s < NS; ++s) sm[buf ^ 1][s][t] = v[s];
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) partials[s * kMaxBlocks] = acc[s];
    }
}

// One block: total of each partials slot in index order, then fin(totals).
template <class Fin, int NS>
__global__ void __launch_bounds__(kBlock) kFinalize(Fin fin, const double* __restrict__ partials,
                                                    int nblocks, Skip skip) {
    if (skip.on()) return;
    __shared__ double sm[kMaxSums * 16];
    finishPartials<Fin, NS>(fin, partials, nblocks, sm);
}

// ------------------------------------------------------ convergence helpers
// solverPerformance::checkConvergence (src/matrix/unapMatrixTest.cpp:3-31)
__device__ __forceinline__ bool checkConv(KState* st, long long iter) {
    const bool c = ((st->finalRes < st->tol ||
                     (st->relTol > kSmallD && st->finalRes <= st->relTol * st->initRes)) &&
                    iter >= st->minIter);
    st->converged = c ? 1 : 0;
    return c;
}

// `++nIter < maxIter && !checkConvergence(..., nIter, ...)` of PCG.cpp:167-180
// (the check is skipped once maxIter is reached, as the && short-circuits).
__device__ __forceinline__ void endIteration(KState* st) {
    st->nIter += 1;
    if (st->hist && st->nIter < st->histCap) st->hist[st->nIter] = st->finalRes;
    if (!(st->nIter < st->maxIter))
        st->done = 1;
    else if (checkConv(st, st->nIter))
        st->done = 1;
}

// ------------------------------------------------------------ state kernels
static __global__ void kStateInit(KState* st, double tol, double relTol, long long maxIter,
                           long long minIter, double* hist, long long histCap) {
    if (threadIdx.x != 0) return;
    st->rho = 1.0e37;  // GREAT (PCG.cpp:57)
    st->rhoOld = 1.0e37;
    st->alpha = st->beta = st->omega = st->pq = 0.0;
    st->initRes = st->finalRes = st->prevRes = 0.0;
    st->tol = tol;
    st->relTol = relTol;
    st->maxIter = maxIter;
    st->minIter = minIter;
    st->nIter = 0;
    st->converged = st->singular = st->done = st->stop = 0;
    st->exitIt = -1;
    st->exitApplied = 0;
    st->r0norm = 0.0;
    st->kShadow = 0;
    st->restart = 0;
    st->noRestart = 1;
    st->hist = hist;
    st->histCap = histCap;
}

// Initial residual norm.  precheck: MGSolve.cpp:87-96 tests convergence before
// the first cycle; the Krylov solvers do not.
struct FinInitRes {
    KState* st;
    int precheck;
    __device__ void operator()(double (&t)[1]) const {
        st->initRes = sqrt(t[0]) + kSmallD;
        st->finalRes = st->initRes;
        st->prevRes = st->initRes;
        if (st->hist && st->histCap > 0) st->hist[0] = st->initRes;
        if (precheck && checkConv(st, st->nIter)) st->done = 1;
    }
};

// ---------------------------------------------------------------- PCG steps
// PCG.cpp:100-117: wArA = w.r; singularity test; beta for the p update.
struct FinPcgRho {
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        st->rhoOld = st->rho;
        st->rho = t[0];
        st->singular = (fabs(st->rho) > kSmallD) ? 0 : 1;
        if (st->singular) {
            st->done = 1;
            return;
        }
        if (st->nIter != 0) st->beta = st->rho / st->rhoOld;
    }
};
struct FinPcgAlpha {  // PCG.cpp:139-142
    KState* st;
    __device__ void operator()(double (&t)[1]) const { st->alpha = st->rho / t[0]; }
};
struct FinResNorm {  // PCG.cpp:151 + loop test
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        st->finalRes = sqrt(t[0]) + kSmallD;
        endIteration(st);
    }
};

// ----------------------------------------------------------- PBiCGStab steps
struct FinBicgRho {  // PBiCGStab.cpp:161-226
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        st->rhoOld = st->rho;
        st->rho = t[0];
        st->singular = (fabs(st->rho) > kSmallD) ? 0 : 1;
        if (st->singular) {
            st->done = 1;
            return;
        }
        if (st->nIter != 0) st->beta = (st->rho / st->rhoOld) * (st->alpha / st->omega);
    }
};
struct FinBicgAlpha {  // PBiCGStab.cpp:254-269
    KState* st;
    __device__ void operator()(double (&t)[1]) const { st->alpha = st->rho / t[0]; }
};
struct FinBicgS {  // PBiCGStab.cpp:276-348: converged on ||s|| -> early exit
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        st->finalRes = sqrt(t[0]) + kSmallD;
        if (checkConv(st, st->nIter + 1)) {
            st->exitIt = st->nIter;
            st->nIter += 1;
            if (st->hist && st->nIter < st->histCap) st->hist[st->nIter] = st->finalRes;
            st->done = 1;
        }
    }
};
struct FinBicgOmega {  // PBiCGStab.cpp:357-386: omega = (t.s) / (t.t)
    KState* st;
    __device__ void operator()(double (&t)[2]) const { st->omega = t[1] / t[0]; }
};

// ---------------------------------------------------- flexible BiCGStab steps
// PBiCGStab's loop (PBiCGStab.cpp:45-459) with a variable preconditioner
// (Vogel 2007: right preconditioning keeps x = x + alpha y + omega z
// consistent whatever M^-1 is at each application) and a restart of the
// shadow residual, r0 = r, once |r0.r| < kShadowRestart |r0| |r|: an FP32
// preconditioner is not an exactly linear operator, and the shadow residual
// loses bi-orthogonality with it.
constexpr double kShadowRestart = 1.0e-6;
struct FinFbicgRho {
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        if (st->nIter == 0) st->r0norm = st->initRes - kSmallD;  // r0 = r at the start
        const bool restart = st->kShadow > 0 && fabs(t[0]) < kShadowRestart * st->r0norm * st->finalRes;
        st->restart = restart ? 1 : 0;
        st->noRestart = restart ? 0 : 1;
        if (restart) return;  // rho comes from the restart's r.r
        st->rhoOld = st->rho;
        st->rho = t[0];
        st->singular = (fabs(st->rho) > kSmallD) ? 0 : 1;
        if (st->singular) {
            st->done = 1;
            return;
        }
        if (st->kShadow != 0) st->beta = (st->rho / st->rhoOld) * (st->alpha / st->omega);
    }
};
struct FinFbicgRestart {  // r0 = r: rho = r.r, |r0| = sqrt(rho), directions restart
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        st->rhoOld = st->rho;
        st->rho = t[0];
        st->r0norm = sqrt(t[0]);
        st->kShadow = 0;
        st->noRestart = 1;
        st->singular = (fabs(st->rho) > kSmallD) ? 0 : 1;
        if (st->singular) st->done = 1;
    }
};
struct FinFbicgOmega {
    KState* st;
    __device__ void operator()(double (&t)[2]) const {
        st->omega = t[1] / t[0];
        st->kShadow += 1;
    }
};

// ---------------------------------------------------------------- FCG steps
// Flexible CG (Notay 2000, FCG(1)): rho = r.z and zq = z.q with q = A p of the
// previous iteration; beta = -zq / (p.q), which A-orthogonalises the new
// direction against the previous one for any preconditioner.
struct FinFcgRz {
    KState* st;
    __device__ void operator()(double (&t)[2]) const {
        st->rhoOld = st->rho;
        st->rho = t[0];
        st->singular = (fabs(st->rho) > kSmallD) ? 0 : 1;
        if (st->singular) {
            st->done = 1;
            return;
        }
        if (st->nIter != 0) st->beta = -t[1] / st->pq;
    }
};
struct FinFcgAlpha {  // pq = p.q kept for the next beta; alpha = rho / pq
    KState* st;
    __device__ void operator()(double (&t)[1]) const {
        st->pq = t[0];
        st->alpha = st->rho / t[0];
    }
};

// ------------------------------------------------------------- eigen steps
// eigenDiagPCG.cpp:87-123: wArA and the VSMALL break
struct FinEigRho {
    KState* st;
    int it;
    __device__ void operator()(double (&t)[1]) const {
        st->rhoOld = st->rho;
        st->rho = t[0];
        if (fabs(st->rho) < kVSmallD) {
            st->stop = 1;
            return;
        }
        if (it == 0) {
            st->betas[0] = 0.0;
        } else {
            st->beta = st->rho / st->rhoOld;
            st->betas[it] = st->beta;
        }
    }
};
struct FinEigAlpha {  // eigenDiagPCG.cpp:166-189
    KState* st;
    int it;
    __device__ void operator()(double (&t)[1]) const {
        st->alpha = st->rho / t[0];
        st->alphas[it] = st->alpha;
    }
};

// computeMaxEig (eigenDiagPCG.cpp:213-315): Lanczos tridiagonal from the CG
// coefficients, Gershgorin range from [0,0], Sturm bisection to 1e-5.
__device__ __forceinline__ void chebyTable(KState* st, double boost, double ratio);

static __global__ void kLanczosMaxEig(KState* st, int m, double boost, double ratio) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double dg[32], of[32];  // T[i][i], T[i][i+1]
    for (int i = 0; i < m; ++i) {
        dg[i] = (i == 0) ? 1.0 / st->alphas[i] : 1.0 / st->alphas[i] + st->betas[i] / st->alphas[i - 1];
        of[i] = (i < m - 1) ? sqrt(st->betas[i + 1]) / st->alphas[i] : 0.0;
    }
    double lo = 0.0, hi = 0.0;
    for (int i = 0; i < m; ++i) {
        const double tr = (i == m - 1) ? 0 : of[i];
        const double tl = (i == 0) ? 0 : of[i - 1];
        const double lmax = dg[i] + fabs(tr) + fabs(tl);
        const double lmin = dg[i] - fabs(tr) - fabs(tl);
        lo = (lo >= lmin) ? lmin : lo;
        hi = (hi <= lmax) ? lmax : hi;
    }
    double xn;
    do {
        xn = (lo + hi) * 0.5;
        int s = 0;
        double p0 = xn - dg[0];
        if (p0 < 0) s = 1;
        double p1 = (xn - dg[1]) * p0 - of[0] * of[0];
        if (p1 * p0 < 0) s++;
        for (int k = 2; k < m; ++k) {
            const double p2 = (xn - dg[k]) * p1 - of[k - 1] * of[k - 1] * p0;
            if (p2 * p1 < 0) s++;
            p0 = p1;
            p1 = p2;
        }
        if (s >= 1)
            lo = xn;
        else
            hi = xn;
    } while ((hi - lo) >= 0.00001);
    st->lambda = xn;
    chebyTable(st, boost, ratio);
}

// Chebyshev coefficients of every sweep from the stored estimate
// (chebySmoother.cpp:61-111), in the reference's recurrence order.
__device__ __forceinline__ void chebyTable(KState* st, double boost, double ratio) {
    const double lam = st->lambda;
    const double mx = lam * boost, mn = lam / ratio;
    const double delta = 2.0 / (mx - mn);
    const double theta = 0.5 * (mx + mn);
    const double s1 = theta * delta;
    double rhok = 1.0 / s1;
    st->cheb[0] = 1.0 / theta;
    st->cheb[1] = 0.0;
    st->cheb[2] = 0.0;
    for (int k = 1; k < 64; ++k) {
        const double rhokp1 = 1.0 / (2.0 * s1 - rhok);
        st->cheb[3 * k + 0] = 1.0 / theta;
        st->cheb[3 * k + 1] = rhokp1 * rhok;
        st->cheb[3 * k + 2] = 2.0 * rhokp1 * delta;
        rhok = rhokp1;
    }
}

}  // namespace unap_hip
