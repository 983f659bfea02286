// coarse.hpp -- the coarsest AMG level solved inside one workgroup.
//
// MGSolver::solveCoarsestLevel (src/solvers/MG/solveCoarsestLevel.cpp:10-55)
// runs PCG + DIC (symmetric) or PBiCGStab + DILU to relTol 1e-2, at most 100
// iterations, from a zero guess.  The coarsest level has a few hundred to a
// few thousand cells, so the whole solve is latency-bound: one 1024-thread
// workgroup keeps every vector in LDS, walks the DIC/DILU level schedule with
// workgroup barriers and reduces with the wave64 shuffle tree.  The arithmetic
// per row is the reference's; reductions differ only in summation order.
#pragma once

#include "ops.hpp"

namespace unap_hip {

struct CoarseArgs {
    DMat<double> A;   // Amul coefficients
    DMat<double> AP;  // preconditioner coefficients (DIC: upper both sides; DILU: as A)
    const double* rD;
    const int* fwdRows;  // null: forward levels are contiguous internal rows
    const int* fwdStart;
    int nFwd;
    const int* bwdRows;
    const int* bwdStart;
    int nBwd;
    const double* src;
    double* corr;
    double relTol, tol;
    int maxIter;
    const int* iperm;  // external -> internal (reference-order reductions)
    int serial;
};

// Sum over the workgroup, broadcast to every thread (fixed order).
template <int NS>
__device__ __forceinline__ void wgSum(double (&v)[NS], double* red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double a = v[s];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) a += __shfl_xor(a, o, 64);
        v[s] = a;
    }
    if (lane == 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) red[s * 16 + wave] = v[s];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        double a = 0.0;
        for (int w = 0; w < nw; ++w) a += red[s * 16 + w];
        v[s] = a;
    }
    __syncthreads();
}

// Sum term(i) over the level: tree over the workgroup, or (serial mode) one
// thread walking the reference's cell order.  Every thread gets the totals.
// The vectors term() reads must be complete (the caller has synchronised).
template <int NS, class F>
__device__ __forceinline__ void wgReduce(const CoarseArgs& a, int n, F term, double (&v)[NS], double* red) {
    if (a.serial) {
        if (threadIdx.x == 0) {
            double acc[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) acc[s] = 0.0;
            for (int e = 0; e < n; ++e) {
                double t[NS];
                term(a.iperm[e], t);
#pragma unroll
                for (int s = 0; s < NS; ++s) acc[s] = acc[s] + t[s];
            }
#pragma unroll
            for (int s = 0; s < NS; ++s) red[48 + s] = acc[s];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NS; ++s) v[s] = red[48 + s];
        __syncthreads();
        return;
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) v[s] = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double t[NS];
        term(i, t);
#pragma unroll
        for (int s = 0; s < NS; ++s) v[s] += t[s];
    }
    wgSum<NS>(v, red);
}

// The coarsest matrix read in place (SELL layout in global memory).
struct CoarseGlobal {
    const CoarseArgs& a;
    __device__ double amul(int i, const double* x) const { return rowAmul<double, double>(a.A, i, x); }
    __device__ void fwd(int i, double* w, const double* r) const { fwdRow<double, double>(a.AP, i, w, r, a.rD); }
    __device__ void bwd(int i, double* w) const { bwdRow<double, double>(a.AP, i, w, a.rD); }
    __device__ int fs(int k) const { return a.fwdStart[k]; }
    __device__ int frow(int q) const { return a.fwdRows ? a.fwdRows[q] : q; }
    __device__ int bs(int k) const { return a.bwdStart[k]; }
    __device__ int brow(int q) const { return a.bwdRows[q]; }
};

// The coarsest matrix in CSR form, built at setup: entries of a row in the
// SELL order (ascending face id), mid = number of neighbour-side entries.
struct CoarseCsr {
    const int* start;
    const uint16_t* col;
    const uint16_t* mid;
    const double* val;
    int nnz;
};

// The same matrix, rD and level schedule copied into LDS at kernel start:
// every row access of the sweeps and products is then an LDS access instead
// of an L2 round trip.  The arithmetic is that of rowAmul / fwdRow / bwdRow.
struct CoarseLocal {
    const double* val;
    const double* diag;
    const double* rD;
    const int* start;
    const uint16_t* col;
    const uint16_t* mid;
    const int* fs_;
    const int* frows;
    const int* bs_;
    const int* brows;
    // acc (+ or - rd *) val[k] * x[col[k]] over k = k0, k0 + step, ... (n
    // entries) in that order.  LDS loads go kLocalBatch at a time (columns
    // and coefficients, then the gathers), positions past the end re-reading
    // the batch's first entry, so a row waits on two LDS round trips a batch
    // rather than two an entry.
    static constexpr int kLocalBatch = 8;
    template <int Op>  // 0: acc + v x;  1: acc - (rd v) x
    __device__ __forceinline__ double run(int k0, int n, int step, double acc, double rd, const double* x) const {
        for (int t0 = 0; t0 < n; t0 += kLocalBatch) {
            int c[kLocalBatch];
            double v[kLocalBatch], xv[kLocalBatch];
#pragma unroll
            for (int j = 0; j < kLocalBatch; ++j) {
                const int k = k0 + step * (t0 + j < n ? t0 + j : t0);
                c[j] = col[k];
                v[j] = val[k];
            }
#pragma unroll
            for (int j = 0; j < kLocalBatch; ++j) xv[j] = x[c[j]];
#pragma unroll
            for (int j = 0; j < kLocalBatch; ++j)
                if (t0 + j < n) acc = Op == 0 ? acc + v[j] * xv[j] : acc - (rd * v[j]) * xv[j];
        }
        return acc;
    }
    __device__ double amul(int i, const double* x) const {
        return run<0>(start[i], start[i + 1] - start[i], 1, 0.0 + diag[i] * x[i], 0.0, x);
    }
    __device__ void fwd(int i, double* w, const double* r) const {
        const double rd = rD[i];
        w[i] = run<1>(start[i], mid[i], 1, rd * r[i], rd, w);
    }
    __device__ void bwd(int i, double* w) const {
        const double rd = rD[i];
        const int b = start[i] + mid[i];
        w[i] = run<1>(start[i + 1] - 1, start[i + 1] - b, -1, w[i], rd, w);
    }
    __device__ int fs(int k) const { return fs_[k]; }
    __device__ int frow(int q) const { return frows ? frows[q] : q; }
    __device__ int bs(int k) const { return bs_[k]; }
    __device__ int brow(int q) const { return brows[q]; }
};

template <class M>
__device__ __forceinline__ void wgPrecond(const CoarseArgs& a, const M& m, double* w, const double* r) {
    for (int k = 0; k < a.nFwd; ++k) {
        for (int q = m.fs(k) + threadIdx.x; q < m.fs(k + 1); q += blockDim.x) m.fwd(m.frow(q), w, r);
        __syncthreads();
    }
    for (int k = 0; k < a.nBwd; ++k) {
        for (int q = m.bs(k) + threadIdx.x; q < m.bs(k + 1); q += blockDim.x) m.bwd(m.brow(q), w);
        __syncthreads();
    }
}

__device__ __forceinline__ bool wgConv(double fin, double init, double tol, double relTol, long long it) {
    return (fin < tol || (relTol > kSmallD && fin <= relTol * init)) && it >= 0;
}

// PCG + DIC (PCG.cpp:35-183) on the coarsest level; x, r, w, p in LDS.
template <class M>
__device__ __forceinline__ void coarsePCG(const CoarseArgs& a, const M& m, double* x, double* r, double* w,
                                          double* p, double* red) {
    const int n = a.A.n;
    const int t = threadIdx.x, nt = blockDim.x;
    double s1[1];
    // x = 0; A x = 0 exactly, so r = b - A x = b
    for (int i = t; i < n; i += nt) {
        x[i] = 0.0;
        r[i] = a.src[i];
        p[i] = 0.0;
    }
    __syncthreads();
    wgReduce<1>(a, n, [&](int i, double(&q)[1]) { q[0] = r[i] * r[i]; }, s1, red);
    const double init = sqrt(s1[0]) + kSmallD;
    double rho = 1.0e37, rhoOld;
    long long nIter = 0;
    for (;;) {
        rhoOld = rho;
        wgPrecond(a, m, w, r);
        wgReduce<1>(a, n, [&](int i, double(&q)[1]) { q[0] = w[i] * r[i]; }, s1, red);
        rho = s1[0];
        if (!(fabs(rho) > kSmallD)) break;
        const double beta = rho / rhoOld;
        for (int i = t; i < n; i += nt) p[i] = (nIter == 0) ? w[i] : w[i] + beta * p[i];
        __syncthreads();
        for (int i = t; i < n; i += nt) w[i] = m.amul(i, p);
        __syncthreads();
        wgReduce<1>(a, n, [&](int i, double(&q)[1]) { q[0] = w[i] * p[i]; }, s1, red);
        const double alpha = rho / s1[0];
        for (int i = t; i < n; i += nt) {
            x[i] += alpha * p[i];
            r[i] -= alpha * w[i];
        }
        __syncthreads();
        wgReduce<1>(a, n, [&](int i, double(&q)[1]) { q[0] = r[i] * r[i]; }, s1, red);
        const double fin = sqrt(s1[0]) + kSmallD;
        ++nIter;
        if (!(nIter < a.maxIter)) break;
        if (wgConv(fin, init, a.tol, a.relTol, nIter)) break;
    }
    for (int i = t; i < n; i += nt) a.corr[i] = x[i];
}

static __global__ void __launch_bounds__(1024) kCoarsePCG(CoarseArgs a, Skip sk) {
    if (sk.on()) return;
    extern __shared__ double lds[];
    const int n = a.A.n;
    coarsePCG(a, CoarseGlobal{a}, lds, lds + n, lds + 2 * n, lds + 3 * n, lds + 4 * n);
}

// LDS bytes of kCoarsePCGLocal for a level of n rows, nnz entries, nf / nb
// forward / backward levels, with or without a forward row list.
inline size_t coarseLocalBytes(int n, int nnz, int nf, int nb, bool fwdRows) {
    const size_t d = sizeof(double) * ((size_t)nnz + 6 * (size_t)n + 64);
    const size_t i = sizeof(int) * ((size_t)n + 1 + (fwdRows ? n : 0) + n + nf + 1 + nb + 1);
    const size_t h = sizeof(uint16_t) * ((size_t)nnz + n);
    return d + i + h + 16;
}

static __global__ void __launch_bounds__(256) kCoarsePCGLocal(CoarseArgs a, CoarseCsr c, Skip sk) {
    if (sk.on()) return;
    extern __shared__ double lds[];
    const int n = a.A.n, nnz = c.nnz, t = threadIdx.x, nt = blockDim.x;
    double* val = lds;
    double* diag = val + nnz;
    double* rD = diag + n;
    double* x = rD + n;
    double* r = x + n;
    double* w = r + n;
    double* p = w + n;
    double* red = p + n;
    int* start = reinterpret_cast<int*>(red + 64);
    int* frows = start + n + 1;
    int* brows = frows + (a.fwdRows ? n : 0);
    int* fs = brows + n;
    int* bs = fs + a.nFwd + 1;
    uint16_t* col = reinterpret_cast<uint16_t*>(bs + a.nBwd + 1);
    uint16_t* mid = col + nnz;
    for (int k = t; k < nnz; k += nt) {
        val[k] = c.val[k];
        col[k] = c.col[k];
    }
    for (int i = t; i < n; i += nt) {
        diag[i] = a.A.diag[i];
        rD[i] = a.rD[i];
        start[i] = c.start[i];
        mid[i] = c.mid[i];
        brows[i] = a.bwdRows[i];
        if (a.fwdRows) frows[i] = a.fwdRows[i];
    }
    if (t == 0) start[n] = c.start[n];
    for (int k = t; k <= a.nFwd; k += nt) fs[k] = a.fwdStart[k];
    for (int k = t; k <= a.nBwd; k += nt) bs[k] = a.bwdStart[k];
    __syncthreads();
    const CoarseLocal m{val, diag, rD, start, col, mid, fs, a.fwdRows ? frows : nullptr, bs, brows};
    coarsePCG(a, m, x, r, w, p, red);
}

// PBiCGStab + DILU (PBiCGStab.cpp:45-459) on the coarsest level.
static __global__ void __launch_bounds__(1024) kCoarseBiCG(CoarseArgs a, Skip sk) {
    if (sk.on()) return;
    extern __shared__ double lds[];
    const int n = a.A.n;
    double* x = lds;
    double* r = x + n;
    double* r0 = r + n;
    double* p = r0 + n;
    double* y = p + n;
    double* Ay = y + n;
    double* s = Ay + n;
    double* z = s + n;
    double* tt = z + n;
    double* red = tt + n;
    const int t = threadIdx.x, nt = blockDim.x;
    double s1[1], s2[2];
    for (int i = t; i < n; i += nt) {
        x[i] = 0.0;
        r[i] = a.src[i];
        r0[i] = r[i];
        p[i] = 0.0;
        Ay[i] = 0.0;
    }
    __syncthreads();
    wgReduce<1>(a, n, [&](int i, double(&q)[1]) { q[0] = r[i] * r[i]; }, s1, red);
    const double init = sqrt(s1[0]) + kSmallD;
    double rho = 0, alpha = 0, omega = 0;
    long long nIter = 0;
    for (;;) {
        const double rhoOld = rho;
        wgReduce<1>(a, n, [&](int i, double(&q)[1]) { q[0] = r0[i] * r[i]; }, s1, red);
        rho = s1[0];
        if (!(fabs(rho) > kSmallD)) break;
        const double beta = (rho / rhoOld) * (alpha / omega);
        for (int i = t; i < n; i += nt)
            p[i] = (nIter == 0) ? r[i] : r[i] + beta * (p[i] - omega * Ay[i]);
        __syncthreads();
        wgPrecond(a, CoarseGlobal{a}, y, p);
        for (int i = t; i < n; i += nt) Ay[i] = rowAmul<double, double>(a.A, i, y);
        __syncthreads();
        wgReduce<1>(a, n, [&](int i, double(&q)[1]) { q[0] = r0[i] * Ay[i]; }, s1, red);
        alpha = rho / s1[0];
        for (int i = t; i < n; i += nt) s[i] = r[i] - alpha * Ay[i];
        __syncthreads();
        wgReduce<1>(a, n, [&](int i, double(&q)[1]) { q[0] = s[i] * s[i]; }, s1, red);
        double fin = sqrt(s1[0]) + kSmallD;
        if (wgConv(fin, init, a.tol, a.relTol, nIter + 1)) {
            for (int i = t; i < n; i += nt) x[i] += alpha * y[i];
            __syncthreads();
            break;
        }
        wgPrecond(a, CoarseGlobal{a}, z, s);
        for (int i = t; i < n; i += nt) tt[i] = rowAmul<double, double>(a.A, i, z);
        __syncthreads();
        wgReduce<2>(a, n, [&](int i, double(&q)[2]) {
            q[0] = tt[i] * tt[i];
            q[1] = tt[i] * s[i];
        }, s2, red);
        omega = s2[1] / s2[0];
        for (int i = t; i < n; i += nt) {
            x[i] += alpha * y[i] + omega * z[i];
            r[i] = s[i] - omega * tt[i];
        }
        __syncthreads();
        wgReduce<1>(a, n, [&](int i, double(&q)[1]) { q[0] = r[i] * r[i]; }, s1, red);
        fin = sqrt(s1[0]) + kSmallD;
        ++nIter;
        if (!(nIter < a.maxIter)) break;
        if (wgConv(fin, init, a.tol, a.relTol, nIter)) break;
    }
    for (int i = t; i < n; i += nt) a.corr[i] = x[i];
}

}  // namespace unap_hip
