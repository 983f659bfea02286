// batch.hip -- many independent LDU systems solved in one persistent launch.
//
// One workgroup runs one system's whole solve: PCG with a DIC preconditioner
// (PCG.cpp:35-183, lduDICPrecond.cpp:45-77) or PBiCGStab with DILU
// (PBiCGStab.cpp:45-459, lduDILUPrecond.cpp:45-86).  The Krylov loop, its dot
// products and norms run in FP64; the preconditioner sweeps run in FP32 over
// the system's level schedule with workgroup barriers between levels.
// Workgroups take systems from an atomic ticket in a largest-first order
// (longest-processing-time scheduling), and exit when the queue is empty.
// The block size, workgroups per CU and rows per wavefront pass are picked by
// a small autotuner that times whole batch solves.
#include "unap_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "device.hpp"
#include "ops.hpp"

namespace unap_hip {

struct BatchSys {
    int n, nf, type, nFwd, nBwd;
    int cellOff;            // cells / rows / internal vectors / sweep positions
    long long ellOff;       // ELL entries (n * ellW per system)
    int ellW;               // ELL width: the system's longest row
    int fwdOff, bwdOff;     // level boundaries
    long long tabOff[2];    // sweep-table entries (forward, backward)
    int tabM[2];            // entry slots per position (forward, backward)
};

// A preconditioner sweep in table form.  Positions p = 0 .. n-1 follow the
// sweep's level order; level k holds positions [start[k], start[k+1]).
// Position p names its row and the number of sweep-side entries (row | cnt <<
// 24), the row's rD, and entry j's column and rD * coefficient (FP32 product,
// the one the sweep forms) at j * n + p.  Every address follows from p alone,
// so the next level's entries are loaded before the level barrier.
struct SweepTab {
    const uint32_t* rc;
    const float* rd;
    const int* col;
    const float* coef;
    int n, m;
};
constexpr int kSweepPre = 4;  // entries a thread holds for its next row

struct BatchArgs {
    const BatchSys* sys;
    const int* order;  // queue order (largest first)
    int nsys;
    unsigned* ticket;
    const uint32_t* rowInfo;
    const int* ellCol;
    const double* ellVal;
    const double* diag;
    const int* fwdStart;
    const int* bwdStart;
    const uint32_t* tabRc[2];  // sweep tables (forward, backward)
    const float* tabRd[2];
    const int* tabCol[2];
    const float* tabCoef[2];
    const double* b;
    double* x;
    double* v;  // 8 FP64 work vectors, each of total length
    float* w32;
    long long total;
    double relTol, tol;
    long long maxIter;
    unap_hip_perf* perf;
};

__device__ __forceinline__ double bsum(double v, double* red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double a = 0.0;
    for (int w = 0; w < nw; ++w) a += red[w];
    __syncthreads();
    return a;
}
__device__ __forceinline__ void bsum2(double& v0, double& v1, double* red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        v0 += __shfl_xor(v0, o, 64);
        v1 += __shfl_xor(v1, o, 64);
    }
    if (lane == 0) {
        red[wave] = v0;
        red[16 + wave] = v1;
    }
    __syncthreads();
    double a = 0.0, c = 0.0;
    for (int w = 0; w < nw; ++w) {
        a += red[w];
        c += red[16 + w];
    }
    __syncthreads();
    v0 = a;
    v1 = c;
}

// The system matrix in ELL form: entry k of row i at k * n + i (rowAmul's
// entry order), rows shorter than w padded.  Every entry address follows from
// the row alone, so a row's loads start together instead of after its slice
// start and row length arrive.
struct EllMat {
    const int* col;
    const double* val;
    const uint32_t* info;  // len | nLowerSide << 16
    const double* diag;
    int n, w;
};
constexpr int kEllBatch = 8;

// rowAmul's arithmetic: acc = 0 + d x_i, then + v_k x[c_k] in entry order.
__device__ __forceinline__ double ellAmul(const EllMat& A, int i, const double* x) {
    const int len = (int)(A.info[i] & 0xffffu);
    double acc = 0.0 + A.diag[i] * x[i];
    for (int k0 = 0; k0 < A.w; k0 += kEllBatch) {
        int c[kEllBatch];
        double v[kEllBatch], xv[kEllBatch];
#pragma unroll
        for (int j = 0; j < kEllBatch; ++j) {
            const size_t e = (size_t)min(k0 + j, A.w - 1) * A.n + i;  // past the width: re-read, not added
            c[j] = A.col[e];
            v[j] = A.val[e];
        }
#pragma unroll
        for (int j = 0; j < kEllBatch; ++j) xv[j] = x[c[j]];
#pragma unroll
        for (int j = 0; j < kEllBatch; ++j)
            if (k0 + j < len) acc = acc + v[j] * xv[j];
    }
    return acc;
}

struct SysView {
    EllMat A;
    SweepTab fwd, bwd;
    const int* fs;
    const int* bs;
    int nFwd, nBwd, n;
};

struct SweepRow {
    int q, cnt;
    float rd;
    int c[kSweepPre];
    float v[kSweepPre];
};

__device__ __forceinline__ void loadSweepRow(SweepRow& R, const SweepTab& T, int p) {
    const uint32_t rc = T.rc[p];
    R.q = (int)(rc & 0xffffffu);
    R.cnt = (int)(rc >> 24);
    R.rd = T.rd[p];
#pragma unroll
    for (int j = 0; j < kSweepPre; ++j)
        if (j < T.m) {
            R.c[j] = T.col[(size_t)j * T.n + p];
            R.v[j] = T.coef[(size_t)j * T.n + p];
        }
}

// acc -= coef_j * w[col_j] in entry order (the DIC/DILU row arithmetic of
// fwdRow / bwdRow: coef_j is the FP32 product rD * coefficient).
__device__ __forceinline__ float sweepRow(const SweepRow& R, const SweepTab& T, int p, float acc, const float* w) {
    float x[kSweepPre];
#pragma unroll
    for (int j = 0; j < kSweepPre; ++j)
        if (j < T.m) x[j] = w[R.c[j]];
#pragma unroll
    for (int j = 0; j < kSweepPre; ++j)
        if (j < R.cnt) acc = acc - R.v[j] * x[j];
    for (int j = kSweepPre; j < R.cnt; ++j) acc = acc - T.coef[(size_t)j * T.n + p] * w[T.col[(size_t)j * T.n + p]];
    return acc;
}

// One sweep over the levels [start[0], start[nl]).  Forward rows start from
// rD * r (r read in FP64), backward rows from their forward value.
template <bool Fwd>
__device__ __forceinline__ void batchSweep(const SweepTab& T, const int* start, int nl, const double* r, float* w) {
    const int nt = (int)blockDim.x;
    SweepRow R;
    int p = start[0] + (int)threadIdx.x;
    if (nl > 0 && p < start[1]) loadSweepRow(R, T, p);
    for (int k = 0; k < nl; ++k) {
        const int e = start[k + 1];
        const int pn = e + (int)threadIdx.x;
        SweepRow Rn;
        if (k + 1 < nl && pn < start[k + 2]) loadSweepRow(Rn, T, pn);
        if (p < e) {
            w[R.q] = sweepRow(R, T, p, Fwd ? R.rd * (float)r[R.q] : w[R.q], w);
            for (int p2 = p + nt; p2 < e; p2 += nt) {
                SweepRow R2;
                loadSweepRow(R2, T, p2);
                w[R2.q] = sweepRow(R2, T, p2, Fwd ? R2.rd * (float)r[R2.q] : w[R2.q], w);
            }
        }
        __syncthreads();
        R = Rn;
        p = pn;
    }
}

// w32 = M^-1 r in FP32 with the level schedule (DIC or DILU arithmetic of
// fwdRow / bwdRow, here reading r in FP64).
__device__ __forceinline__ void batchPrecond(const SysView& s, const double* r, float* w) {
    batchSweep<true>(s.fwd, s.fs, s.nFwd, r, w);
    batchSweep<false>(s.bwd, s.bs, s.nBwd, r, w);
}

// Vector loops over rows i = t, t + nt, ... (< n), VR rows at a time (a
// wavefront covers 64 VR rows a pass; VR is one of the autotuned launch
// parameters): the rows' loads first (rows past n re-read row n - 1 and are
// not used), then their work in row order, the same order as one row at a
// time, so the sums do not depend on VR.  Stores between rows otherwise keep
// each row's loads from starting before the previous row is done.
template <int VR, class V, class Load, class Use>
__device__ __forceinline__ void vecRows(int t, int n, int nt, Load&& load, Use&& use) {
    for (int i0 = t; i0 < n; i0 += VR * nt) {
        V v[VR];
#pragma unroll
        for (int u = 0; u < VR; ++u) v[u] = load(min(i0 + u * nt, n - 1));
#pragma unroll
        for (int u = 0; u < VR; ++u)
            if (i0 + u * nt < n) use(i0 + u * nt, v[u]);
    }
}
struct V1 {
    double a;
};
struct V2 {
    double a, b;
};
struct V3 {
    double a, b, c;
};
struct V4 {
    double a, b, c, d;
};
struct V5 {
    double a, b, c, d, e;
};

// Shadow-residual restart threshold of the FP32-preconditioned BiCGStab.
constexpr double kBicgRestart = 1.0e-6;

__device__ __forceinline__ bool bconv(double fin, double init, double tol, double relTol) {
    return fin < tol || (relTol > kSmallD && fin <= relTol * init);
}

template <int VR>
__device__ void batchPCG(const BatchArgs& a, const SysView& s, int off, unap_hip_perf* pf, double* red) {
    const int n = s.n, t = threadIdx.x, nt = blockDim.x;
    double* x = a.x + off;
    const double* b = a.b + off;
    double* r = a.v + off;
    double* p = a.v + a.total + off;
    double* w = a.v + 2 * a.total + off;  // q = A p, kept into the next iteration's beta
    float* w32 = a.w32 + off;
    double acc = 0.0;
    for (int i = t; i < n; i += nt) {
        const double ri = b[i] - ellAmul(s.A, i, x);
        r[i] = ri;
        p[i] = 0.0;
        w[i] = 0.0;
        acc += ri * ri;
    }
    const double init = sqrt(bsum(acc, red)) + kSmallD;
    double fin = init, prev = init, rho = 1.0e37, pq = 1.0;
    long long it = 0;
    int conv = 0, sing = 0;
    // Flexible CG (Notay 2000, FCG(1)): beta = -(z_k . q_k-1) / (p_k-1 . q_k-1)
    // makes each direction A-orthogonal to the previous one for a
    // preconditioner evaluated in FP32 (not exactly linear or symmetric).
    for (;;) {
        batchPrecond(s, r, w32);
        double rz = 0.0, zq = 0.0;
        vecRows<VR, V3>(t, n, nt, [&](int i) { return V3{(double)w32[i], r[i], w[i]}; },
                    [&](int, const V3& v) {
                        rz += v.a * v.b;
                        zq += v.a * v.c;
                    });
        bsum2(rz, zq, red);
        rho = rz;
        if (!(fabs(rho) > kSmallD)) {
            sing = 1;
            break;
        }
        const double beta = -zq / pq;
        vecRows<VR, V2>(t, n, nt, [&](int i) { return V2{(double)w32[i], p[i]}; },
                    [&](int i, const V2& v) { p[i] = (it == 0) ? v.a : v.a + beta * v.b; });
        __syncthreads();
        acc = 0.0;
        for (int i = t; i < n; i += nt) {
            const double v = ellAmul(s.A, i, p);
            w[i] = v;
            acc += v * p[i];
        }
        pq = bsum(acc, red);
        const double alpha = rho / pq;
        acc = 0.0;
        vecRows<VR, V4>(t, n, nt, [&](int i) { return V4{x[i], p[i], r[i], w[i]}; },
                    [&](int i, const V4& v) {
                        x[i] = v.a + alpha * v.b;
                        const double ri = v.c - alpha * v.d;
                        r[i] = ri;
                        acc += ri * ri;
                    });
        prev = fin;
        fin = sqrt(bsum(acc, red)) + kSmallD;
        ++it;
        if (!(it < a.maxIter)) break;
        if (bconv(fin, init, a.tol, a.relTol)) {
            conv = 1;
            break;
        }
    }
    if (t == 0) {
        pf->initial_residual = init;
        pf->final_residual = fin;
        pf->previous_residual = prev;
        pf->n_iterations = it;
        pf->converged = conv;
        pf->singular = sing;
    }
}

template <int VR>
__device__ void batchBiCG(const BatchArgs& a, const SysView& s, int off, unap_hip_perf* pf, double* red) {
    const int n = s.n, t = threadIdx.x, nt = blockDim.x;
    const long long T = a.total;
    double* x = a.x + off;
    const double* b = a.b + off;
    double* r = a.v + off;
    double* r0 = a.v + T + off;
    double* p = a.v + 2 * T + off;
    double* Ay = a.v + 3 * T + off;
    double* sv = a.v + 4 * T + off;
    double* t_ = a.v + 5 * T + off;
    double* y = a.v + 6 * T + off;
    double* z = a.v + 7 * T + off;
    float* w32 = a.w32 + off;
    double acc = 0.0;
    for (int i = t; i < n; i += nt) {
        const double ri = b[i] - ellAmul(s.A, i, x);
        r[i] = ri;
        r0[i] = ri;
        p[i] = 0.0;
        Ay[i] = 0.0;
        acc += ri * ri;
    }
    const double init = sqrt(bsum(acc, red)) + kSmallD;
    double fin = init, prev = init, rho = 0, alpha = 0, omega = 0;
    double r0norm = init;
    long long it = 0;
    int k = 0;  // iterations since the shadow residual was (re)set
    int conv = 0, sing = 0;
    for (;;) {
        const double rhoOld = rho;
        acc = 0.0;
        vecRows<VR, V2>(t, n, nt, [&](int i) { return V2{r0[i], r[i]}; }, [&](int, const V2& v) { acc += v.a * v.b; });
        rho = bsum(acc, red);
        // The FP32 preconditioner is not an exactly linear operator, and the
        // shadow residual loses bi-orthogonality with it: iteration counts
        // then scatter (73 .. >400 for a system FP64 solves in 53) under
        // last-bit changes of one dot product.  Restarting with r0 = r once
        // |r0.r| < 1e-6 |r0| |r| keeps them at the FP64 count (51-60).
        if (k > 0 && fabs(rho) < kBicgRestart * r0norm * fin) {
            acc = 0.0;
            for (int i = t; i < n; i += nt) {
                r0[i] = r[i];
                acc += r[i] * r[i];
            }
            rho = bsum(acc, red);
            r0norm = sqrt(rho);
            k = 0;
        }
        if (!(fabs(rho) > kSmallD)) {
            sing = 1;
            break;
        }
        const double beta = (rho / rhoOld) * (alpha / omega);
        vecRows<VR, V3>(t, n, nt, [&](int i) { return V3{r[i], p[i], Ay[i]}; },
                    [&](int i, const V3& v) { p[i] = (k == 0) ? v.a : v.a + beta * (v.b - omega * v.c); });
        __syncthreads();
        batchPrecond(s, p, w32);
        vecRows<VR, V1>(t, n, nt, [&](int i) { return V1{(double)w32[i]}; }, [&](int i, const V1& v) { y[i] = v.a; });
        __syncthreads();
        acc = 0.0;
        for (int i = t; i < n; i += nt) {
            const double v = ellAmul(s.A, i, y);
            Ay[i] = v;
            acc += r0[i] * v;
        }
        alpha = rho / bsum(acc, red);
        acc = 0.0;
        vecRows<VR, V2>(t, n, nt, [&](int i) { return V2{r[i], Ay[i]}; },
                    [&](int i, const V2& v) {
                        const double si = v.a - alpha * v.b;
                        sv[i] = si;
                        acc += si * si;
                    });
        fin = sqrt(bsum(acc, red)) + kSmallD;
        if (bconv(fin, init, a.tol, a.relTol)) {
            for (int i = t; i < n; i += nt) x[i] += alpha * y[i];
            ++it;
            conv = 1;
            break;
        }
        batchPrecond(s, sv, w32);
        vecRows<VR, V1>(t, n, nt, [&](int i) { return V1{(double)w32[i]}; }, [&](int i, const V1& v) { z[i] = v.a; });
        __syncthreads();
        double tt = 0.0, ts = 0.0;
        for (int i = t; i < n; i += nt) {
            const double v = ellAmul(s.A, i, z);
            t_[i] = v;
            tt += v * v;
            ts += v * sv[i];
        }
        bsum2(tt, ts, red);
        omega = ts / tt;
        acc = 0.0;
        vecRows<VR, V5>(t, n, nt, [&](int i) { return V5{y[i], z[i], sv[i], t_[i], x[i]}; },
                    [&](int i, const V5& v) {
                        x[i] = v.e + (alpha * v.a + omega * v.b);
                        const double ri = v.c - omega * v.d;
                        r[i] = ri;
                        acc += ri * ri;
                    });
        prev = fin;
        fin = sqrt(bsum(acc, red)) + kSmallD;
        ++it;
        ++k;
        if (!(it < a.maxIter)) break;
        if (bconv(fin, init, a.tol, a.relTol)) {
            conv = 1;
            break;
        }
    }
    if (t == 0) {
        pf->initial_residual = init;
        pf->final_residual = fin;
        pf->previous_residual = prev;
        pf->n_iterations = it;
        pf->converged = conv;
        pf->singular = sing;
    }
}

template <int BS, int VR>
__global__ void __launch_bounds__(BS) kBatchSolve(BatchArgs a) {
    __shared__ double red[32];
    __shared__ int ticket;
    for (;;) {
        if (threadIdx.x == 0) ticket = (int)atomicAdd(a.ticket, 1u);
        __syncthreads();
        const int q = ticket;
        __syncthreads();
        if (q >= a.nsys) return;  // the queue is empty: every workgroup ends here
        const BatchSys d = a.sys[a.order[q]];
        SysView s;
        s.n = d.n;
        s.A = {a.ellCol + d.ellOff, a.ellVal + d.ellOff, a.rowInfo + d.cellOff, a.diag + d.cellOff, d.n, d.ellW};
        SweepTab* tabs[2] = {&s.fwd, &s.bwd};
        for (int dir = 0; dir < 2; ++dir)
            *tabs[dir] = {a.tabRc[dir] + d.cellOff, a.tabRd[dir] + d.cellOff, a.tabCol[dir] + d.tabOff[dir],
                          a.tabCoef[dir] + d.tabOff[dir], d.n, d.tabM[dir]};
        s.fs = a.fwdStart + d.fwdOff;
        s.bs = a.bwdStart + d.bwdOff;
        s.nFwd = d.nFwd;
        s.nBwd = d.nBwd;
        if (d.type == UNAP_HIP_SOLVER_PCG)
            batchPCG<VR>(a, s, d.cellOff, a.perf + a.order[q], red);
        else
            batchBiCG<VR>(a, s, d.cellOff, a.perf + a.order[q], red);
        __syncthreads();
    }
}

__global__ void kResetTicket(unsigned* t) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *t = 0u;
}

}  // namespace unap_hip

using namespace unap_hip;

struct unap_hip_batch_s {
    int nsys = 0;
    long long total = 0;
    std::vector<BatchSys> sysH;
    std::vector<int> permH;       // global internal -> global external cell
    DBuf<BatchSys> sys;
    DBuf<int> order;
    DBuf<unsigned> ticket;
    DBuf<uint32_t> rowInfo;
    DBuf<int> ellCol, fwdStart, bwdStart, perm;
    DBuf<double> ellVal, diag, x, b, v;
    DBuf<float> w32;
    DBuf<uint32_t> tabRc[2];
    DBuf<float> tabRd[2], tabCoef[2];
    DBuf<int> tabCol[2];
    DBuf<unap_hip_perf> perf;
    int blockSize = 0, wgPerCU = 0;  // 0 = not tuned yet
    int rowsPerThread = 4;           // vector-loop rows a thread (64 x this rows a wavefront pass)
    double lastMs = 0;
    int cus = 0;
};

namespace {

template <class F>
int bguard(F&& f) {
    try {
        f();
        setLastError("");
        return UNAP_HIP_OK;
    } catch (const Error& e) {
        setLastError(e.what());
        return e.code;
    } catch (const std::exception& e) {
        setLastError(e.what());
        return UNAP_HIP_ERR_INTERNAL;
    }
}

template <int BS>
void launchShape(const BatchArgs& a, int grid, int vr, hipStream_t st) {
    switch (vr) {
        case 1: kBatchSolve<BS, 1><<<grid, BS, 0, st>>>(a); break;
        case 2: kBatchSolve<BS, 2><<<grid, BS, 0, st>>>(a); break;
        default: kBatchSolve<BS, 4><<<grid, BS, 0, st>>>(a); break;
    }
}

void launchBatch(unap_hip_batch_s& B, double relTol, double tol, long long maxIter, int bs, int wgPerCU, int vr) {
    Ctx& c = Ctx::get();
    BatchArgs a;
    a.sys = B.sys.get();
    a.order = B.order.get();
    a.nsys = B.nsys;
    a.ticket = B.ticket.get();
    a.rowInfo = B.rowInfo.get();
    a.ellCol = B.ellCol.get();
    a.ellVal = B.ellVal.get();
    a.diag = B.diag.get();
    for (int dir = 0; dir < 2; ++dir) {
        a.tabRc[dir] = B.tabRc[dir].get();
        a.tabRd[dir] = B.tabRd[dir].get();
        a.tabCol[dir] = B.tabCol[dir].get();
        a.tabCoef[dir] = B.tabCoef[dir].get();
    }
    a.fwdStart = B.fwdStart.get();
    a.bwdStart = B.bwdStart.get();
    a.b = B.b.get();
    a.x = B.x.get();
    a.v = B.v.get();
    a.w32 = B.w32.get();
    a.total = B.total;
    a.relTol = relTol;
    a.tol = tol;
    a.maxIter = maxIter;
    a.perf = B.perf.get();
    const int grid = std::max(1, std::min(B.nsys, B.cus * wgPerCU));
    kResetTicket<<<1, 1, 0, c.stream>>>(B.ticket.get());
    UH_LAUNCHED();
    switch (bs) {
        case 256: launchShape<256>(a, grid, vr, c.stream); break;
        case 512: launchShape<512>(a, grid, vr, c.stream); break;
        default: launchShape<1024>(a, grid, vr, c.stream); break;
    }
    UH_LAUNCHED();
}

}  // namespace

extern "C" {

int unap_hip_batch_create(int64_t nsys, const int64_t* cellOff, const int64_t* faceOff, const int64_t* la,
                          const int64_t* ua, const double* lower, const double* diag, const double* upper,
                          const int32_t* solver, unap_hip_batch* out) {
    return bguard([&] {
        if (!(nsys > 0 && nsys < (1LL << 31) && cellOff && faceOff && la && ua && diag && upper && solver && out))
            throw Error(UNAP_HIP_ERR_ARG, "null or empty batch argument");
        // The device arrays hold cellOff[nsys] cells and faceOff[nsys] faces
        // from offset 0, indexed with 32-bit ints.
        if (cellOff[0] != 0 || faceOff[0] != 0)
            throw Error(UNAP_HIP_ERR_ARG, "cell_off[0] and face_off[0] must be 0");
        for (int64_t s = 0; s < nsys; ++s)
            if (cellOff[s + 1] <= cellOff[s] || faceOff[s + 1] < faceOff[s])
                throw Error(UNAP_HIP_ERR_ARG, "batch offsets must increase (cells) and not decrease (faces)");
        if (cellOff[nsys] >= (1LL << 31) || faceOff[nsys] >= (1LL << 31))
            throw Error(UNAP_HIP_ERR_ARG, "a batch holds fewer than 2^31 cells and 2^31 faces");
        Ctx& c = Ctx::get();
        auto* B = new unap_hip_batch_s;
        std::unique_ptr<unap_hip_batch_s> guardB(B);
        B->nsys = (int)nsys;
        std::vector<uint32_t> infoAll;
        std::vector<int> ellColAll, fwdAll, bwdAll, permAll;
        std::vector<double> ellValAll, diagAll;
        std::vector<uint32_t> tRc[2];
        std::vector<float> tRd[2], tCoef[2];
        std::vector<int> tCol[2];
        std::vector<double> work(nsys);
        for (int64_t s = 0; s < nsys; ++s) {
            const int64_t n = cellOff[s + 1] - cellOff[s], nf = faceOff[s + 1] - faceOff[s];
            if (n <= 0 || nf < 0) throw Error(UNAP_HIP_ERR_ARG, "empty system in batch");
            if (n >= (1 << 24)) throw Error(UNAP_HIP_ERR_ARG, "batch systems hold fewer than 2^24 cells");
            if (solver[s] != UNAP_HIP_SOLVER_PCG && solver[s] != UNAP_HIP_SOLVER_PBICGSTAB)
                throw Error(UNAP_HIP_ERR_ARG, "batch systems use PCG (DIC) or PBiCGStab (DILU)");
            HostLdu A;
            A.n = (int)n;
            A.l.resize(nf);
            A.u.resize(nf);
            for (int64_t f = 0; f < nf; ++f) {
                const int64_t lf = la[faceOff[s] + f], uf = ua[faceOff[s] + f];
                if (lf < 0 || lf >= n || uf < 0 || uf >= n) throw Error(UNAP_HIP_ERR_ARG, "face address out of range");
                A.l[f] = (int)lf;
                A.u[f] = (int)uf;
            }
            A.d.assign(diag + cellOff[s], diag + cellOff[s] + n);
            A.up.assign(upper + faceOff[s], upper + faceOff[s] + nf);
            A.symm = (lower == nullptr || lower == upper);
            if (!A.symm) A.lo.assign(lower + faceOff[s], lower + faceOff[s] + nf);
            Layout L = buildLayout(A, 1);  // breadth-first rows: the Amul gathers stay local
            if (!L.sweepable)
                throw Error(UNAP_HIP_ERR_FACE_ORDER, "batch system faces must be owner-sorted with owner < neighbour");
            BatchSys d;
            d.n = (int)n;
            d.nf = (int)nf;
            d.type = solver[s];
            d.nFwd = (int)L.fwdStart.size() - 1;
            d.nBwd = (int)L.bwdStart.size() - 1;
            d.cellOff = (int)cellOff[s];
            d.ellOff = (long long)ellColAll.size();
            d.ellW = 0;
            for (int i = 0; i < (int)n; ++i) d.ellW = std::max(d.ellW, (int)(L.rowInfo[i] & 0xffffu));
            d.fwdOff = (int)fwdAll.size();
            d.bwdOff = (int)bwdAll.size();
            const bool dic = (solver[s] == UNAP_HIP_SOLVER_PCG);
            std::vector<float> vp;
            fillSellValues<float>(L, A, dic, vp);
            const std::vector<double> rde = dic ? dicReciprocalD(A) : diluReciprocalD(A);
            std::vector<float> rd32(n);
            for (int i = 0; i < (int)n; ++i) rd32[i] = (float)rde[L.perm[i]];
            // sweep tables: forward over the level-major rows, entries in
            // losort order; backward over bwdRows, owner entries descending
            for (int dir = 0; dir < 2; ++dir) {
                auto rowAt = [&](int p) {
                    if (dir == 1) return L.bwdRows[p];
                    return L.fwdRows.empty() ? p : L.fwdRows[p];
                };
                auto entries = [&](int q, int& k0, int& cnt, int& step) {
                    const int len = (int)(L.rowInfo[q] & 0xffffu), mid = (int)(L.rowInfo[q] >> 16);
                    k0 = dir == 0 ? 0 : len - 1;
                    cnt = dir == 0 ? mid : len - mid;
                    step = dir == 0 ? 1 : -1;
                };
                int M = 0;
                for (int p = 0; p < (int)n; ++p) {
                    int k0, cnt, step;
                    entries(rowAt(p), k0, cnt, step);
                    M = std::max(M, cnt);
                }
                if (M > 255) throw Error(UNAP_HIP_ERR_ARG, "batch rows hold at most 255 faces a side");
                d.tabOff[dir] = (long long)tCol[dir].size();
                d.tabM[dir] = M;
                const size_t o = tCol[dir].size();
                tCol[dir].resize(o + (size_t)M * n);
                tCoef[dir].resize(o + (size_t)M * n, 0.f);
                for (int p = 0; p < (int)n; ++p) {
                    const int q = rowAt(p);
                    int k0, cnt, step;
                    entries(q, k0, cnt, step);
                    tRc[dir].push_back((uint32_t)q | ((uint32_t)cnt << 24));
                    tRd[dir].push_back(rd32[q]);
                    const long long base = L.sliceStart[q >> kSliceShift] + (q & (kSliceRows - 1));
                    for (int j = 0; j < M; ++j) {
                        const size_t at = o + (size_t)j * n + p;
                        if (j < cnt) {
                            const long long e = base + ((long long)(k0 + j * step) << kSliceShift);
                            tCol[dir][at] = L.col[e];
                            tCoef[dir][at] = rd32[q] * vp[e];  // the FP32 product the sweep forms
                        } else {
                            tCol[dir][at] = q;  // padding: a valid address, never added
                        }
                    }
                }
            }
            B->sysH.push_back(d);
            infoAll.insert(infoAll.end(), L.rowInfo.begin(), L.rowInfo.end());
            fwdAll.insert(fwdAll.end(), L.fwdStart.begin(), L.fwdStart.end());
            bwdAll.insert(bwdAll.end(), L.bwdStart.begin(), L.bwdStart.end());
            for (int i = 0; i < (int)n; ++i) permAll.push_back((int)cellOff[s] + L.perm[i]);
            std::vector<double> v;
            fillSellValues<double>(L, A, false, v);
            // ELL copy of the SELL rows: entry k of row i at k * n + i; rows
            // shorter than the width point their padding at themselves
            const size_t eo = ellColAll.size();
            ellColAll.resize(eo + (size_t)d.ellW * n);
            ellValAll.resize(eo + (size_t)d.ellW * n, 0.0);
            for (int i = 0; i < (int)n; ++i) {
                const int len = (int)(L.rowInfo[i] & 0xffffu);
                const long long base = L.sliceStart[i >> kSliceShift] + (i & (kSliceRows - 1));
                for (int k = 0; k < d.ellW; ++k) {
                    const size_t at = eo + (size_t)k * n + i;
                    if (k < len) {
                        const long long e = base + ((long long)k << kSliceShift);
                        ellColAll[at] = L.col[e];
                        ellValAll[at] = v[e];
                    } else {
                        ellColAll[at] = i;
                    }
                }
            }
            std::vector<double> dd;
            fillDiag<double>(L, A, dd);
            diagAll.insert(diagAll.end(), dd.begin(), dd.end());
            work[s] = (double)(n + 2 * nf) * std::cbrt((double)n) * (dic ? 1.0 : 2.0);
        }
        B->total = cellOff[nsys];
        std::vector<int> ord(nsys);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return work[a] > work[b]; });
        B->sys.upload(B->sysH, c.stream);
        B->order.upload(ord, c.stream);
        B->ticket.alloc(1);
        B->rowInfo.upload(infoAll, c.stream);
        if (ellColAll.empty()) {  // diagonal systems only: keep the pointers valid
            ellColAll.push_back(0);
            ellValAll.push_back(0.0);
        }
        B->ellCol.upload(ellColAll, c.stream);
        B->fwdStart.upload(fwdAll, c.stream);
        B->bwdStart.upload(bwdAll, c.stream);
        for (int dir = 0; dir < 2; ++dir) {
            if (tCol[dir].empty()) {  // diagonal systems only: keep the pointers valid
                tCol[dir].push_back(0);
                tCoef[dir].push_back(0.f);
            }
            B->tabRc[dir].upload(tRc[dir], c.stream);
            B->tabRd[dir].upload(tRd[dir], c.stream);
            B->tabCol[dir].upload(tCol[dir], c.stream);
            B->tabCoef[dir].upload(tCoef[dir], c.stream);
        }
        B->perm.upload(permAll, c.stream);
        B->ellVal.upload(ellValAll, c.stream);
        B->diag.upload(diagAll, c.stream);
        B->x.alloc(B->total);
        B->b.alloc(B->total);
        B->v.alloc(8 * B->total);
        B->w32.alloc(B->total);
        B->perf.alloc(nsys);
        c.sync();
        int dev = 0;
        UH_CHECK(hipGetDevice(&dev));
        hipDeviceProp_t p;
        UH_CHECK(hipGetDeviceProperties(&p, dev));
        B->cus = p.multiProcessorCount;
        *out = guardB.release();
    });
}

int unap_hip_batch_destroy(unap_hip_batch B) {
    return bguard([&] { delete B; });
}

int unap_hip_batch_solve(unap_hip_batch B, double* x, const double* b, int where, int64_t maxIter, double relTol,
                         double tol, unap_hip_perf* perf) {
    return bguard([&] {
        if (!(B && x && b)) throw Error(UNAP_HIP_ERR_ARG, "null argument");
        if (maxIter < 0) throw Error(UNAP_HIP_ERR_ARG, "negative iteration limit");
        Ctx& c = Ctx::get();
        const long long T = B->total;
        DBuf<double> stage;
        const double* bs = b;
        const double* xs = x;
        if (where == UNAP_HIP_DEVICE) {
            UH_CHECK(hipDeviceSynchronize());
        } else {
            stage.alloc(2 * T);
            stage.upload(b, T, c.stream);
            UH_CHECK(hipMemcpyAsync(stage.get() + T, x, sizeof(double) * T, hipMemcpyHostToDevice, c.stream));
            bs = stage.get();
            xs = stage.get() + T;
        }
        launchRows(BGather<double, double>{B->perm.get(), bs, B->b.get()}, (int)T);
        launchRows(BGather<double, double>{B->perm.get(), xs, B->x.get()}, (int)T);
        // x0 is kept to restart from it while tuning
        DBuf<double> x0(T);
        UH_CHECK(hipMemcpyAsync(x0.get(), B->x.get(), sizeof(double) * T, hipMemcpyDeviceToDevice, c.stream));
        hipEvent_t e0, e1;
        UH_CHECK(hipEventCreate(&e0));
        UH_CHECK(hipEventCreate(&e1));
        auto timed = [&](int bsz, int wg, int vr) {
            UH_CHECK(hipMemcpyAsync(B->x.get(), x0.get(), sizeof(double) * T, hipMemcpyDeviceToDevice, c.stream));
            UH_CHECK(hipEventRecord(e0, c.stream));
            launchBatch(*B, relTol, tol, maxIter, bsz, wg, vr);
            UH_CHECK(hipEventRecord(e1, c.stream));
            UH_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            UH_CHECK(hipEventElapsedTime(&ms, e0, e1));
            return (double)ms;
        };
        if (B->blockSize == 0) {
            // autotune, one full batch solve a candidate: block shape x
            // workgroups per CU at 4 rows a thread, then the rows a thread
            // (1, 2, 4: 64, 128, 256 rows a wavefront pass) at the best shape
            double best = 1e300;
            B->rowsPerThread = 4;
            for (int bsz : {256, 512, 1024})
                for (int wg : {1, 2, 4}) {
                    if (bsz * wg > 2048) continue;
                    const double ms = timed(bsz, wg, 4);
                    if (ms < best) {
                        best = ms;
                        B->blockSize = bsz;
                        B->wgPerCU = wg;
                    }
                }
            for (int vr : {1, 2}) {
                const double ms = timed(B->blockSize, B->wgPerCU, vr);
                if (ms < best) {
                    best = ms;
                    B->rowsPerThread = vr;
                }
            }
        }
        B->lastMs = timed(B->blockSize, B->wgPerCU, B->rowsPerThread);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (where == UNAP_HIP_DEVICE) {
            launchRows(BScatter<double, double>{B->perm.get(), B->x.get(), x}, (int)T);
        } else {
            launchRows(BScatter<double, double>{B->perm.get(), B->x.get(), stage.get()}, (int)T);
            UH_CHECK(hipMemcpyAsync(x, stage.get(), sizeof(double) * T, hipMemcpyDeviceToHost, c.stream));
        }
        if (perf)
            UH_CHECK(hipMemcpyAsync(perf, B->perf.get(), sizeof(unap_hip_perf) * B->nsys, hipMemcpyDeviceToHost,
                                    c.stream));
        c.sync();
    });
}

int unap_hip_batch_last_ms(unap_hip_batch B, double* ms) {
    return bguard([&] {
        if (!(B && ms)) throw Error(UNAP_HIP_ERR_ARG, "null argument");
        *ms = B->lastMs;
    });
}

int unap_hip_batch_config(unap_hip_batch B, int64_t* block_size, int64_t* wg_per_cu, int64_t* rows_per_wavefront) {
    return bguard([&] {
        if (!B) throw Error(UNAP_HIP_ERR_ARG, "null argument");
        if (block_size) *block_size = B->blockSize;
        if (wg_per_cu) *wg_per_cu = B->wgPerCU;
        if (rows_per_wavefront) *rows_per_wavefront = 64 * B->rowsPerThread;
    });
}

int unap_hip_batch_set_config(unap_hip_batch B, int64_t block_size, int64_t wg_per_cu, int64_t rows_per_wavefront) {
    return bguard([&] {
        if (!B) throw Error(UNAP_HIP_ERR_ARG, "null argument");
        if (!(block_size == 256 || block_size == 512 || block_size == 1024))
            throw Error(UNAP_HIP_ERR_ARG, "block_size must be 256, 512 or 1024");
        if (wg_per_cu < 1 || block_size * wg_per_cu > 2048)
            throw Error(UNAP_HIP_ERR_ARG, "wg_per_cu must be >= 1 with block_size * wg_per_cu <= 2048");
        if (!(rows_per_wavefront == 64 || rows_per_wavefront == 128 || rows_per_wavefront == 256))
            throw Error(UNAP_HIP_ERR_ARG, "rows_per_wavefront must be 64, 128 or 256");
        B->blockSize = (int)block_size;
        B->wgPerCU = (int)wg_per_cu;
        B->rowsPerThread = (int)(rows_per_wavefront / 64);
    });
}

}  // extern "C"
