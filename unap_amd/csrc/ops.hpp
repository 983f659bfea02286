// ops.hpp -- the fused row bodies of the solve path and the level-scheduled
// sweep, transfer and coarsest-level kernels.  Every body cites the reference
// loop whose per-element arithmetic it keeps.
#pragma once

#include "kernels.hpp"

namespace unap_hip {

// ============================================================ permutations
// internal <- external (gather) and external <- internal (scatter), with casts
template <typename TI, typename TO>
struct BGather {
    const int* perm;  // internal -> external
    const TI* in;
    TO* out;
    __device__ void operator()(int i) const { out[i] = (TO)in[perm[i]]; }
};
template <typename TI, typename TO>
struct BScatter {
    const int* perm;
    const TI* in;
    TO* out;
    __device__ void operator()(int i) const { out[perm[i]] = (TO)in[i]; }
};
template <typename T>
struct BFill {
    T* out;
    T v;
    __device__ void operator()(int i) const { out[i] = v; }
};
template <typename TI, typename TO>
struct BCopy {
    const TI* in;
    TO* out;
    __device__ void operator()(int i) const { out[i] = (TO)in[i]; }
};

// ================================================================== Amul
// lduMatrixSpMV.cpp:16-98
template <typename T>
struct BAmul {
    DMat<T> A;
    const T* x;
    T* y;
    __device__ void operator()(int r) const { y[r] = rowAmul<T, T>(A, r, x); }
};

// r = b - A x (PCG.cpp:61-67), optional shadow copy r0 = r (PBiCGStab.cpp:76-79),
// sum r*r for normFactor (unapMatrixSolver.cpp:12-16).
template <typename T>
struct BResidual {
    DMat<T> A;
    const T* x;
    const T* b;
    T* r;
    T* r0;
    __device__ void operator()(int i, double (&a)[1]) const {
        const T ax = rowAmul<T, T>(A, i, x);
        const T ri = b[i] - ax;
        r[i] = ri;
        if (r0) r0[i] = ri;
        a[0] += (double)ri * (double)ri;
    }
    __device__ void terms(int i, double (&t)[1]) const { t[0] = (double)r[i] * (double)r[i]; }
};

// ============================================================ PCG bodies
// w = rD * r and w.r (lduDiagPrecond.cpp:56-59 fused with PCG.cpp:105)
struct BDiagDot {
    const double* rD;
    const double* r;
    double* w;
    __device__ void operator()(int i, double (&a)[1]) const {
        const double wi = rD[i] * r[i];
        w[i] = wi;
        a[0] += wi * r[i];
    }
    __device__ void terms(int i, double (&t)[1]) const { t[0] = w[i] * r[i]; }
};
// w = r (identity preconditioner) and w.r
struct BCopyDot {
    const double* r;
    double* w;
    __device__ void operator()(int i, double (&a)[1]) const {
        const double wi = r[i];
        w[i] = wi;
        a[0] += wi * r[i];
    }
    __device__ void terms(int i, double (&t)[1]) const { t[0] = w[i] * r[i]; }
};
struct BDot {
    const double* a;
    const double* b;
    __device__ void operator()(int i, double (&s)[1]) const { s[0] += a[i] * b[i]; }
    __device__ void terms(int i, double (&t)[1]) const { t[0] = a[i] * b[i]; }
};
// PCG.cpp:119-134
struct BPcgP {
    const double* w;
    double* p;
    const KState* st;
    __device__ void operator()(int i) const {
        if (st->nIter == 0)
            p[i] = w[i];
        else
            p[i] = w[i] + st->beta * p[i];
    }
};
// w = A p and w.p (PCG.cpp:137-139)
struct BAmulDot {
    DMat<double> A;
    const double* p;
    double* w;
    __device__ void operator()(int i, double (&a)[1]) const {
        const double wi = rowAmul<double, double>(A, i, p);
        stS(&w[i], wi);
        a[0] += wi * p[i];
    }
    __device__ void terms(int i, double (&t)[1]) const { t[0] = w[i] * p[i]; }
};
// x += alpha p; r -= alpha w; sum r*r (PCG.cpp:144-151)
struct BPcgXR {
    double* x;
    double* r;
    const double* p;
    const double* w;
    const KState* st;
    float* r32 = nullptr;  // also store r in FP32 (the input of an FP32 preconditioner)
    __device__ void operator()(int i, double (&a)[1]) const {
        const double al = st->alpha;
        stS(&x[i], ldS(&x[i]) + al * ldS(&p[i]));
        const double ri = ldS(&r[i]) - al * ldS(&w[i]);
        stS(&r[i], ri);
        if (r32) r32[i] = (float)ri;
        a[0] += ri * ri;
    }
    __device__ void terms(int i, double (&t)[1]) const { t[0] = r[i] * r[i]; }
};

// ====================================================== PBiCGStab bodies
struct BBicgP {  // PBiCGStab.cpp:131-175
    const double* r;
    const double* Ay;
    double* p;
    const KState* st;
    __device__ void operator()(int i) const {
        if (st->nIter == 0)
            p[i] = r[i];
        else
            p[i] = r[i] + st->beta * (p[i] - st->omega * Ay[i]);
    }
};
// flexible BiCGStab: p = r at the start and after a shadow restart
struct BFbicgP {
    const double* r;
    const double* Ay;
    double* p;
    const KState* st;
    __device__ void operator()(int i) const {
        if (st->kShadow == 0)
            p[i] = r[i];
        else
            p[i] = r[i] + st->beta * (p[i] - st->omega * Ay[i]);
    }
};
// shadow restart: r0 = r and r.r
struct BShadowRestart {
    const double* r;
    double* r0;
    __device__ void operator()(int i, double (&a)[1]) const {
        const double ri = r[i];
        r0[i] = ri;
        a[0] += ri * ri;
    }
    __device__ void terms(int i, double (&t)[1]) const { t[0] = r[i] * r[i]; }
};
struct BAmulDotR0 {  // AyA = A yA; rA0.AyA (PBiCGStab.cpp:248-254)
    DMat<double> A;
    const double* y;
    const double* r0;
    double* Ay;
    __device__ void operator()(int i, double (&a)[1]) const {
        const double v = rowAmul<double, double>(A, i, y);
        Ay[i] = v;
        a[0] += r0[i] * v;
    }
    __device__ void terms(int i, double (&t)[1]) const { t[0] = r0[i] * Ay[i]; }
};
struct BBicgS {  // sA = rA - alpha AyA; ||sA|| (PBiCGStab.cpp:276-281)
    const double* r;
    const double* Ay;
    double* s;
    const KState* st;
    __device__ void operator()(int i, double (&a)[1]) const {
        const double si = r[i] - st->alpha * Ay[i];
        s[i] = si;
        a[0] += si * si;
    }
    __device__ void terms(int i, double (&q)[1]) const { q[0] = s[i] * s[i]; }
};
struct BAmulTT {  // tA = A zA; tA.tA and tA.sA (PBiCGStab.cpp:355-364)
    DMat<double> A;
    const double* z;
    const double* s;
    double* t;
    __device__ void operator()(int i, double (&a)[2]) const {
        const double ti = rowAmul<double, double>(A, i, z);
        t[i] = ti;
        a[0] += ti * ti;
        a[1] += ti * s[i];
    }
    __device__ void terms(int i, double (&q)[2]) const {
        q[0] = t[i] * t[i];
        q[1] = t[i] * s[i];
    }
};
struct BBicgXR {  // PBiCGStab.cpp:388-396
    double* x;
    double* r;
    const double* y;
    const double* z;
    const double* s;
    const double* t;
    const KState* st;
    __device__ void operator()(int i, double (&a)[1]) const {
        const double al = st->alpha, om = st->omega;
        x[i] += al * y[i] + om * z[i];
        const double ri = s[i] - om * t[i];
        r[i] = ri;
        a[0] += ri * ri;
    }
    __device__ void terms(int i, double (&q)[1]) const { q[0] = r[i] * r[i]; }
};
// early exit x += alpha yA (PBiCGStab.cpp:298-312), once: every later launch
// (iterations enqueued past the exit) finds it applied.  The state needs no
// iteration index from the host, so the step replays inside a graph.
static __global__ void kBicgEarlyX(double* x, const double* y, const KState* st, int n) {
    if (st->exitIt < 0 || st->exitApplied) return;
    const double al = st->alpha;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        x[i] += al * y[i];
}
static __global__ void kBicgExitApplied(KState* st) {
    if (threadIdx.x == 0 && st->exitIt >= 0) st->exitApplied = 1;
}

// ================================================================ FCG bodies
// rho = r.z and zq = z.q (q = A p of the previous iteration), z in the
// preconditioner's precision
template <typename TZ>
struct BFcgRz {
    const double* r;
    const TZ* z;
    const double* q;
    int first;
    __device__ void operator()(int i, double (&a)[2]) const {
        const double zi = (double)ldS(&z[i]);
        a[0] += ldS(&r[i]) * zi;
        if (!first) a[1] += zi * ldS(&q[i]);
    }
    __device__ void terms(int i, double (&t)[2]) const {
        t[0] = r[i] * (double)z[i];
        t[1] = first ? 0.0 : (double)z[i] * q[i];
    }
};
template <typename TZ>
struct BFcgP {
    const TZ* z;
    double* p;
    const KState* st;
    __device__ void operator()(int i) const {
        if (st->nIter == 0)
            stS(&p[i], (double)ldS(&z[i]));
        else
            stS(&p[i], (double)ldS(&z[i]) + st->beta * ldS(&p[i]));
    }
};

// ============================================================== MG bodies
// Acf = A f; sums source.f and Acf.f (scalingFactor.cpp:19-43, first half)
template <typename T, int Shape = kShapeWide8>
struct BSfAmul {
    DMat<T> A;
    const T* f;
    const T* src;
    T* Acf;
    __device__ void operator()(int i, double (&a)[2]) const {
        const T v = rowAmul<T, T, T, Shape>(A, i, f);
        stL<Shape>(&Acf[i], v);
        const T fi = f[i];
        a[0] += (double)(ldL<Shape>(&src[i]) * fi);
        a[1] += (double)(v * fi);
    }
    __device__ void terms(int i, double (&t)[2]) const {
        t[0] = (double)(src[i] * f[i]);
        t[1] = (double)(Acf[i] * f[i]);
    }
};
struct FinSF {  // scalingFactor.cpp:63-64
    KState* st;
    int slot;
    __device__ void operator()(double (&t)[2]) const { st->sf[slot] = t[0] / (t[1] + kSmallD); }
};
// field += (source - Acf)/D (scalingFactor.cpp:42), then field *= sf (Vcycle.cpp:82-88,206-212)
template <typename T>
struct BJacobiScale {
    T* f;
    const T* src;
    const T* Acf;
    const T* d;
    const KState* st;
    int slot;
    const T* pre;  // pre-smoothed correction to add back (Vcycle.cpp:224-243), or null
    __device__ void operator()(int i) const {
        const T sf = (T)st->sf[slot];
        T fi = f[i] + (src[i] - Acf[i]) / d[i];
        fi = fi * sf;
        if (pre) fi = fi + pre[i];
        f[i] = fi;
    }
};
// psi += fsf * (finestCorrection + (finestResidual - Apsi)/D) (Vcycle.cpp:266-286
// with the Jacobi update of scalingFactor.cpp:42)
template <typename T>
struct BPsiScaled {
    T* psi;
    T* fc;
    const T* fres;
    const T* Apsi;
    const T* d;
    const KState* st;
    int slot;
    bool psiZero = false;  // psi is known to be zero: write fsf * c without reading it
    __device__ void operator()(int i) const {
        const T fsf = (T)st->sf[slot];
        const T c = fc[i] + (fres[i] - Apsi[i]) / d[i];
        fc[i] = c;
        psi[i] = psiZero ? fsf * c : psi[i] + fsf * c;
    }
};
template <typename T>
struct BAddTo {  // a += b (Vcycle.cpp:300-311)
    T* a;
    const T* b;
    __device__ void operator()(int i) const { a[i] = a[i] + b[i]; }
};
template <typename T>
struct BSubAmul {  // src -= A corr (Vcycle.cpp:100-124)
    DMat<T> A;
    const T* x;
    T* src;
    __device__ void operator()(int i) const {
        const T v = rowAmul<T, T>(A, i, x);
        src[i] = src[i] - v;
    }
};
template <typename T>
struct BResidNoNorm {  // finestResidual = b - Apsi after Apsi = A psi (MGSolve.cpp:127-132)
    DMat<T> A;
    const T* x;
    const T* b;
    T* Apsi;
    T* r;
    __device__ void operator()(int i) const {
        const T v = rowAmul<T, T>(A, i, x);
        Apsi[i] = v;
        r[i] = b[i] - v;
    }
};

// restrictField (unapMatrix.hpp:378-407): children summed in ascending fine
// external id; optionally also writes the fine field cast to TC.  Children
// are fetched kRestrictBatch at a time (indices, then values) and added in
// order: the scattered child loads otherwise serialise on their latency
// (256^3 finest restriction 185 -> 121 us).
constexpr int kRestrictBatch = 8;
template <typename TF, typename TC>
struct BRestrict {
    const int* cs;
    const int* ch;
    const TF* ff;
    TC* cf;
    TC* ffCast;
    __device__ void operator()(int c) const {
        TC acc = (TC)0;
        const int k0 = cs[c], k1 = cs[c + 1];
        for (int k = k0; k < k1; k += kRestrictBatch) {
            int i[kRestrictBatch];
            TF v[kRestrictBatch];
#pragma unroll
            for (int j = 0; j < kRestrictBatch; ++j) i[j] = (k + j < k1) ? ch[k + j] : 0;
#pragma unroll
            for (int j = 0; j < kRestrictBatch; ++j) v[j] = (k + j < k1) ? ff[i[j]] : (TF)0;
#pragma unroll
            for (int j = 0; j < kRestrictBatch; ++j)
                if (k + j < k1) {
                    acc = acc + (TC)v[j];
                    if (ffCast) ffCast[i[j]] = (TC)v[j];
                }
        }
        cf[c] = acc;
    }
};
// prolongField (unapMatrix.hpp:410-438)
template <typename T>
struct BProlong {
    const int* map;
    const T* cf;
    T* ff;
    __device__ void operator()(int i) const { ff[i] = cf[map[i]]; }
};

// ============================================================ Chebyshev
// chebySmoother.cpp:73-88 (sweep 0) and :105-144 (later sweeps), one sweep per
// kernel, Amul fused; x is ping-ponged (x_in -> x_out) because neighbours
// still need the old values.  rD = 1/D is recomputed (the smoother builds a
// lduDiagPrecond on every call, chebySmoother.cpp:21).
template <typename T, int Shape = kShapeWide8>
struct BCheby {
    DMat<T> A;
    const T* xin;
    T* xout;
    const T* b;
    T* w;
    const KState* est;  // coefficient table
    int sweep;
    int last = 0;  // the call's last sweep: w is not read again (the next call starts at sweep 0)
    __device__ void operator()(int i) const {
        const T ax = rowAmul<T, T, T, Shape>(A, i, xin);
        const T rd = (T)1 / ldL<Shape>(&A.diag[i]);
        const T bi = ldL<Shape>(&b[i]);
        T wi;
        if (sweep == 0) {
            wi = rd * (bi - ax) * (T)est->cheb[0];
        } else {
            wi = (T)est->cheb[3 * sweep + 1] * ldL<Shape>(&w[i]) + (T)est->cheb[3 * sweep + 2] * rd * (bi - ax);
        }
        if (!last) stL<Shape>(&w[i], wi);
        stL<Shape>(&xout[i], xin[i] + wi);
    }
};

// ============================================================ eigenDiagPCG
// eigenDiagPCG.cpp:87-111 w = rD r, wArA; :124-165 p update; :166-186 w = A p,
// wApA; :190-197 x, r update.  rD = 1/D.
template <typename T>
struct BEigW {
    const T* d;
    const T* r;
    T* w;
    __device__ void operator()(int i, double (&a)[1]) const {
        const T wi = ((T)1 / d[i]) * r[i];
        w[i] = wi;
        a[0] += (double)(wi * r[i]);
    }
    __device__ void terms(int i, double (&t)[1]) const { t[0] = (double)(w[i] * r[i]); }
};
template <typename T>
struct BEigP {
    const T* w;
    T* p;
    const KState* st;
    int it;
    __device__ void operator()(int i) const {
        if (it == 0)
            p[i] = w[i];
        else
            p[i] = w[i] + (T)st->beta * p[i];
    }
};
template <typename T>
struct BEigAp {
    DMat<T> A;
    const T* p;
    T* w;
    __device__ void operator()(int i, double (&a)[1]) const {
        const T v = rowAmul<T, T>(A, i, p);
        w[i] = v;
        a[0] += (double)(v * p[i]);
    }
    __device__ void terms(int i, double (&t)[1]) const { t[0] = (double)(w[i] * p[i]); }
};
template <typename T>
struct BEigXR {
    T* x;
    T* r;
    const T* p;
    const T* w;
    const KState* st;
    __device__ void operator()(int i) const {
        const T al = (T)st->alpha;
        x[i] += al * p[i];
        r[i] -= al * w[i];
    }
};
template <typename T>
struct BResidPlain {  // r = b - A x (eigenDiagPCG.cpp:64-76)
    DMat<T> A;
    const T* x;
    const T* b;
    T* r;
    __device__ void operator()(int i) const { r[i] = b[i] - rowAmul<T, T>(A, i, x); }
};

// ================================================================ sweeps
// A row's first SELL position and its info word (length | lower-side count << 16).
struct RowEntRaw {
    long long base = 0;
    uint32_t info = 0;
};

// Row helpers.  e(k) = position of entry k of row r.
template <typename T>
struct RowEnt {
    long long base;
    uint32_t info;
    __device__ RowEnt(const DMat<T>& A, int r) {
        base = A.sliceStart[r >> kSliceShift] + (r & (kSliceRows - 1));
        info = A.rowInfo[r];
    }
    __device__ explicit RowEnt(const RowEntRaw& e) : base(e.base), info(e.info) {}
    __device__ int len() const { return (int)(info & 0xffffu); }
    __device__ int mid() const { return (int)(info >> 16); }
    __device__ long long at(int k) const { return base + ((long long)k << kSliceShift); }
};

// Gauss-Seidel row (lduGaussSeidelSmoother.cpp:53-78): b minus the lower-side
// terms (new x) then the upper-side terms (old x), then / diag.
// acc -= c_k * x[col_k] over entries k = k0, k0 + step, ... (kEnd excluded),
// loads issued kUnroll at a time, subtractions in entry order.  c_k is the
// stored coefficient times `scale` when Scaled (DIC/DILU: rD * coef), the
// coefficient itself otherwise.
template <bool Scaled, typename T, typename TV, typename TX>
__device__ __forceinline__ TV rowSub(const DMat<T>& A, const RowEnt<T>& e, int k0, int kEnd, int step, TV acc,
                                     TV scale, const TX* x) {
    const int n = (kEnd - k0) * step;  // entries (step = +-1)
    for (int t0 = 0; t0 < n; t0 += kUnroll) {
        int c[kUnroll];
        T v[kUnroll];
        TX xv[kUnroll];
#pragma unroll
        for (int j = 0; j < kUnroll; ++j) {
            // positions past the row's end re-read its batch's first entry
            // (cached) and are not added: no branch between the loads
            const long long p = e.at(k0 + step * (t0 + j < n ? t0 + j : t0));
            c[j] = A.col[p];
            v[j] = A.val[p];
        }
#pragma unroll
        for (int j = 0; j < kUnroll; ++j) xv[j] = x[c[j]];
#pragma unroll
        for (int j = 0; j < kUnroll; ++j)
            if (t0 + j < n) {
                if (Scaled)
                    acc = acc - (scale * (TV)v[j]) * (TV)xv[j];
                else
                    acc = acc - (TV)v[j] * (TV)xv[j];
            }
    }
    return acc;
}

template <typename T>
__device__ __forceinline__ void gsRow(const DMat<T>& A, int r, T* x, const T* b) {
    RowEnt<T> e(A, r);
    T acc = rowSub<false, T, T, T>(A, e, 0, e.len(), 1, b[r], (T)1, x);
    acc = acc / A.diag[r];
    x[r] = acc;
}
// DIC/DILU forward row (lduDICPrecond.cpp:141-149, lduDILUPrecond.cpp:223-235):
// w = rD r, then w -= (rD * coef) * w[lower neighbour] in losort order.
template <typename T, typename TV>
__device__ __forceinline__ void fwdRow(const DMat<T>& A, int r, TV* w, const TV* rr, const TV* rD) {
    RowEnt<T> e(A, r);
    const TV rd = rD[r];
    w[r] = rowSub<true, T, TV, TV>(A, e, 0, e.mid(), 1, rd * rr[r], rd, w);
}
// backward row (lduDICPrecond.cpp:151-154): owner faces in descending order
template <typename T, typename TV>
__device__ __forceinline__ void bwdRow(const DMat<T>& A, int r, TV* w, const TV* rD) {
    RowEnt<T> e(A, r);
    const TV rd = rD[r];
    w[r] = rowSub<true, T, TV, TV>(A, e, e.len() - 1, e.mid() - 1, -1, w[r], rd, w);
}

enum SweepMode { SWEEP_GS = 0, SWEEP_FWD = 1, SWEEP_BWD = 2 };

template <typename T, typename TV>
struct SweepArgs {
    DMat<T> A;  // val: lower-side coefficients for GS / DILU-forward, upper for DIC / backward
    TV* x;      // GS: x;  FWD/BWD: w
    const TV* b;   // GS: b;  FWD: r
    const TV* rD;  // FWD/BWD
    const int* rows;      // backward: row list, null = contiguous
    const int* levStart;  // level boundaries (device)
    const TV* rDpos = nullptr;  // backward: rD by schedule position (rD[rows[q]]), or null
};

// Row at schedule position q.  Backward sweeps read the backward copy (A is
// DevMatrix::bwdView: position q holds row rows[q]'s owner-side entries,
// last first), the arithmetic of bwdRow.
template <typename T, typename TV, int MODE>
__device__ __forceinline__ void sweepRow(const SweepArgs<T, TV>& a, int q) {
    if (MODE == SWEEP_BWD) {
        const int r = a.rows[q];
        RowEnt<T> e(a.A, q);
        const TV rd = a.rDpos ? a.rDpos[q] : a.rD[r];
        a.x[r] = rowSub<true, T, TV, TV>(a.A, e, 0, e.len(), 1, a.x[r], rd, a.x);
        return;
    }
    const int r = a.rows ? a.rows[q] : q;
    if (MODE == SWEEP_GS)
        gsRow<T>(a.A, r, a.x, a.b);
    else
        fwdRow<T, TV>(a.A, r, a.x, a.b, a.rD);
}

// One level: positions [p0, p1) of the schedule.
template <typename T, typename TV, int MODE>
__global__ void __launch_bounds__(kBlock) kSweepLevel(SweepArgs<T, TV> a, int p0, int p1, Skip sk) {
    if (sk.on()) return;
    const int q = p0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= p1) return;
    sweepRow<T, TV, MODE>(a, q);
}

// One Gauss-Seidel level for long rows (coarse AMG levels, ~10-30 entries):
// LPR lanes a row.  Lane i of a row's group loads, gathers and multiplies
// entries i, i + LPR, ... of a chunk of LPR * S together; the group's first
// lane then subtracts the chunk's products in entry order from LDS, so the
// arithmetic is gsRow's while the row's loads are in flight at once rather
// than kUnroll at a time.  maxLen bounds the rows' entry counts (the chunk
// loop must be uniform for its barriers).
template <typename T, int LPR>
__global__ void __launch_bounds__(256) kGsLevelLanes(SweepArgs<T, T> a, int p0, int p1, int maxLen, Skip sk) {
    if (sk.on()) return;
    constexpr int S = 32 / LPR;
    __shared__ T buf[256 * S];
    const int i = (int)threadIdx.x % LPR;
    T* mine = buf + threadIdx.x * S;
    const T* group = buf + (threadIdx.x - i) * S;
    const int q = p0 + (int)(blockIdx.x * blockDim.x + threadIdx.x) / LPR;
    const bool active = q < p1;
    int r = 0, cnt = 0;
    long long base = 0;
    T acc = (T)0;
    if (active) {
        r = a.rows ? a.rows[q] : q;
        base = a.A.sliceStart[r >> kSliceShift] + (r & (kSliceRows - 1));
        cnt = (int)(a.A.rowInfo[r] & 0xffffu);
        acc = a.b[r];
    }
    for (int t0 = 0; t0 < maxLen; t0 += LPR * S) {
        int c[S];
        T v[S], xv[S];
#pragma unroll
        for (int m = 0; m < S; ++m) {
            const int t = t0 + i + LPR * m;
            const long long p = base + ((long long)(t < cnt ? t : 0) << kSliceShift);
            c[m] = t < cnt ? a.A.col[p] : 0;
            v[m] = t < cnt ? a.A.val[p] : (T)0;
        }
#pragma unroll
        for (int m = 0; m < S; ++m) xv[m] = (t0 + i + LPR * m < cnt) ? a.x[c[m]] : (T)0;
#pragma unroll
        for (int m = 0; m < S; ++m) mine[m] = v[m] * xv[m];
        __syncthreads();
        if (i == 0 && active) {
            // entry t0 + LPR * m + j is slot m of lane j of the group
            const int nn = min(cnt - t0, LPR * S);
            for (int t = 0; t < nn; ++t) acc = acc - group[(t % LPR) * S + t / LPR];
        }
        __syncthreads();
    }
    if (active && i == 0) a.x[r] = acc / a.A.diag[r];
}

// A tail row with everything but its x gathers in registers: its row index,
// initial value (GS: b; forward: rD r; backward: w), scale (rD), diagonal
// (GS), and the columns and coefficients of its first kUnroll entries.  The
// tail loads a level's rows before the barrier that ends the level before it,
// so after that barrier the level waits only for its gathers of the values
// the earlier levels wrote.  The arithmetic is sweepRow's, entry for entry.
template <typename T, typename TV>
struct TailRow {
    int r = -1;  // row written; -1: none
    int n = 0;   // entries to subtract
    RowEntRaw e;
    TV acc, scale;
    T diag;
    int c[kUnroll];
    T v[kUnroll];
};

// Row index, initial value, scale and diagonal of the row at schedule
// position q (q < 0: none).
template <typename T, typename TV, int MODE>
__device__ __forceinline__ void tailLoadHead(const SweepArgs<T, TV>& a, int q, TailRow<T, TV>& t) {
    t.r = -1;
    t.n = 0;
    if (q < 0) return;
    const int r = (MODE == SWEEP_BWD || a.rows) ? a.rows[q] : q;
    const int at = MODE == SWEEP_BWD ? q : r;  // the backward copy is by position
    t.e.base = a.A.sliceStart[at >> kSliceShift] + (at & (kSliceRows - 1));
    t.e.info = a.A.rowInfo[at];
    t.r = r;
    if (MODE == SWEEP_GS) {
        t.n = (int)(t.e.info & 0xffffu);
        t.acc = (TV)a.b[r];
        t.scale = (TV)1;
        t.diag = a.A.diag[r];
    } else if (MODE == SWEEP_FWD) {
        t.n = (int)(t.e.info >> 16);
        const TV rd = a.rD[r];
        t.acc = rd * a.b[r];
        t.scale = rd;
    } else {
        t.n = (int)(t.e.info & 0xffffu);
        t.scale = a.rDpos ? a.rDpos[q] : a.rD[r];
        t.acc = a.x[r];
    }
}

// Columns and coefficients of the row's first kUnroll entries (positions
// past its end re-read the first entry and are not used).
template <typename T, typename TV>
__device__ __forceinline__ void tailLoadEntries(const DMat<T>& A, TailRow<T, TV>& t) {
    if (t.n <= 0) return;
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) {
        const long long p = t.e.base + ((long long)(j < t.n ? j : 0) << kSliceShift);
        t.c[j] = A.col[p];
        t.v[j] = A.val[p];
    }
}

template <typename T, typename TV>
__device__ __forceinline__ void tailGather(const TV* x, const TailRow<T, TV>& t, TV (&xv)[kUnroll]) {
    if (t.n <= 0) return;
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) xv[j] = x[t.c[j]];
}

template <typename T, typename TV, int MODE>
__device__ __forceinline__ void tailFinish(const SweepArgs<T, TV>& a, const TailRow<T, TV>& t,
                                           const TV (&xv)[kUnroll]) {
    if (t.r < 0) return;
    constexpr bool Scaled = MODE != SWEEP_GS;
    TV acc = t.acc;
    if (t.n > 0) {
#pragma unroll
        for (int j = 0; j < kUnroll; ++j)
            if (j < t.n) {
                if (Scaled)
                    acc = acc - (t.scale * (TV)t.v[j]) * xv[j];
                else
                    acc = acc - (TV)t.v[j] * xv[j];
            }
        if (t.n > kUnroll) {
            RowEnt<T> e(t.e);
            acc = rowSub<Scaled, T, TV, TV>(a.A, e, kUnroll, t.n, 1, acc, t.scale, a.x);
        }
    }
    if (MODE == SWEEP_GS) acc = acc / (TV)t.diag;
    a.x[t.r] = acc;
}

// Levels [l0, l1) in one workgroup (the narrow tail of a schedule, or a whole
// small matrix), nSweeps times (GS only; sweeps re-read b every time).  Each
// thread loads its first row of the next level before it finishes its row of
// the current one, so the loads that do not depend on the sweep's values
// overlap the gathers that do; rows past the block size (levels wider than
// the block) run as sweepRow.  Level bounds are uniform scalar loads.
template <typename T, typename TV, int MODE>
__global__ void __launch_bounds__(1024) kSweepTail(SweepArgs<T, TV> a, int l0, int l1, int nSweeps,
                                                   Skip sk) {
    if (sk.on()) return;
    const int nl = l1 - l0;
    const int steps = nSweeps * nl;
    if (steps <= 0) return;
    const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
    auto lev = [&](int k) { return a.levStart[l0 + k]; };
    int p0 = lev(0), p1 = lev(1);
    TailRow<T, TV> cur, nxt;
    tailLoadHead<T, TV, MODE>(a, p0 + tid < p1 ? p0 + tid : -1, cur);
    tailLoadEntries<T, TV>(a.A, cur);
    // the loop is entered with no load in flight, as after each step's barrier
    __syncthreads();
    int k = 0;  // level of this step, relative to l0
    for (int s = 0; s < steps; ++s) {
        // gathers of this level first, then the next level's row heads, so
        // the wait for the gathers does not cover the heads; the next
        // level's entries (which need the heads) after this level's stores
        TV xv[kUnroll];
        tailGather<T, TV>(a.x, cur, xv);
        int n0 = 0, n1 = 0;
        if (s + 1 < steps) {
            k = (k + 1 == nl) ? 0 : k + 1;
            n0 = lev(k);
            n1 = lev(k + 1);
            tailLoadHead<T, TV, MODE>(a, n0 + tid < n1 ? n0 + tid : -1, nxt);
        }
        tailFinish<T, TV, MODE>(a, cur, xv);
        for (int q = p0 + tid + nt; q < p1; q += nt) sweepRow<T, TV, MODE>(a, q);
        if (s + 1 < steps) tailLoadEntries<T, TV>(a.A, nxt);
        __syncthreads();
        cur = nxt;
        p0 = n0;
        p1 = n1;
    }
}

// ================================================== dense coarsest (FP64)
// y = Ainv b for the coarsest level of the mixed-precision hierarchy: one
// wavefront per row, coalesced row reads, FP64 accumulation.  A lane reads
// kDenseBatch of its elements (64 apart) before it adds them, in element
// order; positions past the row's end re-read its last element and are not
// added.
constexpr int kDenseBatch = 8;
template <typename TB>
__global__ void __launch_bounds__(256) kDenseSolve(const double* __restrict__ Ainv, const TB* b, TB* y,
                                                   int n, Skip sk) {
    if (sk.on()) return;
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= n) return;
    const double* a = Ainv + (size_t)row * n;
    double acc = 0.0;
    for (int j0 = lane; j0 < n; j0 += 64 * kDenseBatch) {
        double av[kDenseBatch], bv[kDenseBatch];
#pragma unroll
        for (int u = 0; u < kDenseBatch; ++u) {
            const int j = min(j0 + 64 * u, n - 1);
            av[u] = a[j];
            bv[u] = (double)b[j];
        }
#pragma unroll
        for (int u = 0; u < kDenseBatch; ++u)
            if (j0 + 64 * u < n) acc += av[u] * bv[u];
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) y[row] = (TB)acc;
}

}  // namespace unap_hip
