// unap_hip.hip -- the solve path on the GPU and its C-ABI (include/unap_hip.h).
#include "unap_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>

#include "dense.hpp"
#include "solve.hpp"

namespace unap_hip {

static thread_local std::string g_lastError;
static int g_graphMode = 1;  // unap_hip_set_graph_replay
void setLastError(const std::string& m) { g_lastError = m; }

static Ctx* g_ctx = nullptr;
Ctx& Ctx::get() {
    if (!g_ctx) {
        Ctx* n = new Ctx;
        UH_CHECK(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
        n->partials.alloc((size_t)kMaxSums * kMaxBlocks);
        n->scratchState.alloc(4);
        g_ctx = n;
    }
    return *g_ctx;
}
Ctx* Ctx::peek() { return g_ctx; }

// Pinned host mirror of one KState for flag read-back.
static KState* hostState() {
    static KState* h = nullptr;
    if (!h) UH_CHECK(hipHostMalloc((void**)&h, sizeof(KState), hipHostMallocDefault));
    return h;
}
static const KState& readState(const KState* d) {
    Ctx& c = Ctx::get();
    KState* h = hostState();
    UH_CHECK(hipMemcpyAsync(h, d, sizeof(KState), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    return *h;
}

// ---------------------------------------------------------------- hierarchy
template <typename T>
struct Fine;
template <>
struct Fine<double> {
    static DevMatrix<double>& get(LduObj& A, int order) { return A.dev64(order); }
};
template <>
struct Fine<float> {
    static DevMatrix<float>& get(LduObj& A, int order) { return A.dev32(order); }
};

// Dense inverse by Gauss-Jordan elimination with partial pivoting (host, FP64).
static std::vector<double> denseInverse(std::vector<double> a, int n) {
    std::vector<double> inv((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        double best = std::fabs(a[(size_t)c * n + c]);
        for (int r = c + 1; r < n; ++r) {
            const double v = std::fabs(a[(size_t)r * n + c]);
            if (v > best) {
                best = v;
                piv = r;
            }
        }
        if (best == 0.0) throw Error(-4, "coarsest level matrix is singular");
        if (piv != c) {
            for (int j = 0; j < n; ++j) {
                std::swap(a[(size_t)c * n + j], a[(size_t)piv * n + j]);
                std::swap(inv[(size_t)c * n + j], inv[(size_t)piv * n + j]);
            }
        }
        const double d = 1.0 / a[(size_t)c * n + c];
        for (int j = 0; j < n; ++j) {
            a[(size_t)c * n + j] *= d;
            inv[(size_t)c * n + j] *= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = a[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int j = 0; j < n; ++j) {
                a[(size_t)r * n + j] -= f * a[(size_t)c * n + j];
                inv[(size_t)r * n + j] -= f * inv[(size_t)c * n + j];
            }
        }
    }
    return inv;
}

// ------------------------------------------------------ Galerkin on the GPU
// The segmented-sum lists of every level are built from the host maps once;
// each call then runs two kernels a level, level k's coarse values being the
// fine values of level k + 1.
void AgglHost::galerkin(const HostLdu& fine, long long gen) {
    Ctx& c = Ctx::get();
    const int K = (int)levels.size();
    if ((int)dev.size() != K) {
        dev.clear();
        dev.resize(K);
        for (int k = 0; k < K; ++k) {
            const CoarseLevel& C = levels[k];
            const std::vector<int>& fl = (k == 0) ? fine.l : levels[k - 1].A.l;
            StageTimer tl("Galerkin: segment lists (device)", (long long)C.restrictMap.size());
            AgglDevLevel& D = dev[k];
            galerkinListsDevice(C, fl, D.childStart, D.child, D.innerStart, D.inner, D.cfStart, D.cfFine);
            c.sync();
        }
    }
    fineD.upload(fine.d, c.stream);
    fineUp.upload(fine.up, c.stream);
    if (!fine.symm)
        fineLo.upload(fine.lo, c.stream);
    else
        fineLo.release();
    for (int k = 0; k < K; ++k) {
        AgglDevLevel& D = dev[k];
        const int nc = levels[k].nCoarse, ncf = levels[k].A.nFaces();
        const double* fd = (k == 0) ? fineD.get() : dev[k - 1].d.get();
        const double* fu = (k == 0) ? fineUp.get() : dev[k - 1].up.get();
        const double* flo = fine.symm ? nullptr : ((k == 0) ? fineLo.get() : dev[k - 1].lo.get());
        if (D.d.n != (size_t)nc) D.d.alloc(nc);
        if (D.up.n != (size_t)std::max(ncf, 1)) D.up.alloc(std::max(ncf, 1));
        if (fine.symm)
            D.lo.release();
        else if (D.lo.n != (size_t)std::max(ncf, 1))
            D.lo.alloc(std::max(ncf, 1));
        kGalerkinDiag<<<(nc + 255) / 256, 256, 0, c.stream>>>(nc, D.childStart.get(), D.child.get(),
                                                              D.innerStart.get(), D.inner.get(), fd, fu, flo,
                                                              D.d.get());
        UH_LAUNCHED();
        if (ncf > 0) {
            kGalerkinFaces<<<(ncf + 255) / 256, 256, 0, c.stream>>>(ncf, D.cfStart.get(), D.cfFine.get(), fu, flo,
                                                                    D.up.get(), fine.symm ? nullptr : D.lo.get());
            UH_LAUNCHED();
        }
        levels[k].A.symm = fine.symm;
    }
    c.sync();
    hostValid.assign(K, 0);
    valuesGen = gen;
}

const HostLdu& AgglHost::hostLevel(int k) {
    if (valuesGen < 0) throw Error(UNAP_HIP_ERR_INTERNAL, "coarse values requested before the Galerkin product");
    if ((int)hostValid.size() != (int)levels.size()) hostValid.assign(levels.size(), 0);
    if (!hostValid[k]) {
        HostLdu& A = levels[k].A;
        const AgglDevLevel& D = dev[k];
        A.d.resize(A.n);
        A.up.resize(A.nFaces());
        A.lo.resize(A.symm ? 0 : A.nFaces());
        UH_CHECK(hipMemcpy(A.d.data(), D.d.get(), sizeof(double) * A.n, hipMemcpyDeviceToHost));
        if (A.nFaces() > 0) {
            UH_CHECK(hipMemcpy(A.up.data(), D.up.get(), sizeof(double) * A.nFaces(), hipMemcpyDeviceToHost));
            if (!A.symm)
                UH_CHECK(hipMemcpy(A.lo.data(), D.lo.get(), sizeof(double) * A.nFaces(), hipMemcpyDeviceToHost));
        }
        hostValid[k] = 1;
    }
    return levels[k].A;
}

template <typename T>
void Hierarchy<T>::refresh() {
    if (valuesGen == fineObj->coeffGen) return;
    Ctx& c = Ctx::get();
    if (ag->valuesGen != fineObj->coeffGen) ag->galerkin(fineObj->host, fineObj->coeffGen);
    for (int l = 0; l < L; ++l)
        if (mats[l]) {
            const AgglDevLevel& D = ag->dev[sel[l]];
            mats[l]->fillOnDevice(D.d.get(), D.up.get(), D.lo.p ? D.lo.get() : nullptr, c.stream);
        }
    c.sync();
    setupNumeric();
    valuesGen = fineObj->coeffGen;
}

template <typename T>
void Hierarchy<T>::build(LduObj& A, std::shared_ptr<AgglHost> agg, const AmgParams& p) {
    Ctx& c = Ctx::get();
    StageTimer total("hierarchy (all, after the agglomeration)", A.host.n);
    fineObj = &A;
    order = (p.smoother == UNAP_HIP_SMOOTHER_GS) ? ORDER_SWEEP : ORDER_STREAM;
    fine = &Fine<T>::get(A, order);
    ag = agg;
    prm = p;
    // hierarchy levels: agglomeration levels pairSteps, 2 pairSteps, ...
    // (counted from the finest) and the coarsest one, so the first coarse
    // level is always pairSteps steps down whatever the agglomeration's depth
    sel.clear();
    {
        const int nl = (int)ag->levels.size();
        for (int k = p.pairSteps - 1; k < nl - 1; k += p.pairSteps) sel.push_back(k);
        if (nl > 0) sel.push_back(nl - 1);
    }
    L = (int)sel.size();
    if (L == 0)
        throw Error(-4, "agglomeration produced no coarse level (matrix smaller than the coarsest-level size)");
    scale = A.host.symm && p.scaling >= 1;  // scaleCorrection_(A.symm()), MG.cpp:14
    scaleFinest = A.host.symm && p.scaling >= 2;
    const bool fp32 = sizeof(T) == 4;
    const bool fixed = prm.fixedCycle;
    if (fp32 && !fixed) throw Error(UNAP_HIP_ERR_ARG, "the FP32 hierarchy needs the fixed cycle");
    if (ag->valuesGen != A.coeffGen) {
        StageTimer t("hierarchy: Galerkin (device, with list setup)", A.host.n);
        ag->galerkin(A.host, A.coeffGen);
    }
    mats.clear();
    cs.clear();
    ch.clear();
    f2c.clear();
    chain.clear();
    work.clear();
    mats.resize(L);
    work.resize(L);
    for (int l = 0; l < L; ++l) {
        const Layout& fl = (l == 0) ? A.layout(order) : ag->layoutAt(order, sel[l - 1]);
        const Layout& cl = ag->layoutAt(order, sel[l]);
        // restriction map from the previous hierarchy level: the composition
        // of the agglomeration maps in between
        const int k0 = (l == 0) ? 0 : sel[l - 1] + 1;
        std::vector<int> rm = ag->levels[k0].restrictMap;
        for (int k = k0 + 1; k <= sel[l]; ++k)
            for (int& v : rm) v = ag->levels[k].restrictMap[v];
        Transfer tr;
        {
            StageTimer tt("hierarchy: transfer maps", (long long)rm.size());
            tr = buildTransfer(fl, cl, rm);
        }
        cs.emplace_back();
        ch.emplace_back();
        f2c.emplace_back();
        cs.back().upload(tr.childStart, c.stream);
        ch.back().upload(tr.child, c.stream);
        f2c.back().upload(tr.fineToCoarse, c.stream);
        c.sync();
        buildChain(l, k0, fl);
        if (!(fixed && l == L - 1)) {
            const AgglDevLevel& D = ag->dev[sel[l]];
            mats[l].reset(new DevMatrix<T>);
            mats[l]->buildOnDevice(cl, D.d.get(), D.up.get(), D.lo.p ? D.lo.get() : nullptr, c.stream);
        }
        work[l].ensure(lev(l).nCoarse);
    }
    fw.ensure(fine->n);
    fres.alloc(fine->n);
    if (fp32) src32.alloc(fine->n);
    pre.clear();
    if (prm.nPre) {
        pre.resize(L);
        for (int l = 0; l < L; ++l) pre[l].alloc(lev(l).nCoarse);
    }
    firstCall.assign(L, 1);
    est.alloc(L);
    mgst.alloc(1);
    {
        StageTimer t("hierarchy: numeric setup (coarsest level, estimates)", A.host.n);
        setupNumeric();
    }
    valuesGen = A.coeffGen;
}

// Coarsest-level data from the current Galerkin values: the reference
// cycle's rD and LDS copy, or the fixed cycle's dense inverse and Chebyshev
// estimates.
template <typename T>
void Hierarchy<T>::setupNumeric() {
    Ctx& c = Ctx::get();
    const bool fixed = prm.fixedCycle;
    const HostLdu& CA = ag->hostLevel(sel[L - 1]);
    const Layout& CL = ag->layoutAt(order, sel[L - 1]);
    if (!fixed) {
        // reference path: PCG+DIC / PBiCGStab+DILU to relTol 1e-2 in one workgroup
        coarsest64.reset();
        DevMatrix<double>* cm = reinterpret_cast<DevMatrix<double>*>(mats[L - 1].get());
        if (!cm->sweepable) throw Error(-3, "coarsest level faces are not owner-sorted");
        std::vector<double> rdExt = CA.symm ? dicReciprocalD(CA) : diluReciprocalD(CA);
        std::vector<double> rdInt(CL.n);
        for (int i = 0; i < CL.n; ++i) rdInt[i] = rdExt[CL.perm[i]];
        rDc.upload(rdInt, c.stream);
        c.sync();
        const int nvec = CA.symm ? 4 : 9;
        coarseLds = sizeof(double) * ((size_t)nvec * CL.n + 64);
        if (coarseLds > 160 * 1024)
            throw Error(-4, "coarsest level too large for the one-workgroup solver (raise n_cells_coarsest limit)");
        if (CA.symm)
            UH_CHECK(hipFuncSetAttribute((const void*)kCoarsePCG,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)coarseLds));
        else
            UH_CHECK(hipFuncSetAttribute((const void*)kCoarseBiCG,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)coarseLds));
        // Symmetric coarsest level: a CSR copy that kCoarsePCGLocal keeps in
        // LDS, when it fits the device's per-workgroup limit.
        coarseLocal = false;
        if (CA.symm && CL.n <= 65535) {
            const int n = CL.n;
            std::vector<int> st(n + 1, 0);
            for (int i = 0; i < n; ++i) st[i + 1] = st[i] + (int)(CL.rowInfo[i] & 0xffffu);
            const int nnz = st[n];
            const size_t bytes = coarseLocalBytes(n, nnz, (int)CL.fwdStart.size() - 1, (int)CL.bwdStart.size() - 1,
                                                  !CL.fwdRows.empty());
            int dev = 0, optin = 0;
            UH_CHECK(hipGetDevice(&dev));
            if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess) optin = 0;
            if (bytes <= (size_t)optin &&
                hipFuncSetAttribute((const void*)kCoarsePCGLocal, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)bytes) == hipSuccess) {
                const Layout CLh = hostArrays(CL, c.stream);
                std::vector<double> sv;
                fillSellValues<double>(CLh, CA, false, sv);
                std::vector<double> val(nnz);
                std::vector<uint16_t> col(nnz), mid(n);
                for (int i = 0; i < n; ++i) {
                    const long long base = CL.sliceStart[i >> kSliceShift] + (i & (kSliceRows - 1));
                    for (int k = 0; k < st[i + 1] - st[i]; ++k) {
                        const long long q = base + ((long long)k << kSliceShift);
                        val[st[i] + k] = sv[q];
                        col[st[i] + k] = (uint16_t)CLh.col[q];
                    }
                    mid[i] = (uint16_t)(CL.rowInfo[i] >> 16);
                }
                csrStart.upload(st, c.stream);
                csrVal.upload(val, c.stream);
                csrCol.upload(col, c.stream);
                csrMid.upload(mid, c.stream);
                c.sync();
                coarseLocalLds = bytes;
                coarseLocal = true;
            } else {
                (void)hipGetLastError();  // a refused attribute leaves no sticky error
            }
        }
    } else {
        // fixed cycle: dense FP64 inverse of the coarsest Galerkin matrix,
        // on the GPU for a symmetric matrix, on the host (partial pivoting)
        // otherwise
        const int nc = CL.n;
        if (nc > 16384 || (!CA.symm && nc > 8192))
            throw Error(-4, "coarsest level too large for the dense FP64 solve");
        std::vector<double> D((size_t)nc * nc, 0.0);
        for (int i = 0; i < nc; ++i) D[(size_t)CL.iperm[i] * nc + CL.iperm[i]] = CA.d[i];
        const double* lw = CA.lower();
        for (int f = 0; f < CA.nFaces(); ++f) {
            const int a = CL.iperm[CA.l[f]], b = CL.iperm[CA.u[f]];
            D[(size_t)a * nc + b] += CA.up[f];
            D[(size_t)b * nc + a] += lw[f];
        }
        if (CA.symm) {
            denseInverseSPD(D, nc, Ainv, c.stream);
        } else {
            std::vector<double> inv = denseInverse(D, nc);
            Ainv.upload(inv, c.stream);
            c.sync();
        }
        // Chebyshev estimates once, here: x = 0 and a fixed pseudo-random b
        if (prm.smoother == UNAP_HIP_SMOOTHER_CHEBY) {
            for (int si = 0; si < L; ++si) {
                DevMatrix<T>& M = (si == 0) ? *fine : *mats[si - 1];
                LevelWork<T>& W = (si == 0) ? fw : work[si - 1];
                std::vector<T> bh(M.n);
                uint64_t s = 0x5eed + (uint64_t)si;
                for (int i = 0; i < M.n; ++i) {
                    s = s * 6364136223846793005ull + 1442695040888963407ull;
                    bh[i] = (T)((double)(s >> 11) * 0x1.0p-53 * 2.0 - 1.0);
                }
                DBuf<T> bb;
                bb.upload(bh, c.stream);
                launchRows(BFill<T>{W.a.get(), (T)0}, M.n);
                eigenEstimate<T>(M, W.a.get(), bb.get(), 10, est.get() + si, W.b.get(), W.c.get(),
                                 W.d.get(), Skip{nullptr, nullptr}, prm.chebyBoost, prm.ratioOf(si));
                c.sync();
            }
            for (auto& f : firstCall) f = 0;
        }
    }
}

template <typename T>
void Hierarchy<T>::smooth(int si, T* x, const T* b, int nSweeps, Skip sk) {
    DevMatrix<T>& M = (si == 0) ? *fine : *mats[si - 1];
    LevelWork<T>& W = (si == 0) ? fw : work[si - 1];
    if (prm.smoother == UNAP_HIP_SMOOTHER_GS) {
        gsSmooth<T>(M, x, b, nSweeps, sk);
        return;
    }
    // work vectors free while smoothing: finest a,b,c,d; coarse c,d,e
    T* v1 = (si == 0) ? W.a.get() : W.c.get();
    T* v2 = (si == 0) ? W.b.get() : W.d.get();
    T* v3 = (si == 0) ? W.c.get() : W.e.get();
    if (firstCall[si] && !prm.fixedCycle) {
        eigenEstimate<T>(M, x, b, 10, est.get() + si, v1, v2, v3, sk, prm.chebyBoost, prm.ratioOf(si));
        firstCall[si] = 0;
        return;
    }
    chebySweeps<T>(M, x, b, nSweeps, est.get() + si, v1, v2, sk);
}

template <typename T>
void Hierarchy<T>::coarsestSolve(Skip sk) {
    Ctx& c = Ctx::get();
    const int l = L - 1;
    T* corr = work[l].a.get();
    const T* src = work[l].b.get();
    const int nc = lev(l).nCoarse;
    if (prm.fixedCycle) {
        const int g = (nc + 3) / 4;
        kDenseSolve<T><<<g, 256, 0, c.stream>>>(Ainv.get(), src, corr, nc, sk);
        UH_LAUNCHED();
        return;
    }
    DevMatrix<double>* cm = reinterpret_cast<DevMatrix<double>*>(mats[l].get());
    const HostLdu& CA = lev(l).A;
    CoarseArgs a;
    a.A = cm->view();
    a.AP = CA.symm ? cm->viewU() : cm->view();
    a.rD = rDc.get();
    a.fwdRows = cm->fwdRows.p;
    a.fwdStart = cm->fwdStart.get();
    a.nFwd = (int)cm->fwdStartH.size() - 1;
    a.bwdRows = cm->bwdRows.get();
    a.bwdStart = cm->bwdStart.get();
    a.nBwd = (int)cm->bwdStartH.size() - 1;
    a.src = reinterpret_cast<const double*>(src);
    a.corr = reinterpret_cast<double*>(corr);
    a.relTol = 1.0e-2;
    a.tol = 0.0;
    a.maxIter = 100;
    a.iperm = cm->iperm.get();
    a.serial = c.serialReductions;
    if (CA.symm && coarseLocal)
        kCoarsePCGLocal<<<1, 256, coarseLocalLds, c.stream>>>(
            a, CoarseCsr{csrStart.get(), csrCol.get(), csrMid.get(), csrVal.get(), (int)csrVal.n}, sk);
    else if (CA.symm)
        kCoarsePCG<<<1, 1024, coarseLds, c.stream>>>(a, sk);
    else
        kCoarseBiCG<<<1, 1024, coarseLds, c.stream>>>(a, sk);
    UH_LAUNCHED();
}

// With pair steps above 2 a coarse cell has ~20 children (4 steps), spread
// over several breadth-first layers of the fine level, and a one-launch
// restriction fetches about ten times its bytes.  Restricting two pairwise
// levels at a time (about 4.6 children each), through the agglomeration
// levels in between in their own breadth-first orders, keeps the gathers in
// cache (256^3, four steps: 240 -> 129 us).  The sums associate differently
// from the one-launch sum (the same terms).
template <typename T>
void Hierarchy<T>::buildChain(int l, int k0, const Layout& fl) {
    chain.emplace_back();
    Chain& C = chain.back();
    const int k1 = sel[l];
    if (k1 - k0 + 1 <= 2) return;  // one step
    Ctx& c = Ctx::get();
    const Layout* from = &fl;
    for (int a = k0; a <= k1;) {
        const int b = std::min(a + 1, k1);  // this sub-step: agglomeration levels a..b
        const Layout& to = ag->layoutAt(order, b);
        std::vector<int> rm = ag->levels[a].restrictMap;
        for (int k = a + 1; k <= b; ++k)
            for (int& v : rm) v = ag->levels[k].restrictMap[v];
        Transfer tr = buildTransfer(*from, to, rm);
        C.cs.emplace_back();
        C.ch.emplace_back();
        C.cs.back().upload(tr.childStart, c.stream);
        C.ch.back().upload(tr.child, c.stream);
        c.sync();
        C.n.push_back(to.n);
        if (b < k1) {
            C.buf.emplace_back();
            C.buf.back().alloc(to.n);
        }
        from = &to;
        a = b + 1;
    }
}

template <typename T>
void Hierarchy<T>::restrictInto(int l, const T* src, T* dst, Skip sk) {
    const Chain& C = chain[l];
    if (C.cs.empty()) {
        launchRowsXcd(BRestrict<T, T>{cs[l].get(), ch[l].get(), src, dst, nullptr}, nl(l), sk);
        return;
    }
    const T* in = src;
    for (size_t j = 0; j < C.cs.size(); ++j) {
        T* out = (j + 1 < C.cs.size()) ? C.buf[j].get() : dst;
        launchRowsXcd(BRestrict<T, T>{C.cs[j].get(), C.ch[j].get(), in, out, nullptr}, C.n[j], sk);
        in = out;
    }
}

template <typename T>
void Hierarchy<T>::vcycle(T* psi, const T* source, T* fRes, bool psiZero, Skip sk) {
    const int coarsest = L - 1;
    // restrict the finest residual (Vcycle.cpp:43-45)
    restrictInto(0, fRes, work[0].b.get(), sk);
    for (int l = 0; l < coarsest; ++l) {
        LevelWork<T>& W = work[l];
        const int n = nl(l);
        if (prm.nPre) {
            launchRows(BFill<T>{W.a.get(), (T)0}, n, sk);
            smooth(l + 1, W.a.get(), W.b.get(), prm.nPre + prm.growth * l, sk);
            if (scale && l < coarsest - 1) {
                sfAmul(mat(l), W.a.get(), W.b.get(), W.c.get(), FinSF{mgst.get(), l}, sk);
                launchRows(BJacobiScale<T>{W.a.get(), W.b.get(), W.c.get(), mat(l).diag.get(), mgst.get(), l,
                                           nullptr},
                           n, sk);
            }
            launchRows(BSubAmul<T>{mat(l).view(), W.a.get(), W.b.get()}, n, sk);
        }
        restrictInto(l + 1, W.b.get(), work[l + 1].b.get(), sk);
    }
    coarsestSolve(sk);
    for (int l = coarsest - 1; l >= 0; --l) {
        LevelWork<T>& W = work[l];
        const int n = nl(l);
        if (prm.nPre) launchRows(BCopy<T, T>{W.a.get(), pre[l].get()}, n, sk);
        launchRowsXcd(BProlong<T>{f2c[l + 1].get(), work[l + 1].a.get(), W.a.get()}, n, sk);
        if (scale && l < coarsest - 1) {
            sfAmul(mat(l), W.a.get(), W.b.get(), W.c.get(), FinSF{mgst.get(), l}, sk);
            launchRows(BJacobiScale<T>{W.a.get(), W.b.get(), W.c.get(), mat(l).diag.get(), mgst.get(), l,
                                       prm.nPre ? pre[l].get() : nullptr},
                       n, sk);
        } else if (prm.nPre) {
            launchRows(BAddTo<T>{W.a.get(), pre[l].get()}, n, sk);
        }
        smooth(l + 1, W.a.get(), W.b.get(), prm.nPost + prm.growth * l, sk);
    }
    const int n = fine->n;
    if (!scaleFinest && psiZero) {
        // unscaled correction onto a zero psi: prolong straight into it
        launchRowsXcd(BProlong<T>{f2c[0].get(), work[0].a.get(), psi}, n, sk);
    } else {
        launchRowsXcd(BProlong<T>{f2c[0].get(), work[0].a.get(), fw.a.get()}, n, sk);
        if (scaleFinest) {
            sfAmul(*fine, fw.a.get(), fRes, fw.b.get(), FinSF{mgst.get(), 63}, sk);
            launchRows(BPsiScaled<T>{psi, fw.a.get(), fRes, fw.b.get(), fine->diag.get(), mgst.get(), 63, psiZero},
                       n, sk);
        } else {
            launchRows(BAddTo<T>{psi, fw.a.get()}, n, sk);
        }
    }
    smooth(0, psi, source, prm.nFinest, sk);
}

// precondition(w, r): w = 0; nVcycles V-cycles of MGSolver::solve(w, A, r)
// with tol = relTol = 0 (the residual after the last cycle is never used).
template <typename T>
void Hierarchy<T>::precondition(const double* r, T* w, Skip sk, const T* r32) {
    const int n = fine->n;
    // The reference starts from w = 0 and adds the first correction; a fixed
    // cycle writes the first correction directly (no fill).
    const bool psiZero = prm.fixedCycle;
    if (!psiZero) launchRows(BFill<T>{w, (T)0}, n, sk);
    const bool fp32 = sizeof(T) == 4;
    T* source = fp32 ? src32.get() : reinterpret_cast<T*>(const_cast<double*>(r));
    if (fp32) {
        // FP32 hierarchy: the residual in FP32, from the caller (stored by the
        // Krylov update) or cast here once, coalesced
        if (r32)
            source = const_cast<T*>(r32);
        else
            launchRows(BCopy<double, T>{r, src32.get()}, n, sk);
    }
    for (int cyc = 0; cyc < prm.nVcycles; ++cyc) {
        if (cyc == 0) {
            vcycle(w, source, source, psiZero, sk);
        } else {
            launchRows(BResidNoNorm<T>{fine->view(), w, source, fw.b.get(), fres.get()}, n, sk);
            vcycle(w, source, fres.get(), false, sk);
        }
    }
}

template struct Hierarchy<double>;
template struct Hierarchy<float>;

struct BDiagMul {
    const double* rD;
    const double* r;
    double* w;
    __device__ void operator()(int i) const { w[i] = rD[i] * r[i]; }
};

void applyPrecond64(PrecondObj& P, const double* r, double* w, Skip sk) {
    const int n = P.A->n();
    switch (P.type) {
        case UNAP_HIP_PRECOND_NONE:
            launchRows(BCopy<double, double>{r, w}, n, sk);
            break;
        case UNAP_HIP_PRECOND_DIAG:
            launchRows(BDiagMul{P.rD.get(), r, w}, n, sk);
            break;
        case UNAP_HIP_PRECOND_DIC:
        case UNAP_HIP_PRECOND_DILU:
            if (P.lu32) throw Error(-4, "FP32 DIC/DILU preconditioners need the FCG or PBiCGStab solver");
            if (P.type == UNAP_HIP_PRECOND_DIC)  // upper on both sides once the storage is asymmetric
                P.A->dev64(P.order).buildValU(P.A->host, P.A->layout(P.order), Ctx::get().stream);
            luSweeps<double, double>(P.A->dev64(P.order), P.type == UNAP_HIP_PRECOND_DIC, w, r, P.rD.get(), sk,
                                     P.rDpos.get());
            break;
        case UNAP_HIP_PRECOND_AMG:
            if (!P.amg64) throw Error(-4, "FP32 AMG preconditioner needs the FCG solver");
            P.amg64->precondition(r, w, sk);
            break;
        default:
            throw Error(-1, "unknown preconditioner");
    }
}

// ------------------------------------------------------------- Krylov drivers
struct SolveBufs {
    int n = -1;
    DBuf<double> x, b, v[8];
    DBuf<float> z32, r32;
    DBuf<double> hist;
    DBuf<KState> st;
    void ensure(int nn, long long histCap, bool f32) {
        if (nn != n) {
            x.alloc(nn);
            b.alloc(nn);
            for (auto& q : v) q.alloc(nn);
            n = nn;
            z32.release();
            r32.release();
        }
        if (f32 && !z32.p) {
            z32.alloc(nn);
            r32.alloc(nn);
        }
        if ((long long)hist.n < histCap) hist.alloc(histCap);
        if (!st.p) st.alloc(1);
    }
};

// Iterations to enqueue before the next flag read-back: the count that reaches
// the tolerance at the observed mean convergence rate, between 1 and 16.
// Iterations enqueued past convergence are skipped on the device.
static int nextBatch(const KState& h, int cur) {
    int b = std::min(cur * 2, 16);
    if (h.nIter >= 1 && h.finalRes > 0 && h.initRes > 0 && h.relTol > 0 && h.finalRes < h.initRes) {
        const double rate = std::log(h.finalRes / h.initRes) / (double)h.nIter;  // < 0
        const double need = std::log(h.relTol * h.initRes / h.finalRes) / rate;
        if (std::isfinite(need)) b = (int)std::ceil(std::max(1.0, need));
    }
    const long long left = h.maxIter - h.nIter;
    if (left > 0 && left < b) b = (int)left;
    return std::max(1, std::min(b, 16));
}

// Iterations replayed from a HIP graph.  Iteration 0 of every solve runs as
// ordinary launches: it does the one-off work whose host side differs (the
// reference cycle's first-call eigenvalue estimates, the backward-sweep
// copies built on first use).  Iteration 1 is captured and the graph replays
// every later one: the kernels read their iteration-dependent scalars from
// the device state, so one graph serves them all.  The graph is kept on the
// solver and reused by later solves while everything its kernel arguments
// point at is unchanged (the key below).
static bool graphsEnabled() {
    static const int on = [] {
        const char* e = std::getenv("UNAP_HIP_GRAPHS");
        return (e && e[0] == '0') ? 0 : 1;
    }();
    return on && g_graphMode;
}

struct Driver {
    SolverObj& S;
    LduObj& A;
    DevMatrix<double>& M;  // the matrix in the preconditioner's cell order
    SolveBufs& B;
    KState* st;
    int n;
    std::vector<unsigned long long> key;  // what the captured launches depend on
    Driver(SolverObj& s, LduObj& a, DevMatrix<double>& m, SolveBufs& b)
        : S(s), A(a), M(m), B(b), st(b.st.get()), n(a.n()) {}

    template <class Enqueue>
    void iterate(Enqueue&& oneIteration) {
        Ctx& c = Ctx::get();
        GraphCache& G = S.graph;
        const bool graphs = graphsEnabled() && !key.empty();
        if (!graphs || G.key != key) G.reset();
        long long it = 0;
        int batch = 1;
        for (;;) {
            for (int k = 0; k < batch; ++k, ++it) {
                if (!graphs || it == 0) {
                    oneIteration(it);
                    continue;
                }
                if (!G.exec) {
                    hipGraph_t g = nullptr;
                    UH_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
                    try {
                        oneIteration(it);
                    } catch (...) {
                        (void)hipStreamEndCapture(c.stream, &g);
                        if (g) (void)hipGraphDestroy(g);
                        throw;
                    }
                    UH_CHECK(hipStreamEndCapture(c.stream, &g));
                    const hipError_t e = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(g);
                    UH_CHECK(e);
                    G.key = key;
                }
                UH_CHECK(hipGraphLaunch(G.exec, c.stream));
            }
            const KState& h = readState(st);
            if (h.done) return;
            batch = nextBatch(h, batch);
        }
    }

    void pcg() {
        double *x = B.x.get(), *b = B.b.get(), *r = B.v[0].get(), *w = B.v[1].get(), *p = B.v[2].get();
        launchReduce<1>(BResidual<double>{M.view(), x, b, r, nullptr}, n, M.iperm.get(), FinInitRes{st, 0});
        const Skip sk{&st->done, nullptr};
        PrecondObj* P = S.P;
        iterate([&](long long) {
            if (!P || P->type == UNAP_HIP_PRECOND_NONE)
                launchReduce<1>(BCopyDot{r, w}, n, M.iperm.get(), FinPcgRho{st}, sk);
            else if (P->type == UNAP_HIP_PRECOND_DIAG)
                launchReduce<1>(BDiagDot{P->rD.get(), r, w}, n, M.iperm.get(), FinPcgRho{st}, sk);
            else {
                applyPrecond64(*P, r, w, sk);
                launchReduce<1>(BDot{w, r}, n, M.iperm.get(), FinPcgRho{st}, sk);
            }
            launchRows(BPcgP{w, p, st}, n, sk);
            launchReduce<1>(BAmulDot{M.view(), p, w}, n, M.iperm.get(), FinPcgAlpha{st}, sk);
            launchReduce<1>(BPcgXR{x, r, p, w, st}, n, M.iperm.get(), FinResNorm{st}, sk);
        });
    }

    void precondInto(const double* in, double* out, Skip sk) {
        if (!S.P) {
            launchRows(BCopy<double, double>{in, out}, n, sk);
        } else if (S.P->lu32) {
            // FP32 sweeps between an FP32 copy of the input and the output
            launchRows(BCopy<double, float>{in, B.r32.get()}, n, sk);
            S.P->applyLu32(B.r32.get(), B.z32.get(), sk);
            launchRows(BCopy<float, double>{B.z32.get(), out}, n, sk);
        } else {
            applyPrecond64(*S.P, in, out, sk);
        }
    }

    // flexible: the shadow-residual restart of FinFbicgRho (no reference counterpart)
    void bicg(bool flexible = false) {
        double *x = B.x.get(), *b = B.b.get();
        double *r = B.v[0].get(), *r0 = B.v[1].get(), *p = B.v[2].get(), *y = B.v[3].get();
        double *Ay = B.v[4].get(), *s = B.v[5].get(), *z = B.v[6].get(), *t = B.v[7].get();
        launchReduce<1>(BResidual<double>{M.view(), x, b, r, r0}, n, M.iperm.get(), FinInitRes{st, 0});
        launchRows(BFill<double>{Ay, 0.0}, n);
        launchRows(BFill<double>{p, 0.0}, n);
        const Skip sk{&st->done, nullptr};
        const Skip skRestart{&st->done, &st->noRestart};
        iterate([&](long long it) {
            if (flexible) {
                launchReduce<1>(BDot{r0, r}, n, M.iperm.get(), FinFbicgRho{st}, sk);
                launchReduce<1>(BShadowRestart{r, r0}, n, M.iperm.get(), FinFbicgRestart{st}, skRestart);
                launchRows(BFbicgP{r, Ay, p, st}, n, sk);
            } else {
                launchReduce<1>(BDot{r0, r}, n, M.iperm.get(), FinBicgRho{st}, sk);
                launchRows(BBicgP{r, Ay, p, st}, n, sk);
            }
            precondInto(p, y, sk);
            launchReduce<1>(BAmulDotR0{M.view(), y, r0, Ay}, n, M.iperm.get(), FinBicgAlpha{st}, sk);
            launchReduce<1>(BBicgS{r, Ay, s, st}, n, M.iperm.get(), FinBicgS{st}, sk);
            {
                const Launch L = rowLaunch(n);
                kBicgEarlyX<<<L.grid, kBlock, 0, Ctx::get().stream>>>(x, y, st, n);
                UH_LAUNCHED();
                kBicgExitApplied<<<1, 64, 0, Ctx::get().stream>>>(st);
                UH_LAUNCHED();
            }
            precondInto(s, z, sk);
            if (flexible)
                launchReduce<2>(BAmulTT{M.view(), z, s, t}, n, M.iperm.get(), FinFbicgOmega{st}, sk);
            else
                launchReduce<2>(BAmulTT{M.view(), z, s, t}, n, M.iperm.get(), FinBicgOmega{st}, sk);
            launchReduce<1>(BBicgXR{x, r, y, z, s, t, st}, n, M.iperm.get(), FinResNorm{st}, sk);
        });
    }

    // Flexible CG (Notay 2000, FCG(1)); q = A p survives into the next
    // iteration's beta.  r32: an FP32 copy of r kept up to date by the
    // residual update (the input of an FP32 preconditioner), or null.
    template <typename TZ>
    void fcgWith(TZ* z, float* r32, std::function<void(const double*, TZ*, Skip)> prec) {
        double *x = B.x.get(), *b = B.b.get(), *r = B.v[0].get(), *q = B.v[1].get(), *p = B.v[2].get();
        launchReduce<1>(BResidual<double>{M.view(), x, b, r, nullptr}, n, M.iperm.get(), FinInitRes{st, 0});
        if (r32) launchRows(BCopy<double, float>{r, r32}, n);
        const Skip sk{&st->done, nullptr};
        iterate([&](long long it) {
            prec(r, z, sk);
            launchReduce<2>(BFcgRz<TZ>{r, z, q, it == 0 ? 1 : 0}, n, M.iperm.get(), FinFcgRz{st}, sk);
            launchRows(BFcgP<TZ>{z, p, st}, n, sk);
            launchReduce<1>(BAmulDot{M.view(), p, q}, n, M.iperm.get(), FinFcgAlpha{st}, sk);
            launchReduce<1>(BPcgXR{x, r, p, q, st, r32}, n, M.iperm.get(), FinResNorm{st}, sk);
        });
    }

    void fcg() {
        if (S.P && S.P->lu32) {
            PrecondObj* P = S.P;
            float* r32 = B.r32.get();
            fcgWith<float>(B.z32.get(), r32,
                           [P, r32](const double*, float* z, Skip sk) { P->applyLu32(r32, z, sk); });
        } else if (S.P && S.P->amg32) {
            Hierarchy<float>* H = S.P->amg32.get();
            float* r32 = B.r32.get();
            fcgWith<float>(B.z32.get(), r32,
                           [H, r32](const double* r, float* z, Skip sk) { H->precondition(r, z, sk, r32); });
        } else {
            fcgWith<double>(B.v[3].get(), nullptr,
                            [this](const double* r, double* z, Skip sk) { precondInto(r, z, sk); });
        }
    }

    void mg(Hierarchy<double>& H) {
        double *x = B.x.get(), *b = B.b.get(), *fres = B.v[0].get();
        H.initSmoothers();
        launchReduce<1>(BResidual<double>{M.view(), x, b, fres, nullptr}, n, M.iperm.get(), FinInitRes{st, 1});
        const Skip sk{&st->done, nullptr};
        iterate([&](long long) {
            H.vcycle(x, b, fres, false, sk);
            launchReduce<1>(BResidual<double>{M.view(), x, b, fres, nullptr}, n, M.iperm.get(), FinResNorm{st}, sk);
        });
    }
};

static std::shared_ptr<AgglHost> makeAggl(const HostLdu& A, const double* weights, bool normalise,
                                          int maxLevels, int nCoarsest) {
    std::vector<double> w(A.nFaces());
    if (weights) {
        for (int f = 0; f < A.nFaces(); ++f) w[f] = weights[f];
    } else if (normalise) {
        // |upper| / ||upper||_2 (fortranInterface.cpp:383-395)
        double s = 0;
        for (int f = 0; f < A.nFaces(); ++f) s += A.up[f] * A.up[f];
        const double avg = std::sqrt(s);
        for (int f = 0; f < A.nFaces(); ++f) w[f] = std::fabs(A.up[f]) / avg;
    } else {
        for (int f = 0; f < A.nFaces(); ++f) w[f] = std::fabs(A.up[f]);  // MGPreconditioner.cpp:14-19
    }
    auto ag = std::make_shared<AgglHost>();
    buildAggl(A, w, maxLevels, nCoarsest, *ag);
    return ag;
}

// Default-weight hierarchies are cached on the matrix (coefficients are part
// of the key only through set_coeffs, which clears the cache).
static std::shared_ptr<AgglHost> agglFor(LduObj& L, const double* weights, bool normalise, int maxLevels,
                                         int nCoarsest) {
    if (weights) return makeAggl(L.host, weights, normalise, maxLevels, nCoarsest);
    if (L.agglCache && L.agglMaxLevels == maxLevels && L.agglNCoarsest == nCoarsest &&
        L.agglNormalised == (int)normalise)
        return L.agglCache;
    L.agglCache = makeAggl(L.host, nullptr, normalise, maxLevels, nCoarsest);
    L.agglMaxLevels = maxLevels;
    L.agglNCoarsest = nCoarsest;
    L.agglNormalised = (int)normalise;
    return L.agglCache;
}

}  // namespace unap_hip

// ====================================================================== C-ABI
using namespace unap_hip;

struct unap_hip_ldu_s {
    LduObj o;
};
struct unap_hip_precond_s {
    PrecondObj o;
};
struct unap_hip_smoother_s {
    SmootherObj o{1};
    LevelWork<double> w;
};
struct unap_hip_aggl_s {
    LduObj* A;
    std::shared_ptr<AgglHost> ag;
    std::vector<DBuf<int>> cs, ch, f2c;
};
struct unap_hip_solver_s {
    SolverObj o;
    SolveBufs bufs;
};

// Scratch of unap_hip_ldu_spmv_noalloc: two vectors of up to
// kNoAllocMaxCells doubles (32 MiB), one copy on every device the library
// runs on, held until the process ends.  Host code never uses the array's
// name as an address: each device's address comes from hipGetSymbolAddress,
// taken once per device on the first call there.  Calls that overlap on one
// device share the array and race.
constexpr long long kNoAllocMaxCells = 1LL << 21;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) double scratch[2 * kNoAllocMaxCells];

namespace {

double* g_scratchPtr[kNoAllocMaxDevices] = {};

double* noAllocScratch() {
    int dev = 0;
    UH_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) throw Error(UNAP_HIP_ERR_ARG, "device index out of range");
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        UH_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratchPtr[dev] = static_cast<double*>(p);
    }
    return g_scratchPtr[dev];
}

template <class F>
int guard(F&& f) {
    try {
        f();
        g_lastError.clear();
        return UNAP_HIP_OK;
    } catch (const Error& e) {
        g_lastError = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_lastError = "host out of memory";
        return UNAP_HIP_ERR_INTERNAL;
    } catch (const std::exception& e) {
        g_lastError = e.what();
        return UNAP_HIP_ERR_INTERNAL;
    }
}

void need(bool c, const char* msg) {
    if (!c) throw Error(UNAP_HIP_ERR_ARG, msg);
}

// External-order vector -> internal-order device vector.
void toInternal(const DevMatrix<double>& M, const double* src, int where, double* dst, DBuf<double>& stage) {
    Ctx& c = Ctx::get();
    const double* s = src;
    if (where == UNAP_HIP_HOST) {
        if ((int)stage.n < M.n) stage.alloc(M.n);
        stage.upload(src, M.n, c.stream);
        s = stage.get();
    }
    launchRows(BGather<double, double>{M.perm.get(), s, dst}, M.n);
}
void toExternal(const DevMatrix<double>& M, const double* src, int where, double* dst, DBuf<double>& stage) {
    Ctx& c = Ctx::get();
    if (where == UNAP_HIP_HOST) {
        if ((int)stage.n < M.n) stage.alloc(M.n);
        launchRows(BScatter<double, double>{M.perm.get(), src, stage.get()}, M.n);
        UH_CHECK(hipMemcpyAsync(dst, stage.get(), sizeof(double) * M.n, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    } else {
        launchRows(BScatter<double, double>{M.perm.get(), src, dst}, M.n);
        c.sync();
    }
}

DBuf<double>& stageBuf() {
    static DBuf<double>* s = new DBuf<double>;
    return *s;
}

void checkLdu(unap_hip_ldu A) { need(A != nullptr, "null matrix handle"); }
// A matrix whose topology changed to another face count has no valid
// coefficients until unap_hip_ldu_set_coeffs.
void checkLduReady(unap_hip_ldu A) {
    checkLdu(A);
    if (A->o.coeffsPending)
        throw Error(UNAP_HIP_ERR_STATE, "the matrix's topology changed its face count: set its coefficients first");
}

}  // namespace

extern "C" {

const char* unap_hip_last_error(void) { return g_lastError.c_str(); }

int unap_hip_set_graph_replay(int on) {
    return guard([&] {
        need(on == 0 || on == 1, "graph replay is 0 or 1");
        g_graphMode = on;
    });
}

int unap_hip_set_reduction_order(int order) {
    return guard([&] {
        need(order == UNAP_HIP_REDUCE_TREE || order == UNAP_HIP_REDUCE_REFERENCE, "unknown reduction order");
        Ctx::get().serialReductions = order;
    });
}

int unap_hip_device_info(char* name, int64_t name_len, int64_t* cus) {
    return guard([&] {
        int dev = 0;
        UH_CHECK(hipGetDevice(&dev));
        hipDeviceProp_t p;
        UH_CHECK(hipGetDeviceProperties(&p, dev));
        if (name && name_len > 0) {
            std::snprintf(name, (size_t)name_len, "%s (%s)", p.name, p.gcnArchName);
        }
        if (cus) *cus = p.multiProcessorCount;
    });
}

int unap_hip_ldu_create(int64_t n, int64_t nf, const int64_t* la, const int64_t* ua, const double* lower,
                        const double* diag, const double* upper, unap_hip_ldu* out) {
    return guard([&] {
        need(out != nullptr, "out is null");
        need(n > 0 && n < (1LL << 31) && nf >= 0 && 2 * nf < (1LL << 31), "sizes out of range");
        need(diag && (nf == 0 || (la && ua && upper)), "null array");
        for (int64_t f = 0; f < nf; ++f)
            need(la[f] >= 0 && la[f] < n && ua[f] >= 0 && ua[f] < n && la[f] != ua[f], "face address out of range");
        auto* h = new unap_hip_ldu_s;
        HostLdu& A = h->o.host;
        A.n = (int)n;
        A.l.assign(la, la + nf);
        A.u.assign(ua, ua + nf);
        A.d.assign(diag, diag + n);
        A.up.assign(upper, upper + nf);
        A.symm = (lower == nullptr || lower == upper);
        if (!A.symm) A.lo.assign(lower, lower + nf);
        *out = h;
    });
}

int unap_hip_ldu_destroy(unap_hip_ldu A) {
    return guard([&] { delete A; });
}

int unap_hip_ldu_info(unap_hip_ldu A, int64_t* n, int64_t* nf, int* symm, int64_t* fl, int64_t* bl) {
    return guard([&] {
        checkLdu(A);
        if (n) *n = A->o.host.n;
        if (nf) *nf = A->o.host.nFaces();
        if (symm) *symm = A->o.host.symm ? 1 : 0;
        if ((fl || bl) && A->o.nFwdLevels < 0) levelCounts(A->o.host, A->o.nFwdLevels, A->o.nBwdLevels);
        if (fl) *fl = A->o.nFwdLevels;
        if (bl) *bl = A->o.nBwdLevels;
    });
}

int unap_hip_ldu_get(unap_hip_ldu A, int64_t* la, int64_t* ua, double* lower, double* diag, double* upper) {
    return guard([&] {
        checkLduReady(A);
        const HostLdu& H = A->o.host;
        const int nf = H.nFaces();
        const double* lw = H.lower();
        for (int f = 0; f < nf; ++f) {
            if (la) la[f] = H.l[f];
            if (ua) ua[f] = H.u[f];
            if (upper) upper[f] = H.up[f];
            if (lower) lower[f] = lw[f];
        }
        if (diag) std::memcpy(diag, H.d.data(), sizeof(double) * H.n);
    });
}

int unap_hip_ldu_set_coeffs(unap_hip_ldu A, const double* diag, const double* upper, const double* lower) {
    return guard([&] {
        checkLdu(A);
        HostLdu& H = A->o.host;
        need(diag && (H.nFaces() == 0 || upper), "null array");
        A->o.coeffsPending = false;
        H.d.assign(diag, diag + H.n);
        H.up.assign(upper, upper + H.nFaces());
        H.symm = (lower == nullptr || lower == upper);
        H.lo.clear();
        if (!H.symm) H.lo.assign(lower, lower + H.nFaces());
        Ctx& c = Ctx::get();
        // New preconditioners agglomerate from the new weights; existing
        // hierarchies keep their maps and redo their Galerkin levels at their
        // next use (Hierarchy::refresh).  DIC/DILU/diagonal preconditioners
        // keep the rD they computed at creation, as the reference's do
        // (lduDICPrecond.cpp:4-13).
        A->o.agglCache.reset();
        A->o.coeffGen += 1;
        for (LduView& v : A->o.views) {
            if (v.d64) v.d64->updateValues(H, *v.layout, c.stream);
            if (v.d32) v.d32->updateValues(H, *v.layout, c.stream);
        }
    });
}

int unap_hip_ldu_set_topology(unap_hip_ldu A, int64_t nf, const int64_t* la, const int64_t* ua) {
    return guard([&] {
        checkLdu(A);
        HostLdu& H = A->o.host;
        const int64_t n = H.n;
        need(nf >= 0 && 2 * nf < (1LL << 31), "sizes out of range");
        need(nf == 0 || (la && ua), "null array");
        for (int64_t f = 0; f < nf; ++f)
            need(la[f] >= 0 && la[f] < n && ua[f] >= 0 && ua[f] < n && la[f] != ua[f], "face address out of range");
        if (Ctx* c = Ctx::peek()) c->sync();  // device copies about to be freed may be in use
        LduObj& o = A->o;
        const bool sameCount = nf == H.nFaces();
        H.l.assign(la, la + nf);
        H.u.assign(ua, ua + nf);
        if (!sameCount) {
            // the coefficient arrays no longer match the faces (the
            // reference keeps them until setMatrixCoeffients): refuse use
            H.up.clear();
            H.lo.clear();
            o.coeffsPending = true;
        }
        for (LduView& v : o.views) {
            v.d64.reset();
            v.d32.reset();
            v.layout.reset();
        }
        o.agglCache.reset();
        o.topoGen += 1;
        o.coeffGen += 1;
        o.nFwdLevels = o.nBwdLevels = -1;
    });
}

int unap_hip_ldu_spmv(unap_hip_ldu A, const double* psi, double* Apsi, int where) {
    return guard([&] {
        checkLduReady(A);
        need(psi && Apsi, "null vector");
        DevMatrix<double>& M = A->o.dev64(ORDER_STREAM);
        DBuf<double> xi(M.n), yi(M.n);
        toInternal(M, psi, where, xi.get(), stageBuf());
        launchRows(BAmul<double>{M.view(), xi.get(), yi.get()}, M.n);
        toExternal(M, yi.get(), where, Apsi, stageBuf());
    });
}

int unap_hip_ldu_spmv_noalloc(unap_hip_ldu A, const double* psi, double* Apsi) {
    return guard([&] {
        checkLduReady(A);
        need(psi && Apsi, "null vector");
        const int n = A->o.n();
        need(n <= kNoAllocMaxCells, "unap_hip_ldu_spmv_noalloc: the matrix has more cells than the scratch holds "
                                    "(at most 2097152)");
        Ctx* c = Ctx::peek();
        const LduView& v = A->o.views[layoutMode(ORDER_STREAM)];
        if (!c || !v.d64)
            throw Error(UNAP_HIP_ERR_STATE, "unap_hip_ldu_spmv_noalloc: call unap_hip_ldu_spmv on this matrix once "
                                            "first (it builds the device layout, which allocates)");
        const DevMatrix<double>& M = *v.d64;
        double* xi = noAllocScratch();  // [0, n): psi in internal order, [n, 2n): A psi
        double* yi = xi + n;
        launchRows(BGather<double, double>{M.perm.get(), psi, xi}, n);
        launchRows(BAmul<double>{M.view(), xi, yi}, n);
        launchRows(BScatter<double, double>{M.perm.get(), yi, Apsi}, n);
        c->sync();
    });
}

int unap_hip_amg_params_default(unap_hip_amg_params* p) {
    return guard([&] {
        need(p != nullptr, "null params");
        p->smoother = UNAP_HIP_SMOOTHER_GS;
        p->n_cells_coarsest = 50;
        p->max_levels = 50;
        p->n_vcycles = 2;
        p->n_pre_sweeps = 0;
        p->n_post_sweeps = 2;
        p->n_finest_sweeps = 2;
        p->precision = 64;
        p->level_sweep_growth = 1;
        p->pair_steps = 1;
        p->cycle = UNAP_HIP_CYCLE_AUTO;
        p->scaling = 2;
        p->cheby_eig_ratio = 1.5;
        p->cheby_boost = 1.1;
        p->cheby_eig_ratio_finest = 0.0;
    });
}

static AmgParams toParams(const unap_hip_amg_params* p) {
    AmgParams a;
    if (!p) return a;
    a.smoother = (int)p->smoother;
    a.nCoarsest = (int)p->n_cells_coarsest;
    a.maxLevels = (int)p->max_levels;
    a.nVcycles = (int)p->n_vcycles;
    a.nPre = (int)p->n_pre_sweeps;
    a.nPost = (int)p->n_post_sweeps;
    a.nFinest = (int)p->n_finest_sweeps;
    a.precision = (int)p->precision;
    a.growth = (int)p->level_sweep_growth;
    a.pairSteps = (int)p->pair_steps;
    a.scaling = (int)p->scaling;
    a.chebyRatio = p->cheby_eig_ratio;
    a.chebyBoost = p->cheby_boost;
    a.chebyRatioFinest = p->cheby_eig_ratio_finest;
    need(a.scaling >= 0 && a.scaling <= 2, "scaling must be 0, 1 or 2");
    need(std::isfinite(a.chebyRatio) && a.chebyRatio > 1.0 && std::isfinite(a.chebyBoost) && a.chebyBoost > 0.0,
         "cheby_eig_ratio must exceed 1 and cheby_boost be positive");
    need(a.chebyRatioFinest == 0.0 || (std::isfinite(a.chebyRatioFinest) && a.chebyRatioFinest > 1.0),
         "cheby_eig_ratio_finest must be 0 or exceed 1");
    need(a.growth == 0 || a.growth == 1, "level_sweep_growth must be 0 or 1");
    need(a.pairSteps >= 1 && a.pairSteps <= 8, "pair_steps must be 1 to 8");
    need(p->cycle >= -1 && p->cycle <= 1, "unknown cycle");
    a.fixedCycle = (p->cycle == UNAP_HIP_CYCLE_AUTO) ? (a.precision == 32) : (p->cycle == UNAP_HIP_CYCLE_FIXED);
    need(a.fixedCycle || a.precision == 64, "the FP32 hierarchy needs the fixed cycle");
    need(a.smoother == 1 || a.smoother == 2, "smoother must be GS (1) or Chebyshev (2)");
    need(a.precision == 64 || a.precision == 32, "precision must be 64 or 32");
    need(a.nVcycles >= 1 && a.nPre >= 0 && a.nPost >= 0 && a.nFinest >= 0, "bad sweep counts");
    need(a.maxLevels >= 2 && a.maxLevels <= 60 && a.nCoarsest >= 1, "bad level limits");
    return a;
}

int unap_hip_precond_create(unap_hip_ldu A, int type, const unap_hip_amg_params* amg, const double* fw,
                            unap_hip_precond* out) {
    return guard([&] {
        checkLduReady(A);
        need(out != nullptr, "out is null");
        need(type >= 0 && type <= 4, "unknown preconditioner type");
        // the parameters are checked before anything touches the device
        const AmgParams prm = (type == UNAP_HIP_PRECOND_AMG) ? toParams(amg) : AmgParams{};
        std::unique_ptr<unap_hip_precond_s> P(new unap_hip_precond_s);
        PrecondObj& o = P->o;
        o.type = type;
        o.A = &A->o;
        o.topoGen = A->o.topoGen;
        const HostLdu& H = A->o.host;
        Ctx& c = Ctx::get();
        if (type == UNAP_HIP_PRECOND_DIC || type == UNAP_HIP_PRECOND_DILU) o.order = ORDER_SWEEP;
        if (type == UNAP_HIP_PRECOND_DIAG || type == UNAP_HIP_PRECOND_DIC || type == UNAP_HIP_PRECOND_DILU) {
            const Layout& L = A->o.layout(o.order);
            if (type == UNAP_HIP_PRECOND_DIAG) {
                o.rDext.resize(H.n);
                for (int i = 0; i < H.n; ++i) o.rDext[i] = 1.0 / H.d[i];
            } else {
                if (!H.ownerSorted())
                    throw Error(UNAP_HIP_ERR_FACE_ORDER, "DIC/DILU need faces sorted by owner with owner < neighbour");
                o.rDext = (type == UNAP_HIP_PRECOND_DIC) ? dicReciprocalD(H) : diluReciprocalD(H);
            }
            std::vector<double> rin(H.n);
            for (int i = 0; i < H.n; ++i) rin[i] = o.rDext[L.perm[i]];
            o.rD.upload(rin, c.stream);
            o.lu32 = (type != UNAP_HIP_PRECOND_DIAG) && amg != nullptr && amg->precision == 32;
            need(amg == nullptr || amg->precision == 32 || amg->precision == 64, "precision must be 32 or 64");
            std::vector<double> rpos;
            if (type != UNAP_HIP_PRECOND_DIAG) {
                rpos.resize(H.n);
                for (int q = 0; q < H.n; ++q) rpos[q] = rin[L.bwdRows[q]];
            }
            std::vector<float> r32, r32pos;  // host sources of the copies: alive until the sync below
            if (o.lu32) {
                r32.assign(rin.begin(), rin.end());
                o.rD32.upload(r32, c.stream);
                r32pos.assign(rpos.begin(), rpos.end());
                o.rD32pos.upload(r32pos, c.stream);
                if (type == UNAP_HIP_PRECOND_DIC) A->o.dev32(o.order).buildValU(H, L, c.stream);
            } else if (type != UNAP_HIP_PRECOND_DIAG) {
                o.rDpos.upload(rpos, c.stream);
                if (type == UNAP_HIP_PRECOND_DIC) A->o.dev64(o.order).buildValU(H, L, c.stream);
            }
            c.sync();
        } else if (type == UNAP_HIP_PRECOND_AMG) {
            o.order = (prm.smoother == UNAP_HIP_SMOOTHER_GS) ? ORDER_SWEEP : ORDER_STREAM;
            auto ag = agglFor(A->o, fw, false, prm.maxLevels, prm.nCoarsest);
            if (prm.precision == 32) {
                o.amg32.reset(new Hierarchy<float>);
                o.amg32->build(A->o, ag, prm);
            } else {
                o.amg64.reset(new Hierarchy<double>);
                o.amg64->build(A->o, ag, prm);
            }
        }
        *out = P.release();
    });
}

int unap_hip_precond_destroy(unap_hip_precond P) {
    return guard([&] { delete P; });
}

int unap_hip_precond_apply(unap_hip_precond P, const double* r, double* w, int where) {
    return guard([&] {
        need(P && r && w, "null argument");
        PrecondObj& o = P->o;
        if (o.A->coeffsPending || o.topoGen != o.A->topoGen)
            throw Error(UNAP_HIP_ERR_STATE, "preconditioner was built for an earlier topology of this matrix");
        DevMatrix<double>& M = o.A->dev64(o.order);
        DBuf<double> ri(M.n), wi(M.n);
        toInternal(M, r, where, ri.get(), stageBuf());
        if (o.amg32) o.amg32->refresh();
        if (o.amg64) o.amg64->refresh();
        if (o.amg32) {
            DBuf<float> w32(M.n);
            o.amg32->initSmoothers();
            o.amg32->precondition(ri.get(), w32.get(), Skip{nullptr, nullptr});
            launchRows(BCopy<float, double>{w32.get(), wi.get()}, M.n);
        } else if (o.lu32) {
            DBuf<float> r32(M.n), w32(M.n);
            launchRows(BCopy<double, float>{ri.get(), r32.get()}, M.n);
            o.applyLu32(r32.get(), w32.get(), Skip{nullptr, nullptr});
            launchRows(BCopy<float, double>{w32.get(), wi.get()}, M.n);
        } else {
            if (o.amg64) o.amg64->initSmoothers();
            applyPrecond64(o, ri.get(), wi.get(), Skip{nullptr, nullptr});
        }
        toExternal(M, wi.get(), where, w, stageBuf());
    });
}

int unap_hip_precond_rd(unap_hip_precond P, double* rD) {
    return guard([&] {
        need(P && rD, "null argument");
        if (P->o.rDext.empty()) throw Error(UNAP_HIP_ERR_STATE, "this preconditioner has no rD");
        std::memcpy(rD, P->o.rDext.data(), sizeof(double) * P->o.rDext.size());
    });
}

int unap_hip_dense_inverse_spd(int64_t n, const double* a, double* inv) {
    return guard([&] {
        need(n > 0 && n <= 16384 && a && inv, "bad dense inverse arguments");
        std::vector<double> h(a, a + (size_t)n * n);
        DBuf<double> d;
        Ctx& c = Ctx::get();
        denseInverseSPD(h, (int)n, d, c.stream);
        UH_CHECK(hipMemcpy(inv, d.get(), sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost));
    });
}

int unap_hip_precond_amg_levels(unap_hip_precond P, int64_t cap, int64_t* nLevels, int64_t* nCells,
                                int64_t* nFaces) {
    return guard([&] {
        need(P != nullptr && nLevels != nullptr, "null argument");
        const PrecondObj& o = P->o;
        const AgglHost* ag = o.amg32 ? o.amg32->ag.get() : (o.amg64 ? o.amg64->ag.get() : nullptr);
        if (!ag) throw Error(UNAP_HIP_ERR_STATE, "not an AMG preconditioner");
        const std::vector<int>& sel = o.amg32 ? o.amg32->sel : o.amg64->sel;
        const int64_t L = 1 + (int64_t)sel.size();
        *nLevels = L;
        for (int64_t l = 0; l < std::min<int64_t>(L, cap); ++l) {
            const HostLdu& H = (l == 0) ? o.A->host : ag->levels[sel[l - 1]].A;
            if (nCells) nCells[l] = H.n;
            if (nFaces) nFaces[l] = H.nFaces();
        }
    });
}

int unap_hip_smoother_create(int type, unap_hip_smoother* out) {
    return guard([&] {
        need(out != nullptr, "out is null");
        need(type == UNAP_HIP_SMOOTHER_GS || type == UNAP_HIP_SMOOTHER_CHEBY, "unknown smoother");
        auto* s = new unap_hip_smoother_s;
        s->o.type = type;
        *out = s;
    });
}

int unap_hip_smoother_destroy(unap_hip_smoother S) {
    return guard([&] { delete S; });
}

int unap_hip_smoother_init(unap_hip_smoother S) {
    return guard([&] {
        need(S != nullptr, "null smoother");
        S->o.firstCall = true;
    });
}

int unap_hip_smoother_smooth(unap_hip_smoother S, unap_hip_ldu A, double* x, const double* b, int64_t nSweeps,
                             int where) {
    return guard([&] {
        need(S && A && x && b, "null argument");
        need(nSweeps >= 0, "negative sweep count");
        checkLduReady(A);
        DevMatrix<double>& M = A->o.dev64(S->o.type == UNAP_HIP_SMOOTHER_GS ? ORDER_SWEEP : ORDER_STREAM);
        DBuf<double> xi(M.n), bi(M.n);
        toInternal(M, x, where, xi.get(), stageBuf());
        toInternal(M, b, where, bi.get(), stageBuf());
        S->w.ensure(M.n);
        const Skip none{nullptr, nullptr};
        if (S->o.type == UNAP_HIP_SMOOTHER_GS) {
            gsSmooth<double>(M, xi.get(), bi.get(), (int)nSweeps, none);
        } else if (S->o.firstCall) {
            eigenEstimate<double>(M, xi.get(), bi.get(), 10, S->o.est.get(), S->w.a.get(), S->w.b.get(),
                                  S->w.c.get(), none);
            S->o.firstCall = false;
        } else {
            chebySweeps<double>(M, xi.get(), bi.get(), (int)nSweeps, S->o.est.get(), S->w.a.get(), S->w.b.get(),
                                none);
        }
        toExternal(M, xi.get(), where, x, stageBuf());
    });
}

int unap_hip_eigen_estimate(unap_hip_ldu A, double* x, const double* b, int64_t nPCGs, int where, double* lam) {
    return guard([&] {
        need(A && x && b && lam, "null argument");
        checkLduReady(A);
        DevMatrix<double>& M = A->o.dev64(ORDER_STREAM);
        DBuf<double> xi(M.n), bi(M.n), p(M.n), w(M.n), r(M.n);
        toInternal(M, x, where, xi.get(), stageBuf());
        toInternal(M, b, where, bi.get(), stageBuf());
        KState* est = Ctx::get().scratchState.get();
        eigenEstimate<double>(M, xi.get(), bi.get(), (int)nPCGs, est, p.get(), w.get(), r.get(),
                              Skip{nullptr, nullptr});
        toExternal(M, xi.get(), where, x, stageBuf());
        *lam = readState(est).lambda;
    });
}

int unap_hip_aggl_create(unap_hip_ldu A, const double* weights, int64_t maxLevels, int64_t nCoarsest,
                         unap_hip_aggl* out) {
    return guard([&] {
        need(A && weights && out, "null argument");
        need(maxLevels >= 1 && maxLevels <= 60 && nCoarsest >= 1, "bad level limits");
        checkLduReady(A);
        auto* g = new unap_hip_aggl_s;
        g->A = &A->o;
        try {
            g->ag = makeAggl(A->o.host, weights, false, (int)maxLevels, (int)nCoarsest);
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}

int unap_hip_aggl_destroy(unap_hip_aggl G) {
    return guard([&] { delete G; });
}

int unap_hip_aggl_nlevels(unap_hip_aggl G, int64_t* n) {
    return guard([&] {
        need(G && n, "null argument");
        *n = (int64_t)G->ag->levels.size();
    });
}

static const CoarseLevel& level(unap_hip_aggl G, int64_t l) {
    need(G != nullptr, "null agglomeration");
    need(l >= 0 && l < (int64_t)G->ag->levels.size(), "level out of range");
    return G->ag->levels[l];
}

int unap_hip_aggl_level_size(unap_hip_aggl G, int64_t l, int64_t* nc, int64_t* nf) {
    return guard([&] {
        const CoarseLevel& C = level(G, l);
        if (nc) *nc = C.nCoarse;
        if (nf) *nf = C.A.nFaces();
    });
}

int unap_hip_aggl_restrict_addressing(unap_hip_aggl G, int64_t l, int64_t* out) {
    return guard([&] {
        const CoarseLevel& C = level(G, l);
        need(out != nullptr, "null output");
        for (size_t i = 0; i < C.restrictMap.size(); ++i) out[i] = C.restrictMap[i];
    });
}

int unap_hip_aggl_face_restrict_addressing(unap_hip_aggl G, int64_t l, int64_t* out) {
    return guard([&] {
        const CoarseLevel& C = level(G, l);
        need(out != nullptr, "null output");
        for (size_t i = 0; i < C.faceRestrictMap.size(); ++i) out[i] = C.faceRestrictMap[i];
    });
}

int unap_hip_aggl_coarse_matrix(unap_hip_aggl G, int64_t l, int64_t* la, int64_t* ua, double* lower, double* diag,
                                double* upper, int* symm) {
    return guard([&] {
        const CoarseLevel& C = level(G, l);
        if (lower || diag || upper) {
            // Galerkin values (on the GPU) of the matrix's current coefficients
            if (G->ag->valuesGen != G->A->coeffGen) G->ag->galerkin(G->A->host, G->A->coeffGen);
            G->ag->hostLevel((int)l);
        }
        const HostLdu& A = C.A;
        const int nf = A.nFaces();
        for (int f = 0; f < nf; ++f) {
            if (la) la[f] = A.l[f];
            if (ua) ua[f] = A.u[f];
            if (upper) upper[f] = A.up[f];
            if (lower) lower[f] = A.lower()[f];
        }
        if (diag)
            for (int i = 0; i < A.n; ++i) diag[i] = A.d[i];
        if (symm) *symm = A.symm ? 1 : 0;
    });
}

// restrict (fine level l -> coarse level l) / prolong (coarse l -> fine) on the
// GPU through the same device transfer structures the V-cycle uses.
static void ensureTransfers(unap_hip_aggl G) {
    if (!G->cs.empty()) return;
    Ctx& c = Ctx::get();
    for (size_t l = 0; l < G->ag->levels.size(); ++l) {
        const Layout& fl = (l == 0) ? G->A->layout(ORDER_STREAM) : G->ag->layoutAt(ORDER_STREAM, (int)l - 1);
        Transfer tr = buildTransfer(fl, G->ag->layoutAt(ORDER_STREAM, (int)l), G->ag->levels[l].restrictMap);
        G->cs.emplace_back();
        G->ch.emplace_back();
        G->f2c.emplace_back();
        G->cs.back().upload(tr.childStart, c.stream);
        G->ch.back().upload(tr.child, c.stream);
        G->f2c.back().upload(tr.fineToCoarse, c.stream);
        c.sync();
    }
}

int unap_hip_aggl_restrict_field(unap_hip_aggl G, int64_t l, const double* ff, double* cf, int where) {
    return guard([&] {
        level(G, l);
        need(ff && cf, "null vector");
        ensureTransfers(G);
        Ctx& c = Ctx::get();
        const Layout& fl = (l == 0) ? G->A->layout(ORDER_STREAM) : G->ag->layoutAt(ORDER_STREAM, (int)l - 1);
        const Layout& cl = G->ag->layoutAt(ORDER_STREAM, (int)l);
        const int nf = fl.n, nc = cl.n;
        DBuf<int> fperm, cperm;
        fperm.upload(fl.perm, c.stream);
        cperm.upload(cl.perm, c.stream);
        DBuf<double> fi(nf), ci(nc), stage(std::max(nf, nc));
        const double* src = ff;
        if (where == UNAP_HIP_HOST) {
            stage.upload(ff, nf, c.stream);
            src = stage.get();
        }
        launchRows(BGather<double, double>{fperm.get(), src, fi.get()}, nf);
        launchRows(BRestrict<double, double>{G->cs[l].get(), G->ch[l].get(), fi.get(), ci.get(), nullptr}, nc);
        if (where == UNAP_HIP_HOST) {
            launchRows(BScatter<double, double>{cperm.get(), ci.get(), stage.get()}, nc);
            UH_CHECK(hipMemcpyAsync(cf, stage.get(), sizeof(double) * nc, hipMemcpyDeviceToHost, c.stream));
        } else {
            launchRows(BScatter<double, double>{cperm.get(), ci.get(), cf}, nc);
        }
        c.sync();
    });
}

int unap_hip_aggl_prolong_field(unap_hip_aggl G, int64_t l, const double* cf, double* ff, int where) {
    return guard([&] {
        level(G, l);
        need(ff && cf, "null vector");
        ensureTransfers(G);
        Ctx& c = Ctx::get();
        const Layout& fl = (l == 0) ? G->A->layout(ORDER_STREAM) : G->ag->layoutAt(ORDER_STREAM, (int)l - 1);
        const Layout& cl = G->ag->layoutAt(ORDER_STREAM, (int)l);
        const int nf = fl.n, nc = cl.n;
        DBuf<int> fperm, cperm;
        fperm.upload(fl.perm, c.stream);
        cperm.upload(cl.perm, c.stream);
        DBuf<double> fi(nf), ci(nc), stage(std::max(nf, nc));
        const double* src = cf;
        if (where == UNAP_HIP_HOST) {
            stage.upload(cf, nc, c.stream);
            src = stage.get();
        }
        launchRows(BGather<double, double>{cperm.get(), src, ci.get()}, nc);
        launchRows(BProlong<double>{G->f2c[l].get(), ci.get(), fi.get()}, nf);
        if (where == UNAP_HIP_HOST) {
            launchRows(BScatter<double, double>{fperm.get(), fi.get(), stage.get()}, nf);
            UH_CHECK(hipMemcpyAsync(ff, stage.get(), sizeof(double) * nf, hipMemcpyDeviceToHost, c.stream));
        } else {
            launchRows(BScatter<double, double>{fperm.get(), fi.get(), ff}, nf);
        }
        c.sync();
    });
}

}  // extern "C"

template <typename T>
static void profileKernel(LduObj& L, int kernel, long long reps, double* avgMs, double* bytes) {
    Ctx& c = Ctx::get();
    const int order = (kernel == UNAP_HIP_KERNEL_GS || kernel == UNAP_HIP_KERNEL_DIC) ? ORDER_SWEEP : ORDER_STREAM;
    DevMatrix<T>& M = Fine<T>::get(L, order);
    DevMatrix<double>& M64 = L.dev64(order);
    const Layout& lay = L.layout(order);
    const int n = M.n;
    const double s = sizeof(T), nn = n, f = L.host.nFaces();
    DBuf<T> x(n), y(n), b(n), w(n);
    DBuf<double> x64(n), r64(n), p64(n), w64(n);
    launchRows(BFill<T>{x.get(), (T)0.5}, n);
    launchRows(BFill<T>{b.get(), (T)1}, n);
    launchRows(BFill<T>{w.get(), (T)0.25}, n);
    for (auto* v : {&x64, &r64, &p64, &w64}) launchRows(BFill<double>{v->get(), 0.5}, n);
    KState* est = c.scratchState.get();
    kStateInit<<<1, 1, 0, c.stream>>>(est, 0, 0, 1, 0, nullptr, 0);
    kLanczosMaxEig<<<1, 1, 0, c.stream>>>(est, 2, 1.1, 1.5);  // some finite table
    DBuf<double> rD;
    std::vector<double> rdh, rin;  // host source of the copy: alive until the timing's synchronisation
    if (kernel == UNAP_HIP_KERNEL_DIC) {
        if (!lay.sweepable) throw Error(-3, "DIC needs owner-sorted faces");
        rdh = dicReciprocalD(L.host);
        rin.resize(n);
        for (int i = 0; i < n; ++i) rin[i] = rdh[lay.perm[i]];
        rD.upload(rin, c.stream);
        M64.buildValU(L.host, lay, c.stream);
    }
    const Skip none{nullptr, nullptr};
    long long launch = 0;
    auto once = [&]() {
        const bool odd = (launch++ & 1) != 0;
        switch (kernel) {
            case UNAP_HIP_KERNEL_AMUL:
                launchRows(BAmul<T>{M.view(), x.get(), y.get()}, n);
                break;
            case UNAP_HIP_KERNEL_CHEBY:  // as the large levels' sweeps run in a solve: x ping-ponged
                launchRowsXcd(BCheby<T>{M.view(), odd ? y.get() : x.get(), odd ? x.get() : y.get(), b.get(), w.get(),
                                        est, 1},
                              n);
                break;
            case UNAP_HIP_KERNEL_GS:
                gsSmooth<T>(M, x.get(), b.get(), 1, none);
                break;
            case UNAP_HIP_KERNEL_DIC:
                luSweeps<double, double>(M64, true, w64.get(), r64.get(), rD.get(), none, nullptr);
                break;
            case UNAP_HIP_KERNEL_RESIDUAL:
                launchReduce<1>(BResidual<T>{M.view(), x.get(), b.get(), y.get(), nullptr}, n, M.iperm.get(),
                                FinInitRes{est, 0});
                break;
            case UNAP_HIP_KERNEL_AXPY_NORM:
                launchReduce<1>(BPcgXR{x64.get(), r64.get(), p64.get(), w64.get(), est}, n, M.iperm.get(),
                                FinResNorm{est});
                break;
            case UNAP_HIP_KERNEL_DOT:
                launchReduce<1>(BDot{x64.get(), p64.get()}, n, M.iperm.get(), FinPcgAlpha{est});
                break;
            case UNAP_HIP_KERNEL_COPY:
                launchRows(BCopy<T, T>{x.get(), y.get()}, n);
                break;
            default:
                throw Error(-1, "unknown kernel id");
        }
    };
    const double s8 = (kernel == UNAP_HIP_KERNEL_DIC || kernel == UNAP_HIP_KERNEL_AXPY_NORM ||
                       kernel == UNAP_HIP_KERNEL_DOT) ? 8.0 : s;
    switch (kernel) {
        case UNAP_HIP_KERNEL_AMUL: *bytes = (3 * s + 4) * nn + (2 * s + 8) * f; break;
        case UNAP_HIP_KERNEL_CHEBY: *bytes = (6 * s + 4) * nn + (2 * s + 8) * f; break;
        case UNAP_HIP_KERNEL_GS: *bytes = (4 * s + 4) * nn + (2 * s + 8) * f; break;
        case UNAP_HIP_KERNEL_DIC: *bytes = (6 * s8 + 12) * nn + (2 * s8 + 8) * f; break;
        case UNAP_HIP_KERNEL_RESIDUAL: *bytes = (4 * s + 4) * nn + (2 * s + 8) * f; break;
        case UNAP_HIP_KERNEL_AXPY_NORM: *bytes = 6 * s8 * nn; break;
        case UNAP_HIP_KERNEL_DOT: *bytes = 2 * s8 * nn; break;
        case UNAP_HIP_KERNEL_COPY: *bytes = 2 * s * nn; break;
    }
    for (int i = 0; i < 3; ++i) once();
    hipEvent_t e0, e1;
    UH_CHECK(hipEventCreate(&e0));
    UH_CHECK(hipEventCreate(&e1));
    UH_CHECK(hipEventRecord(e0, c.stream));
    for (long long i = 0; i < reps; ++i) once();
    UH_CHECK(hipEventRecord(e1, c.stream));
    UH_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    UH_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avgMs = ms / (double)reps;
}

static int64_t layoutMismatches(const Layout& a, const Layout& b) {
    int64_t m = 0;
    m += a.n != b.n;
    m += a.nf != b.nf;
    m += a.nslices != b.nslices;
    m += a.nnzPadded != b.nnzPadded;
    m += a.maxRowLen != b.maxRowLen;
    m += a.sweepable != b.sweepable;
    m += a.perm != b.perm;
    m += a.iperm != b.iperm;
    m += a.sliceStart != b.sliceStart;
    m += a.rowInfo != b.rowInfo;
    m += a.col != b.col;
    m += a.src != b.src;
    m += a.fwdRows != b.fwdRows;
    m += a.fwdStart != b.fwdStart;
    m += a.bwdRows != b.bwdRows;
    m += a.bwdStart != b.bwdStart;
    return m;
}

extern "C" {

int unap_hip_setup_self_check(unap_hip_ldu A, const double* weights, int64_t* mismatches) {
    return guard([&] {
        checkLduReady(A);
        need(mismatches != nullptr, "null argument");
        Ctx::get();
        const HostLdu& H = A->o.host;
        int64_t m = 0;
        for (int mode = 0; mode < 3; ++mode) {
            const Layout host = buildLayout(H, mode);
            Layout dev;
            if (!buildLayoutDevice(H, mode, kSortWindow, dev)) {
                m += 1;
                continue;
            }
            m += !dev.dev || !dev.col.empty() || !dev.src.empty();  // col and src stay in HBM
            m += layoutMismatches(host, hostArrays(dev, Ctx::get().stream));
        }
        std::vector<double> w(H.nFaces());
        for (int f = 0; f < H.nFaces(); ++f) w[f] = weights ? weights[f] : std::fabs(H.up[f]);
        const AgglHooks dev = deviceAgglHooks(true);
        const std::vector<CoarseLevel> a = agglomerate(H, w, 50, 50, nullptr);
        const std::vector<CoarseLevel> b = agglomerate(H, w, 50, 50, &dev);
        m += a.size() != b.size();
        for (size_t k = 0; k < std::min(a.size(), b.size()); ++k) {
            m += a[k].nCoarse != b[k].nCoarse;
            m += a[k].restrictMap != b[k].restrictMap;
            m += a[k].faceRestrictMap != b[k].faceRestrictMap;
            m += a[k].A.l != b[k].A.l;
            m += a[k].A.u != b[k].A.u;
            // Galerkin segment lists
            const std::vector<int>& fl = k == 0 ? H.l : a[k - 1].A.l;
            std::vector<int> hl[6];
            galerkinListsHost(a[k], fl, hl[0], hl[1], hl[2], hl[3], hl[4], hl[5]);
            DBuf<int> dl[6];
            galerkinListsDevice(a[k], fl, dl[0], dl[1], dl[2], dl[3], dl[4], dl[5]);
            for (int q = 0; q < 6; ++q) {
                std::vector<int> got(hl[q].size());
                if (!got.empty() && dl[q].n >= got.size())
                    UH_CHECK(hipMemcpy(got.data(), dl[q].get(), sizeof(int) * got.size(), hipMemcpyDeviceToHost));
                m += dl[q].n < got.size() || got != hl[q];
            }
        }
        *mismatches = m;
    });
}

int unap_hip_profile_kernel(unap_hip_ldu A, int kernel, int precision, int64_t reps, double* avgMs, double* bytes) {
    return guard([&] {
        need(A && avgMs && bytes, "null argument");
        checkLduReady(A);
        need(reps >= 1, "reps must be positive");
        need(precision == 64 || precision == 32, "precision must be 64 or 32");
        if (precision == 64)
            profileKernel<double>(A->o, kernel, reps, avgMs, bytes);
        else
            profileKernel<float>(A->o, kernel, reps, avgMs, bytes);
    });
}

int unap_hip_solver_create(int type, unap_hip_solver* out) {
    return guard([&] {
        need(out != nullptr, "out is null");
        need(type >= 1 && type <= 5, "unknown solver type");
        auto* s = new unap_hip_solver_s;
        s->o.type = type;
        *out = s;
    });
}

int unap_hip_solver_destroy(unap_hip_solver S) {
    return guard([&] { delete S; });
}

int unap_hip_solver_set_precond(unap_hip_solver S, unap_hip_precond P) {
    return guard([&] {
        need(S != nullptr, "null solver");
        S->o.P = P ? &P->o : nullptr;
    });
}

int unap_hip_solver_set_limits(unap_hip_solver S, int64_t maxIter, int64_t minIter, double relTol, double tol) {
    return guard([&] {
        need(S != nullptr, "null solver");
        need(maxIter >= 0 && minIter >= 0, "negative iteration limit");
        S->o.maxIter = maxIter;
        S->o.minIter = minIter;
        S->o.relTol = relTol;
        S->o.tol = tol;
    });
}

int unap_hip_solver_set_mg(unap_hip_solver S, const unap_hip_amg_params* amg, const double* weights) {
    return guard([&] {
        need(S != nullptr, "null solver");
        S->o.mg = toParams(amg);
        need(S->o.mg.precision == 64, "the standalone MG solver runs in FP64");
        S->o.mgSet = true;
        S->o.mgH.reset();
        S->o.mgFor = 0;
        S->o.mgHasWeights = weights != nullptr;
        S->o.mgWeights.clear();
        S->o.mgWeightsPtr = weights;  // copied at the next solve, when the face count is known
    });
}

// Residual history entries kept on the device per solve: the initial residual
// and one per iteration, up to this many (a bound on the buffer whatever
// max_iter is).
static constexpr long long kMaxHistory = 1LL << 20;

int unap_hip_solver_solve(unap_hip_solver S, unap_hip_ldu A, double* x, const double* b, int where,
                          unap_hip_perf* perf) {
    return guard([&] {
        need(S && A && x && b, "null argument");
        checkLduReady(A);
        SolverObj& o = S->o;
        LduObj& L = A->o;
        need(!o.P || o.P->A == &L, "preconditioner was built for another matrix");
        if (o.P && o.P->topoGen != L.topoGen)
            throw Error(UNAP_HIP_ERR_STATE, "preconditioner was built for an earlier topology of this matrix");
        need(!(o.P && o.P->amg32 && o.type != UNAP_HIP_SOLVER_FCG), "the FP32 AMG preconditioner needs FCG");
        need(!(o.P && o.P->lu32 && o.type != UNAP_HIP_SOLVER_FCG && o.type != UNAP_HIP_SOLVER_PBICGSTAB &&
               o.type != UNAP_HIP_SOLVER_FBICGSTAB),
             "FP32 DIC/DILU preconditioners need FCG, PBiCGStab or FBiCGStab");
        need(o.type != UNAP_HIP_SOLVER_MG || o.mgSet, "MG solver needs unap_hip_solver_set_mg first");
        Ctx& c = Ctx::get();
        if (where == UNAP_HIP_DEVICE) UH_CHECK(hipDeviceSynchronize());
        int order = ORDER_STREAM;
        if (o.P) order = o.P->order;
        else if (o.type == UNAP_HIP_SOLVER_MG && o.mg.smoother == UNAP_HIP_SMOOTHER_GS) order = ORDER_SWEEP;
        DevMatrix<double>& M = L.dev64(order);
        const long long histCap = std::min(o.maxIter, kMaxHistory) + 2;
        S->bufs.ensure(M.n, histCap, o.P && (o.P->amg32 || o.P->lu32));
        if (o.type == UNAP_HIP_SOLVER_MG && o.mgHasWeights && o.mgWeightsPtr) {
            // the caller's weights are read once, here, and kept
            o.mgWeights.assign(o.mgWeightsPtr, o.mgWeightsPtr + L.host.nFaces());
            o.mgWeightsPtr = nullptr;
        }
        if (o.type == UNAP_HIP_SOLVER_MG && o.mgHasWeights)
            need((int64_t)o.mgWeights.size() == L.host.nFaces(), "MG weights were given for a matrix with another face count");
        if (o.type == UNAP_HIP_SOLVER_MG && (!o.mgH || o.mgFor != L.id || o.mgTopoGen != L.topoGen)) {
            auto ag = agglFor(L, o.mgHasWeights ? o.mgWeights.data() : nullptr, true, o.mg.maxLevels, o.mg.nCoarsest);
            o.mgH.reset(new Hierarchy<double>);
            o.mgH->build(L, ag, o.mg);
            o.mgFor = L.id;
            o.mgTopoGen = L.topoGen;
        }
        // coefficients updated since the hierarchy last ran: new Galerkin levels
        if (o.type == UNAP_HIP_SOLVER_MG) o.mgH->refresh();
        if (o.P && o.P->amg64) o.P->amg64->refresh();
        if (o.P && o.P->amg32) o.P->amg32->refresh();
        if (o.P && o.P->amg64) o.P->amg64->initSmoothers();
        hipEvent_t e0, e1;
        UH_CHECK(hipEventCreate(&e0));
        UH_CHECK(hipEventCreate(&e1));
        toInternal(M, x, where, S->bufs.x.get(), stageBuf());
        toInternal(M, b, where, S->bufs.b.get(), stageBuf());
        UH_CHECK(hipEventRecord(e0, c.stream));
        KState* st = S->bufs.st.get();
        kStateInit<<<1, 1, 0, c.stream>>>(st, o.tol, o.relTol, o.maxIter, o.minIter, S->bufs.hist.get(), histCap);
        UH_LAUNCHED();
        Driver d(o, L, M, S->bufs);
        d.key = {L.id, (unsigned long long)L.coeffGen, (unsigned long long)L.topoGen,
                 (unsigned long long)(uintptr_t)o.P, (unsigned long long)o.type,
                 (unsigned long long)c.serialReductions, (unsigned long long)(uintptr_t)S->bufs.x.get(),
                 (unsigned long long)(uintptr_t)S->bufs.hist.get(), (unsigned long long)(uintptr_t)o.mgH.get()};
        switch (o.type) {
            case UNAP_HIP_SOLVER_PCG: d.pcg(); break;
            case UNAP_HIP_SOLVER_PBICGSTAB: d.bicg(); break;
            case UNAP_HIP_SOLVER_FBICGSTAB: d.bicg(true); break;
            case UNAP_HIP_SOLVER_FCG: d.fcg(); break;
            case UNAP_HIP_SOLVER_MG: d.mg(*o.mgH); break;
        }
        UH_CHECK(hipEventRecord(e1, c.stream));
        toExternal(M, S->bufs.x.get(), where, x, stageBuf());
        float ms = 0;
        UH_CHECK(hipEventElapsedTime(&ms, e0, e1));
        o.lastMs = ms;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const KState& h = readState(st);
        const long long nh = std::min<long long>(h.nIter + 1, histCap);
        o.history.resize(nh);
        UH_CHECK(hipMemcpy(o.history.data(), S->bufs.hist.get(), sizeof(double) * nh, hipMemcpyDeviceToHost));
        if (perf) {
            perf->initial_residual = h.initRes;
            perf->final_residual = h.finalRes;
            perf->previous_residual = h.prevRes;
            perf->n_iterations = h.nIter;
            perf->converged = h.converged;
            perf->singular = h.singular;
        }
    });
}

int unap_hip_solver_history(unap_hip_solver S, double* out, int64_t cap, int64_t* len) {
    return guard([&] {
        need(S != nullptr, "null solver");
        const auto& h = S->o.history;
        if (len) *len = (int64_t)h.size();
        if (out) std::memcpy(out, h.data(), sizeof(double) * std::min<size_t>(h.size(), (size_t)std::max<int64_t>(cap, 0)));
    });
}

int unap_hip_solver_last_ms(unap_hip_solver S, double* ms) {
    return guard([&] {
        need(S && ms, "null argument");
        *ms = S->o.lastMs;
    });
}

}  // extern "C"
