// solve.hpp -- host orchestration of the solve path: matrices, preconditioners,
// smoothers, the agglomeration hierarchy with its V-cycle, and the Krylov/MG
// drivers.  All vectors handled here are device vectors in internal order.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <type_traits>
#include <vector>

#include "coarse.hpp"
#include "device.hpp"

namespace unap_hip {

// ------------------------------------------------------------------ matrix
// Cell orders of the device copies.  Sweeps (GS, DIC, DILU) run level by
// level and want each level's rows contiguous (level-major order); Amul and
// every row-parallel kernel want neighbours close together (breadth-first
// order).  On a 256^3 Poisson matrix: Amul 4.1 TB/s breadth-first against
// 1.9 TB/s level-major; DIC 0.82 TB/s level-major against 0.45 TB/s
// breadth-first.  Each consumer asks for the order its hot kernel wants.
enum CellOrder { ORDER_SWEEP = 0, ORDER_STREAM = 1 };
inline int layoutMode(int order) {
    const int forced = forcedOrderMode();
    return forced >= 0 ? forced : (order == ORDER_SWEEP ? 0 : 1);
}

// A matrix's device layout: built on the device when a HIP context exists
// and the matrix has at least deviceSetupMin() cells, on the host otherwise
// (the two give the same layout).
inline Layout buildLayoutAuto(const HostLdu& A, int orderMode) {
    const int forced = forcedOrderMode();
    if (forced >= 0) orderMode = forced;
    const int mn = deviceSetupMin();
    if (mn >= 0 && A.n >= mn && Ctx::peek()) {
        StageTimer t("layout on the device (all)", A.n);
        Layout L;
        if (buildLayoutDevice(A, orderMode, kSortWindow, L)) return L;
    }
    return buildLayout(A, orderMode);
}

// galerkinListsHost's lists, built on the device into the buffers the
// Galerkin kernels read (setup_dev.hip)
void galerkinListsDevice(const CoarseLevel& C, const std::vector<int>& fineL, DBuf<int>& cst, DBuf<int>& cl,
                         DBuf<int>& ist, DBuf<int>& il, DBuf<int>& fst, DBuf<int>& ff);
// the agglomeration's device hooks: pairing lists and coarse addressing on the
// device for levels of at least deviceSetupMin() cells when a HIP context exists
inline AgglHooks deviceAgglHooks(bool always = false) {
    auto on = [always](const HostLdu& A) {
        const int mn = deviceSetupMin();
        return always || (mn >= 0 && A.n >= mn && Ctx::peek() != nullptr);
    };
    AgglHooks h;
    h.lists = [on](const HostLdu& A, const std::vector<double>& wt, std::vector<int>& rs, std::vector<int>& nbr,
                   std::vector<double>& wk) {
        if (!on(A)) return false;
        pairingListsDevice(A, wt, rs, nbr, wk);
        return true;
    };
    h.addressing = [on](const HostLdu& F, CoarseLevel& C, const std::vector<double>& w, std::vector<double>& cw) {
        if (!on(F)) return false;
        coarseAddressingDevice(F, C, w, cw);
        return true;
    };
    return h;
}

struct AgglHost;
struct LduView {
    std::unique_ptr<Layout> layout;
    std::unique_ptr<DevMatrix<double>> d64;
    std::unique_ptr<DevMatrix<float>> d32;
};
struct LduObj {
    HostLdu host;
    int nFwdLevels = -1, nBwdLevels = -1;  // sweep level counts; -1 until asked for
    // bumped by every coefficient update: hierarchies built on this matrix
    // recompute their Galerkin levels when it moves
    long long coeffGen = 0;
    // bumped by every topology update: objects built on the old addressing
    // refuse to run
    long long topoGen = 0;
    bool coeffsPending = false;  // a topology of another face count awaits its coefficients
    // process-unique id (caches keyed on a matrix never confuse two matrices
    // that lived at one address)
    const unsigned long long id = nextId();
    static unsigned long long nextId() {
        static unsigned long long k = 0;
        return ++k;
    }
    // agglomeration built from the default weights, shared by every
    // preconditioner / solver of this matrix with the same level limits
    std::shared_ptr<AgglHost> agglCache;
    int agglMaxLevels = -1, agglNCoarsest = -1, agglNormalised = -1;
    LduView views[3];  // by layout mode, built on first use
    const Layout& layout(int order) {
        LduView& v = views[layoutMode(order)];
        if (!v.layout) v.layout.reset(new Layout(buildLayoutAuto(host, layoutMode(order))));
        return *v.layout;
    }
    DevMatrix<double>& dev64(int order) {
        LduView& v = views[layoutMode(order)];
        if (!v.d64) {
            const Layout& L = layout(order);
            v.d64.reset(new DevMatrix<double>);
            v.d64->build(host, L, Ctx::get().stream);
        }
        return *v.d64;
    }
    DevMatrix<float>& dev32(int order) {
        LduView& v = views[layoutMode(order)];
        if (!v.d32) {
            const Layout& L = layout(order);
            v.d32.reset(new DevMatrix<float>);
            v.d32->build(host, L, Ctx::get().stream);
        }
        return *v.d32;
    }
    int n() const { return host.n; }
};

// Levels of a schedule small enough to run in one workgroup: the first level
// index from which every level holds at most maxRows rows.  The workgroup
// sits on one CU, whose gather rate bounds a level at about 1 us per 1000
// entries (MI355X, coarse AMG levels), while a level of its own launch costs
// about 5 us over the whole chip: a level goes to the tail when its rows hold
// at most kTailEntries sweep entries (4096 measured best of 2048 / 4096 /
// 8192 on 128^3 AMG-GS and DILU solves).
constexpr int kTailEntries = 4096;
inline int tailRows(double entriesPerRow) {
    const double r = kTailEntries / std::max(1.0, entriesPerRow);
    return (int)std::min(2048.0, std::max(64.0, r));
}
inline int tailStart(const std::vector<int>& start, int maxRows) {
    const int nl = (int)start.size() - 1;
    int k = nl;
    while (k > 0 && start[k] - start[k - 1] <= maxRows) --k;
    return k;
}

// One pass over a level schedule: wide levels one kernel each, the narrow
// tail in one workgroup.
// Gauss-Seidel levels of rows this long on average take 16 lanes a row
// (kGsLevelLanes); shorter rows one thread a row.  On 128^3 AMG levels with
// 12-20 entries a row a level launch drops from ~7.5 to ~5.3 us; at 9 entries
// the two are even.
constexpr double kLanesRowLen = 11.0;

template <typename T, typename TV, int MODE>
void runSchedule(SweepArgs<T, TV> a, const std::vector<int>& levStartH, int maxRows, Skip sk,
                 double entriesPerRow = 0.0, int maxLen = 0) {
    Ctx& c = Ctx::get();
    const int nl = (int)levStartH.size() - 1;
    const int tail = tailStart(levStartH, maxRows);
    for (int k = 0; k < tail; ++k) {
        const int p0 = levStartH[k], p1 = levStartH[k + 1];
        if (p1 <= p0) continue;
        if constexpr (MODE == SWEEP_GS && std::is_same<T, TV>::value) {
            if (entriesPerRow >= kLanesRowLen) {
                constexpr int LPR = 16;
                const long long threads = (long long)(p1 - p0) * LPR;
                kGsLevelLanes<T, LPR><<<(unsigned)((threads + 255) / 256), 256, 0, c.stream>>>(a, p0, p1, maxLen, sk);
                UH_LAUNCHED();
                continue;
            }
        }
        const int g = (p1 - p0 + kBlock - 1) / kBlock;
        kSweepLevel<T, TV, MODE><<<g, kBlock, 0, c.stream>>>(a, p0, p1, sk);
        UH_LAUNCHED();
    }
    if (tail < nl) {
        kSweepTail<T, TV, MODE><<<1, 1024, 0, c.stream>>>(a, tail, nl, 1, sk);
        UH_LAUNCHED();
    }
}

// Gauss-Seidel smoothing, nSweeps forward sweeps (lduGaussSeidelSmoother.cpp:40-79).
template <typename T>
void gsSmooth(DevMatrix<T>& M, T* x, const T* b, int nSweeps, Skip sk) {
    if (!M.sweepable) throw Error(-3, "Gauss-Seidel needs faces sorted by owner with owner < neighbour");
    SweepArgs<T, T> a{M.view(), x, b, nullptr, M.fwdRows.p, M.fwdStart.get()};
    const int nl = (int)M.fwdStartH.size() - 1;
    const int maxRows = tailRows(M.entriesPerRow);  // a GS row reads all its entries
    if (tailStart(M.fwdStartH, maxRows) == 0) {
        kSweepTail<T, T, SWEEP_GS><<<1, 1024, 0, Ctx::get().stream>>>(a, 0, nl, nSweeps, sk);
        UH_LAUNCHED();
        return;
    }
    for (int s = 0; s < nSweeps; ++s)
        runSchedule<T, T, SWEEP_GS>(a, M.fwdStartH, maxRows, sk, M.entriesPerRow, M.maxRowLen);
}

// DIC / DILU application w = M^-1 r: forward then backward schedule.
// rDpos: rD in backward-schedule order (rD[bwdRows[q]]), read contiguously by
// the backward sweep, or null (then rD is gathered by row).
template <typename T, typename TV>
void luSweeps(DevMatrix<T>& M, bool dic, TV* w, const TV* r, const TV* rD, Skip sk, const TV* rDpos) {
    if (!M.sweepable) throw Error(-3, "DIC/DILU needs faces sorted by owner with owner < neighbour");
    const DMat<T> Av = dic ? M.viewU() : M.view();
    SweepArgs<T, TV> f{Av, w, r, rD, M.fwdRows.p, M.fwdStart.get()};
    const int maxRows = tailRows(0.5 * M.entriesPerRow);  // one side of each row
    runSchedule<T, TV, SWEEP_FWD>(f, M.fwdStartH, maxRows, sk);
    SweepArgs<T, TV> bk{M.bwdView(Ctx::get().stream), w, nullptr, rD, M.bwdRows.get(), M.bwdStart.get(), rDpos};
    runSchedule<T, TV, SWEEP_BWD>(bk, M.bwdStartH, maxRows, sk);
}

// Work vectors of one level.
template <typename T>
struct LevelWork {
    DBuf<T> a, b, c, d, e;
    void ensure(int n) {
        if ((int)a.n != n) {
            a.alloc(n);
            b.alloc(n);
            c.alloc(n);
            d.alloc(n);
            e.alloc(n);
        }
    }
};

// Chebyshev smoothing (chebySmoother.cpp:13-149).  est holds the coefficient
// table.  x is ping-ponged through tmp; the result ends in x.
template <typename T>
void chebySweeps(DevMatrix<T>& M, T* x, const T* b, int nSweeps, const KState* est, T* w, T* tmp,
                 Skip sk) {
    if (nSweeps > 64) throw Error(-1, "Chebyshev: at most 64 sweeps per call");
    T* cur = x;
    T* nxt = tmp;
    for (int s = 0; s < nSweeps; ++s) {
        if (M.shape == kShapeBatch4)
            launchRows(BCheby<T, kShapeBatch4>{M.view(), cur, nxt, b, w, est, s, s == nSweeps - 1}, M.n, sk);
        else  // the large levels: rows in XCD slabs, so each L2 holds the x halo of one slab
            launchRowsXcd(BCheby<T>{M.view(), cur, nxt, b, w, est, s, s == nSweeps - 1}, M.n, sk);
        std::swap(cur, nxt);
    }
    if (cur != x) launchRows(BCopy<T, T>{cur, x}, M.n, sk);  // a kernel, not a copy node: graph-capturable
}

// Acf = A f with the sums of scalingFactor, in the matrix's rowAmul shape.
template <typename T, typename Fin>
void sfAmul(DevMatrix<T>& M, const T* f, const T* src, T* Acf, Fin fin, Skip sk) {
    if (M.shape == kShapeBatch4)
        launchReduce<2>(BSfAmul<T, kShapeBatch4>{M.view(), f, src, Acf}, M.n, M.iperm.get(), fin, sk);
    else
        launchReduce<2>(BSfAmul<T>{M.view(), f, src, Acf}, M.n, M.iperm.get(), fin, sk);
}

// eigenDiagPCG (eigenDiagPCG.cpp:41-212): nPCG Jacobi-PCG steps updating x,
// then the Lanczos estimate and the Chebyshev coefficient table.
template <typename T>
void eigenEstimate(DevMatrix<T>& M, T* x, const T* b, int nPCG, KState* est, T* p, T* w, T* r,
                   Skip sk, double boost = 1.1, double ratio = 1.5) {
    if (nPCG < 2 || nPCG > 32) throw Error(-1, "eigen estimate: 2..32 PCG steps");
    Ctx& c = Ctx::get();
    kStateInit<<<1, 1, 0, c.stream>>>(est, 0, 0, 0, 0, nullptr, 0);
    UH_LAUNCHED();
    launchRows(BResidPlain<T>{M.view(), x, b, r}, M.n, sk);
    const Skip sk2{sk.a, &est->stop};
    for (int it = 0; it < nPCG; ++it) {
        launchReduce<1>(BEigW<T>{M.diag.get(), r, w}, M.n, M.iperm.get(), FinEigRho{est, it}, sk2);
        launchRows(BEigP<T>{w, p, est, it}, M.n, sk2);
        launchReduce<1>(BEigAp<T>{M.view(), p, w}, M.n, M.iperm.get(), FinEigAlpha{est, it}, sk2);
        launchRows(BEigXR<T>{x, r, p, w, est}, M.n, sk2);
    }
    kLanczosMaxEig<<<1, 1, 0, c.stream>>>(est, nPCG, boost, ratio);
    UH_LAUNCHED();
}

// ----------------------------------------------------------------- smoother
struct SmootherObj {
    int type;                // 1 GS, 2 Chebyshev
    bool firstCall = true;   // Chebyshev: next call only estimates (chebySmoother.cpp:22-42)
    DBuf<KState> est;
    SmootherObj(int t) : type(t), est(1) {}
};

// ----------------------------------------------------- agglomeration levels
// Device side of one agglomeration level: the segmented-sum lists of the
// Galerkin product (galerkin.hpp) and the coarse coefficients in the coarse
// level's external order.
struct AgglDevLevel {
    DBuf<int> childStart, child;   // per coarse cell: fine cells, ascending
    DBuf<int> innerStart, inner;   // per coarse cell: fine faces inside it, ascending
    DBuf<int> cfStart, cfFine;     // per coarse face: fine faces, ascending (~f: swapped sides)
    DBuf<double> d, up, lo;        // coarse diag, upper, lower (lo empty when symmetric)
};

struct AgglHost {
    std::vector<CoarseLevel> levels;
    // Galerkin coefficients on the device for the fine values of generation
    // valuesGen of the fine matrix (-1: not formed); hostValid[k]: level k's
    // values are also in levels[k].A
    std::vector<AgglDevLevel> dev;
    DBuf<double> fineD, fineUp, fineLo;
    long long valuesGen = -1;
    std::vector<char> hostValid;
    // Galerkin products of every level from the fine coefficients (fine.symm
    // decides the symmetric or asymmetric form, as lduAgglomerateMatrix.cpp:29
    // does per level).
    void galerkin(const HostLdu& fine, long long gen);
    // levels[k].A with its Galerkin values (copied from the device once per
    // generation)
    const HostLdu& hostLevel(int k);
    // device layouts of the coarse levels per layout mode, each built when a
    // consumer first asks for it (a hierarchy that keeps every second
    // agglomeration level never builds the others)
    std::vector<Layout> layouts[3];
    std::vector<char> built[3];
    const Layout& layoutAt(int order, int k) {
        const int m = layoutMode(order);
        if (layouts[m].size() != levels.size()) {
            layouts[m].assign(levels.size(), Layout());
            built[m].assign(levels.size(), 0);
        }
        if (!built[m][k]) {
            layouts[m][k] = buildLayoutAuto(levels[k].A, m);
            built[m][k] = 1;
        }
        return layouts[m][k];
    }
};
inline void buildAggl(const HostLdu& fine, const std::vector<double>& w, int maxLevels, int nCoarsest,
                      AgglHost& out) {
    {
        StageTimer t("agglomeration (every level)", fine.n);
        // pairing lists and coarse addressing on the device where a context
        // exists (the greedy pass itself stays sequential on the host)
        const AgglHooks hooks = deviceAgglHooks();
        out.levels = agglomerate(fine, w, maxLevels, nCoarsest, &hooks);
    }
    for (auto& v : out.layouts) v.clear();
    for (auto& v : out.built) v.clear();
    out.dev.clear();
    out.valuesGen = -1;
    out.hostValid.assign(out.levels.size(), 0);
}

// -------------------------------------------------------------- hierarchy
struct AmgParams {
    int smoother = 1, nPre = 0, nPost = 2, nFinest = 2, nVcycles = 2, precision = 64;
    int growth = 1;  // coarse level l smooths nPost + l (pre: nPre + l) times
    int maxLevels = 50, nCoarsest = 50;
    int pairSteps = 1;  // agglomeration steps per hierarchy level
    bool fixedCycle = false;  // estimates at setup, dense FP64 coarsest solve
    int scaling = 2;          // 2 all but the two coarsest levels, 1 coarse levels only, 0 none
    // Chebyshev interval [lambda / ratioOf(level), lambda * chebyBoost] of every
    // level (chebySmoother.cpp:63-64; the reference's 1.5 and 1.1 by default)
    double chebyRatio = 1.5, chebyBoost = 1.1;
    double chebyRatioFinest = 0;  // > 0: the finest level's ratio in place of chebyRatio
    double ratioOf(int si) const { return (si == 0 && chebyRatioFinest > 0) ? chebyRatioFinest : chebyRatio; }
};

// The V-cycle of MGSolver (Vcycle.cpp:27-332) in precision T on a resident
// hierarchy.  For T = float the coarsest level is solved directly in FP64.
template <typename T>
struct Hierarchy {
    LduObj* fineObj = nullptr;
    DevMatrix<T>* fine = nullptr;
    std::shared_ptr<AgglHost> ag;
    std::vector<std::unique_ptr<DevMatrix<T>>> mats;  // coarse levels (not the coarsest for FP32)
    std::vector<DBuf<int>> cs, ch, f2c;                // transfer into level l from its fine level
    // Restriction into level l through the agglomeration levels it skips
    // (pair steps > 2): sub-steps of two pairwise levels, each a segmented
    // sum over its own child lists, with a vector for every intermediate level
    struct Chain {
        std::vector<DBuf<int>> cs, ch;
        std::vector<DBuf<T>> buf;  // intermediate levels (one fewer than the sub-steps)
        std::vector<int> n;        // cells of each sub-step's coarse side
    };
    std::vector<Chain> chain;
    std::vector<LevelWork<T>> work;                   // corr(a) src(b) t1(c) t2(d) per coarse level
    LevelWork<T> fw;                                  // finest: fCorr(a) Apsi(b) w(c) tmp(d)
    DBuf<T> fres, pre0, src32;                        // finest residual copy, spare, source cast
    DBuf<double> coarseSrc64, coarseCorr64;           // FP32 path: coarsest in FP64
    std::vector<DBuf<T>> pre;                         // pre-smoothed corrections (nPre > 0)
    AmgParams prm;
    int order = ORDER_STREAM;  // GS hierarchies sweep: ORDER_SWEEP
    bool scale = true;        // coarse-level correction scaling
    bool scaleFinest = true;  // and on the finest level
    int L = 0;
    // smoothers: index 0 finest, l+1 coarse level l
    std::vector<char> firstCall;
    DBuf<KState> est;    // one per smoother
    DBuf<KState> mgst;   // scaling factors
    // coarsest (FP64 reference path)
    std::unique_ptr<DevMatrix<double>> coarsest64;
    DBuf<double> rDc;
    size_t coarseLds = 0;
    // LDS-resident copy of a symmetric coarsest level (kCoarsePCGLocal)
    bool coarseLocal = false;
    size_t coarseLocalLds = 0;
    DBuf<int> csrStart;
    DBuf<uint16_t> csrCol, csrMid;
    DBuf<double> csrVal;
    // coarsest (FP32 path): dense inverse in internal order
    DBuf<double> Ainv;

    std::vector<int> sel;  // hierarchy coarse level l = agglomeration level sel[l]
    const CoarseLevel& lev(int l) const { return ag->levels[sel[l]]; }
    int nl(int l) const { return l < 0 ? fine->n : lev(l).nCoarse; }
    DevMatrix<T>& mat(int l) { return l < 0 ? *fine : *mats[l]; }

    // fine coefficient generation the coarse values and coarsest data follow
    long long valuesGen = -1;

    void build(LduObj& A, std::shared_ptr<AgglHost> agg, const AmgParams& p);
    // After a coefficient update of the fine matrix: the Galerkin levels again
    // from the new fine values (MGSolve.cpp:108-111 does it on every solve),
    // with the agglomeration maps kept, and the coarsest-level data and
    // estimates derived from them.  No-op while the values are current.
    void refresh();
    void setupNumeric();
    // The reference re-arms the Chebyshev estimate on every solve; a fixed
    // cycle keeps the estimates taken at setup.
    void initSmoothers() {
        if (!prm.fixedCycle)
            for (auto& f : firstCall) f = 1;
    }
    void smooth(int si, T* x, const T* b, int nSweeps, Skip sk);
    void coarsestSolve(Skip sk);
    void vcycle(T* psi, const T* source, T* fRes, bool psiZero, Skip sk);
    // chain[l] for level l, whose fine level is the layout fl and whose
    // composite map starts at agglomeration level k0
    void buildChain(int l, int k0, const Layout& fl);
    // dst (level l) = restriction of src (its fine level)
    void restrictInto(int l, const T* src, T* dst, Skip sk);
    // w = M^-1 r.  r32: r already in FP32 (FP32 hierarchy), or null.
    void precondition(const double* r, T* w, Skip sk, const T* r32 = nullptr);
};

// ------------------------------------------------------------ preconditioner
struct PrecondObj {
    int type = 0;
    int order = ORDER_STREAM;  // cell order of the vectors it works on
    LduObj* A = nullptr;
    long long topoGen = 0;           // topology generation of A at creation
    DBuf<double> rD;                 // internal order
    std::vector<double> rDext;       // external order (rD())
    std::unique_ptr<Hierarchy<double>> amg64;
    std::unique_ptr<Hierarchy<float>> amg32;
    bool lu32 = false;   // DIC / DILU with FP32 sweeps
    DBuf<float> rD32;    // their rD, internal order
    DBuf<double> rDpos;  // DIC / DILU: rD in backward-schedule order
    DBuf<float> rD32pos;
    int precision() const { return (amg32 || lu32) ? 32 : 64; }
    // w = M^-1 r with FP32 sweeps (lu32), on FP32 vectors
    void applyLu32(const float* r, float* w, Skip sk) {
        if (type == UNAP_HIP_PRECOND_DIC)  // upper on both sides once the storage is asymmetric
            A->dev32(order).buildValU(A->host, A->layout(order), Ctx::get().stream);
        luSweeps<float, float>(A->dev32(order), type == UNAP_HIP_PRECOND_DIC, w, r, rD32.get(), sk, rD32pos.get());
    }
};

// precondition(w, r) in FP64 (types 0-3 and the FP64 hierarchy).
void applyPrecond64(PrecondObj& P, const double* r, double* w, Skip sk);

// ----------------------------------------------------------------- solvers
// An instantiated graph of one solver iteration and what it was captured for.
struct GraphCache {
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned long long> key;
    GraphCache() = default;
    GraphCache(const GraphCache&) = delete;
    GraphCache& operator=(const GraphCache&) = delete;
    ~GraphCache() { reset(); }
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        key.clear();
    }
};

struct SolverObj {
    int type = 1;
    PrecondObj* P = nullptr;
    long long maxIter = 1000, minIter = 0;
    double relTol = 1e-6, tol = 0.0;
    // standalone MG
    AmgParams mg;
    bool mgSet = false;
    std::vector<double> mgWeights;
    bool mgHasWeights = false;
    const double* mgWeightsPtr = nullptr;  // caller-owned, read at the next solve
    std::unique_ptr<Hierarchy<double>> mgH;
    unsigned long long mgFor = 0;  // id of the matrix mgH was built for
    long long mgTopoGen = -1;      // and its topology generation
    // graph of one iteration, replayed by later iterations and solves
    GraphCache graph;
    // results
    std::vector<double> history;
    double lastMs = 0;
};

}  // namespace unap_hip
