// unap_oracle.cpp -- CPU restatement of UNAP's solve path (test infrastructure
// only; see unap_oracle.hpp).  Each function cites the reference file:line whose
// serial x86 arithmetic it restates, operation for operation.
#include "unap_oracle.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

namespace unap_oracle {

static const double SMALL_ = 1.0e-37;   // src/base/unap.hpp:88
static const double VSMALL_ = 1.0e-128; // src/base/unap.hpp:89
static const double GREAT_ = 1.0e+37;   // src/base/unap.hpp:92

// src/matrix/unapMatrixTest.cpp:3-31
bool checkConvergence(Perf& p, double tol, double relTol, label iter, label minIter) {
    p.converged = ((p.final_ < tol || (relTol > SMALL_ && p.final_ <= relTol * p.initial)) &&
                   iter >= minIter);
    return p.converged;
}

// src/matrix/unapMatrixTest.cpp:82-97
bool checkSingularity(Perf& p, double residual) {
    p.singular = !(residual > SMALL_);
    return p.singular;
}

// src/matrix/lduMatrix/lduMatrixSpMV.cpp:16-98 (x86 branch :83-91)
void Ldu::spmv(double* y, const double* x) const {
    const double* lw = lower();
    const double* uw = upper();
    for (label c = 0; c < n; ++c) y[c] = 0.0;
    for (label c = 0; c < n; ++c) y[c] += d[c] * x[c];
    const label nf = nFaces();
    for (label f = 0; f < nf; ++f) {
        y[u[f]] += lw[f] * x[l[f]];
        y[l[f]] += uw[f] * x[u[f]];
    }
}

// src/matrix/lduMatrix/lduMatrix.cpp:269-296 (initial fill with nFaces, :273)
const std::vector<label>& Ldu::ownerStart() const {
    if (ownerStart_.empty()) {
        ownerStart_.assign(n + 1, nFaces());
        ownerStart_[0] = 0;
        label nOwnStart = 0, i = 1;
        for (label f = 0; f < nFaces(); ++f) {
            const label cur = l[f];
            if (cur > nOwnStart) {
                while (i <= cur) ownerStart_[i++] = f;
                nOwnStart = cur;
            }
        }
    }
    return ownerStart_;
}

// src/matrix/lduMatrix/lduMatrix.cpp:310-367: faces grouped by neighbour,
// ascending, face order kept inside a group.
const std::vector<label>& Ldu::losort() const {
    if (losort_.empty() && nFaces() > 0) {
        std::vector<label> cnt(n + 1, 0);
        for (label f = 0; f < nFaces(); ++f) cnt[u[f] + 1]++;
        for (label c = 0; c < n; ++c) cnt[c + 1] += cnt[c];
        losort_.assign(nFaces(), -1);
        for (label f = 0; f < nFaces(); ++f) losort_[cnt[u[f]]++] = f;
    }
    return losort_;
}

// src/matrix/lduMatrix/lduMatrix.cpp:381-414 (initial fill 0, :383)
const std::vector<label>& Ldu::losortStart() const {
    if (losortStart_.empty()) {
        losortStart_.assign(n + 1, 0);
        const std::vector<label>& ls = losort();
        label nLs = 0, i = 0;
        for (label k = 0; k < nFaces(); ++k) {
            const label cur = u[ls[k]];
            if (cur > nLs) {
                while (i <= cur) losortStart_[i++] = k;
                nLs = cur;
            }
        }
        losortStart_[n] = nFaces();
    }
    return losortStart_;
}

// src/base/vectors/unapVectorOperations.cpp:6-34
double dot(const double* a, const double* b, label n) {
    double res = 0.0;
    for (label i = 0; i < n; ++i) res += a[i] * b[i];
    return res;
}

// src/base/vectors/unapVector.hpp:478-500
double sumSqr(const double* a, label n) {
    double s = 0;
    for (label i = 0; i < n; ++i) s += a[i] * a[i];
    return s;
}

// src/matrix/unapMatrixSolver.cpp:12-16
double normFactor(const double* a, label n) { return std::sqrt(sumSqr(a, n)) + SMALL_; }

// src/matrix/unapMatrixPrecond.cpp:4-17
void Precond::apply(double* w, const double* r, label n) const {
    for (label i = 0; i < n; ++i) w[i] = r[i];
}

// src/preconditioners/lduPreconditioners/lduDiagonal/lduDiagPrecond.cpp:13-73
DiagPrecond::DiagPrecond(const Ldu& A) : rd(A.n) {
    for (label i = 0; i < A.n; ++i) rd[i] = 1.0 / A.d[i];
}
void DiagPrecond::apply(double* w, const double* r, label n) const {
    for (label i = 0; i < n; ++i) w[i] = rd[i] * r[i];
}

// src/preconditioners/lduPreconditioners/lduDIC/lduDICPrecond.cpp:4-41
DICPrecond::DICPrecond(const Ldu& M) : A(&M), rd(M.d) {
    const double* up = M.upper();
    for (label f = 0; f < M.nFaces(); ++f) rd[M.u[f]] -= up[f] * up[f] / rd[M.l[f]];
    for (label i = 0; i < M.n; ++i) rd[i] = 1.0 / rd[i];
}

// src/preconditioners/lduPreconditioners/lduDIC/lduDICPrecond.cpp:45-77
void DICPrecond::apply(double* w, const double* r, label n) const {
    const double* up = A->upper();
    const label* uu = A->u.data();
    const label* ll = A->l.data();
    const label nf = A->nFaces();
    for (label i = 0; i < n; ++i) w[i] = rd[i] * r[i];
    for (label f = 0; f < nf; ++f) w[uu[f]] -= rd[uu[f]] * up[f] * w[ll[f]];
    for (label f = nf - 1; f >= 0; --f) w[ll[f]] -= rd[ll[f]] * up[f] * w[uu[f]];
}

// src/preconditioners/lduPreconditioners/lduDILU/lduDILUPrecond.cpp:4-41
DILUPrecond::DILUPrecond(const Ldu& M) : A(&M), rd(M.d) {
    const double* up = M.upper();
    const double* lw = M.lower();
    for (label f = 0; f < M.nFaces(); ++f) rd[M.u[f]] -= up[f] * lw[f] / rd[M.l[f]];
    for (label i = 0; i < M.n; ++i) rd[i] = 1.0 / rd[i];
}

// src/preconditioners/lduPreconditioners/lduDILU/lduDILUPrecond.cpp:45-86
void DILUPrecond::apply(double* w, const double* r, label n) const {
    const double* up = A->upper();
    const double* lw = A->lower();
    const label* uu = A->u.data();
    const label* ll = A->l.data();
    const std::vector<label>& ls = A->losort();
    const label nf = A->nFaces();
    for (label i = 0; i < n; ++i) w[i] = rd[i] * r[i];
    for (label k = 0; k < nf; ++k) {
        const label s = ls[k];
        w[uu[s]] -= rd[uu[s]] * lw[s] * w[ll[s]];
    }
    for (label f = nf - 1; f >= 0; --f) w[ll[f]] -= rd[ll[f]] * up[f] * w[uu[f]];
}

// src/smoothers/lduSmoothers/lduGaussSeidelSmoother/lduGaussSeidelSmoother.cpp:3-80
// (serial run: the interface terms leave bPrime = b - 0)
void GSSmoother::smooth(double* x, const Ldu& A, const double* b, label nSweeps) const {
    const label n = A.n;
    std::vector<double> bPrime(n, 0.0);
    const double* up = A.upper();
    const double* lw = A.lower();
    const label* uu = A.u.data();
    const std::vector<label>& os = A.ownerStart();
    for (label sweep = 0; sweep < nSweeps; ++sweep) {
        for (label i = 0; i < n; ++i) bPrime[i] = 0.0;
        for (label i = 0; i < n; ++i) bPrime[i] = b[i] - bPrime[i];
        label fEnd = os[0];
        for (label c = 0; c < n; ++c) {
            const label fStart = fEnd;
            fEnd = os[c + 1];
            double cur = bPrime[c];
            for (label f = fStart; f < fEnd; ++f) cur -= up[f] * x[uu[f]];
            cur /= A.d[c];
            for (label f = fStart; f < fEnd; ++f) bPrime[uu[f]] -= lw[f] * cur;
            x[c] = cur;
        }
    }
}

// w = 0, then one GSSmoother sweep of A w = r.
void GSPrecond::apply(double* w, const double* r, label n) const {
    for (label i = 0; i < n; ++i) w[i] = 0.0;
    GSSmoother().smooth(w, *A, r, 1);
}

// src/matrix/eigenvalue/eigenDiagPCG/eigenDiagPCG.cpp:48-212 (PCG loops) and
// :213-308 (tridiagonal, Gershgorin range, Sturm bisection)
double eigenDiagPCG(const Ldu& A, double* x, const double* b, const double* rD, label nPCGs,
                    std::vector<double>* alphasOut, std::vector<double>* betasOut) {
    const label n = A.n;
    std::vector<double> alphas(nPCGs, 0.0), betas(nPCGs, 0.0);
    std::vector<double> p(n, 0.0), w(n, 0.0), r(n, 0.0);
    double wArA = GREAT_, wArAold = wArA;
    A.spmv(w.data(), x);
    for (label i = 0; i < n; ++i) r[i] = b[i] - w[i];
    for (label it = 0; it < nPCGs; ++it) {
        wArAold = wArA;
        wArA = 0.0;
        for (label i = 0; i < n; ++i) {
            w[i] = rD[i] * r[i];
            wArA += w[i] * r[i];
        }
        if (std::fabs(wArA) < VSMALL_) break;
        if (it == 0) {
            betas[it] = 0.0;
            for (label i = 0; i < n; ++i) p[i] = w[i];
        } else {
            const double beta = wArA / wArAold;
            betas[it] = beta;
            for (label i = 0; i < n; ++i) p[i] = w[i] + beta * p[i];
        }
        A.spmv(w.data(), p.data());
        double wApA = 0.0;
        for (label i = 0; i < n; ++i) wApA += w[i] * p[i];
        const double alpha = wArA / wApA;
        alphas[it] = alpha;
        for (label i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * w[i];
        }
    }
    if (alphasOut) *alphasOut = alphas;
    if (betasOut) *betasOut = betas;
    return maxEigFromLanczos(alphas, betas);
}

double maxEigFromLanczos(const std::vector<double>& alphas, const std::vector<double>& betas) {
    const label m = (label)alphas.size();
    std::vector<std::vector<double>> T(m, std::vector<double>(m, 0.0));
    for (label i = 0; i < m; ++i) {
        if (i == 0)
            T[i][i] = 1.0 / alphas[i];
        else
            T[i][i] = 1.0 / alphas[i] + betas[i] / alphas[i - 1];
        if (i < m - 1) T[i][i + 1] = std::sqrt(betas[i + 1]) / alphas[i];
        if (i > 0) T[i][i - 1] = T[i - 1][i];
    }
    // Gershgorin range, starting from [0, 0]
    double lo = 0.0, hi = 0.0;
    for (label i = 0; i < m; ++i) {
        const double tr = (i == m - 1) ? 0 : T[i][i + 1];
        const double tl = (i == 0) ? 0 : T[i][i - 1];
        const double lmax = T[i][i] + std::fabs(tr) + std::fabs(tl);
        const double lmin = T[i][i] - std::fabs(tr) - std::fabs(tl);
        lo = (lo >= lmin) ? lmin : lo;
        hi = (hi <= lmax) ? lmax : hi;
    }
    // Sturm-sequence bisection for the largest eigenvalue (k = 1)
    double xn;
    std::vector<double> pv(m);
    do {
        xn = (lo + hi) * 0.5;
        label s = 0;
        pv[0] = xn - T[0][0];
        if (pv[0] < 0) s = 1;
        pv[1] = (xn - T[1][1]) * pv[0] - T[0][1] * T[0][1];
        if (pv[1] * pv[0] < 0) s++;
        for (label k = 2; k < m; ++k) {
            pv[k] = (xn - T[k][k]) * pv[k - 1] - T[k - 1][k] * T[k - 1][k] * pv[k - 2];
            if (pv[k] * pv[k - 1] < 0) s++;
        }
        if (s >= 1)
            lo = xn;
        else
            hi = xn;
    } while ((hi - lo) >= 0.00001);
    return xn;
}

// src/smoothers/chebySmoother/chebySmoother.cpp:13-149 with the defaults of
// src/smoothers/chebySmoother/chebySmoother.hpp:41-51
void ChebySmoother::smooth(double* x, const Ldu& A, const double* b, label nSweeps) const {
    DiagPrecond pre(A);
    const double* rD = pre.rd.data();
    if (firstCall) {
        maxEig = eigenDiagPCG(A, x, b, rD, nDiagPCGs);
        firstCall = false;
        return;
    }
    const label n = A.n;
    std::vector<double> p(n, 0.0), w(n, 0.0);
    const double mx = maxEig * boost;
    const double mn = maxEig / eigRatio;
    const double delta = 2.0 / (mx - mn);
    const double theta = 0.5 * (mx + mn);
    const double s1 = theta * delta;
    double rhok = 1.0 / s1;
    A.spmv(w.data(), x);
    const double Rtheta = 1.0 / theta;
    for (label c = 0; c < n; ++c) {
        w[c] = rD[c] * (b[c] - w[c]) * Rtheta;
        x[c] += w[c];
    }
    for (label sweep = 1; sweep < nSweeps; ++sweep) {
        const double rhokp1 = 1.0 / (2.0 * s1 - rhok);
        const double dtemp1 = rhokp1 * rhok;
        const double dtemp2 = 2.0 * rhokp1 * delta;
        rhok = rhokp1;
        A.spmv(p.data(), x);
        for (label c = 0; c < n; ++c) {
            w[c] = dtemp1 * w[c] + dtemp2 * rD[c] * (b[c] - p[c]);
            x[c] += w[c];
        }
    }
}

// src/matrix/lduMatrix/lduAgglomeration/lduAgglomerate.cpp:94-271 (forward_
// is always true, lduAgglomeration.cpp:3)
std::vector<label> Agglomeration::pairLevel(label& nCoarse, const Ldu& A, const double* w) const {
    const label nF = A.n;
    const label nf = A.nFaces();
    const label* ua = A.u.data();
    const label* la = A.l.data();
    // per-cell face lists: faces where the cell is the neighbour, then the
    // faces where it is the owner, each ascending (:111-155)
    std::vector<label> off(nF + 1, 0), nb(nF, 0), cellFaces(2 * nf);
    for (label f = 0; f < nf; ++f) nb[ua[f]]++;
    for (label f = 0; f < nf; ++f) nb[la[f]]++;
    for (label c = 0; c < nF; ++c) off[c + 1] = off[c] + nb[c];
    for (label c = 0; c < nF; ++c) nb[c] = 0;
    for (label f = 0; f < nf; ++f) cellFaces[off[ua[f]] + nb[ua[f]]++] = f;
    for (label f = 0; f < nf; ++f) cellFaces[off[la[f]] + nb[la[f]]++] = f;

    std::vector<label> map(nF, -1);
    nCoarse = 0;
    for (label c = 0; c < nF; ++c) {
        if (map[c] >= 0) continue;
        label match = -1;
        double maxW = SMALL_;
        for (label k = off[c]; k < off[c + 1]; ++k) {
            const label f = cellFaces[k];
            if (map[ua[f]] < 0 && map[la[f]] < 0 && w[f] > maxW) {
                match = f;
                maxW = w[f];
            }
        }
        if (match >= 0) {
            map[ua[match]] = nCoarse;
            map[la[match]] = nCoarse;
            nCoarse++;
        } else {
            label cm = -1;
            double cmW = SMALL_;
            for (label k = off[c]; k < off[c + 1]; ++k) {
                const label f = cellFaces[k];
                if (w[f] > cmW) {
                    cm = f;
                    cmW = w[f];
                }
            }
            if (cm >= 0) {
                const label a = map[ua[cm]], b2 = map[la[cm]];
                map[c] = a > b2 ? a : b2;
            }
        }
    }
    for (label c = 0; c < nF; ++c)
        if (map[c] < 0) map[c] = nCoarse++;
    return map;
}

// src/matrix/lduMatrix/lduAgglomeration/lduAgglomerationAddressing.cpp:8-178
void Agglomeration::buildCoarseAddressing(label lev) {
    const Ldu& A = matrixLevel(lev);
    const std::vector<label>& rm = restrictAddr[lev];
    const label nc = nCells[lev];
    const label nf = A.nFaces();
    std::vector<label> fr(nf);
    std::vector<std::vector<label>> ccFaces(nc);  // provisional coarse faces per coarse owner
    std::vector<label> initNei;
    initNei.reserve(nf);
    for (label f = 0; f < nf; ++f) {
        const label ru = rm[A.u[f]], rl = rm[A.l[f]];
        if (ru == rl) {
            fr[f] = -(ru + 1);
            continue;
        }
        label cOwn = ru, cNei = rl;
        if (ru > rl) {
            cOwn = rl;
            cNei = ru;
        }
        bool found = false;
        for (label k : ccFaces[cOwn]) {
            if (initNei[k] == cNei) {
                found = true;
                fr[f] = k;
                break;
            }
        }
        if (!found) {
            const label k = (label)initNei.size();
            ccFaces[cOwn].push_back(k);
            initNei.push_back(cNei);
            fr[f] = k;
        }
    }
    // renumber owner-major, neighbours in first-appearance order (:133-163)
    const label ncf = (label)initNei.size();
    std::vector<label> cOwner(ncf), cNeigh(ncf), faceMap(ncf);
    label k2 = 0;
    for (label c = 0; c < nc; ++c)
        for (label k : ccFaces[c]) {
            cOwner[k2] = c;
            cNeigh[k2] = initNei[k];
            faceMap[k] = k2;
            k2++;
        }
    for (label f = 0; f < nf; ++f)
        if (fr[f] >= 0) fr[f] = faceMap[fr[f]];
    faceRestrictAddr.push_back(fr);
    Ldu C;
    C.n = nc;
    C.l = cOwner;
    C.u = cNeigh;
    C.symm = true;
    coarse.push_back(C);
}

// src/matrix/lduMatrix/lduAgglomeration/lduAgglomerate.cpp:4-91 with
// continueAgglomerating (lduAgglomeration.cpp:26-43; serial LOR = local test)
void Agglomeration::agglomerate(const double* weights) {
    coarse.clear();
    restrictAddr.clear();
    faceRestrictAddr.clear();
    nCells.clear();
    std::vector<double> w(weights, weights + fine->nFaces());
    label nCreated = 0;
    while (nCreated < maxLevels - 1) {
        label nC = -1;
        std::vector<label> map = pairLevel(nC, matrixLevel(nCreated), w.data());
        if (!(nC >= nCoarsest)) break;
        nCells.push_back(nC);
        restrictAddr.push_back(map);
        buildCoarseAddressing(nCreated);
        std::vector<double> cw(coarse[nCreated].nFaces(), 0.0);
        restrictFaceField(cw.data(), w.data(), nCreated);
        w.swap(cw);
        nCreated++;  // mergeLevels_ = 1: never combineLevels (:64)
    }
}

// src/matrix/unapMatrix.hpp:378-407
void Agglomeration::restrictField(double* cf, const double* ff, label lev) const {
    const std::vector<label>& m = restrictAddr[lev];
    for (label c = 0; c < nCells[lev]; ++c) cf[c] = 0;
    for (label i = 0; i < (label)m.size(); ++i) cf[m[i]] += ff[i];
}

// src/matrix/unapMatrix.hpp:410-438
void Agglomeration::prolongField(double* ff, const double* cf, label lev) const {
    const std::vector<label>& m = restrictAddr[lev];
    for (label i = 0; i < (label)m.size(); ++i) ff[i] = cf[m[i]];
}

// src/matrix/unapMatrix.hpp:441-475
void Agglomeration::restrictFaceField(double* cf, const double* ff, label lev) const {
    const std::vector<label>& m = faceRestrictAddr[lev];
    for (label i = 0; i < coarse[lev].nFaces(); ++i) cf[i] = 0;
    for (label i = 0; i < (label)m.size(); ++i)
        if (m[i] >= 0) cf[m[i]] += ff[i];
}

// src/matrix/lduMatrix/lduAgglomeration/lduAgglomerateMatrix.cpp:3-104
void Agglomeration::agglomerateMatrix(label lev) {
    const Ldu& F = matrixLevel(lev);
    Ldu& C = coarse[lev];
    const std::vector<label>& fr = faceRestrictAddr[lev];
    C.d.assign(nCells[lev], 0.0);
    restrictField(C.d.data(), F.d.data(), lev);
    const label ncf = C.nFaces();
    if (!F.symm) {
        const double* fu = F.upper();
        const double* fl = F.lower();
        C.symm = false;
        C.up.assign(ncf, 0.0);
        C.lo.assign(ncf, 0.0);
        const std::vector<label>& ra = restrictAddr[lev];
        for (label f = 0; f < (label)fr.size(); ++f) {
            const label cf = fr[f];
            if (cf >= 0) {
                if (C.l[cf] == ra[F.l[f]]) {
                    C.up[cf] += fu[f];
                    C.lo[cf] += fl[f];
                } else if (C.u[cf] == ra[F.l[f]]) {
                    C.up[cf] += fl[f];
                    C.lo[cf] += fu[f];
                } else {
                    throw std::runtime_error("agglomerateMatrix: inconsistent addressing");
                }
            } else {
                C.d[-1 - cf] += fu[f] + fl[f];
            }
        }
    } else {
        const double* fu = F.upper();
        C.symm = true;
        C.lo.clear();
        C.up.assign(ncf, 0.0);
        for (label f = 0; f < (label)fr.size(); ++f) {
            const label cf = fr[f];
            if (cf >= 0)
                C.up[cf] += fu[f];
            else
                C.d[-1 - cf] += 2 * fu[f];
        }
    }
}

// src/solvers/PCG/PCG.cpp:35-183
Perf PCG::solve(double* x, const Ldu& A, const double* b) const {
    Perf perf;
    const label n = A.n;
    std::vector<double> p(n, 0.0), w(n, 0.0), r(n, 0.0);
    double wArA = GREAT_, wArAold = wArA;
    A.spmv(w.data(), x);
    for (label i = 0; i < n; ++i) r[i] = b[i] - w[i];
    perf.initial = normFactor(r.data(), n);
    perf.final_ = perf.initial;
    if (history) history->assign(1, perf.initial);
    do {
        wArAold = wArA;
        M->apply(w.data(), r.data(), n);
        wArA = dot(w.data(), r.data(), n);
        if (checkSingularity(perf, std::fabs(wArA))) break;
        if (perf.nIter == 0) {
            for (label i = 0; i < n; ++i) p[i] = w[i];
        } else {
            const double beta = wArA / wArAold;
            for (label i = 0; i < n; ++i) p[i] = w[i] + beta * p[i];
        }
        A.spmv(w.data(), p.data());
        const double wApA = dot(w.data(), p.data(), n);
        const double alpha = wArA / wApA;
        for (label i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * w[i];
        }
        perf.final_ = normFactor(r.data(), n);
        if (history) history->push_back(perf.final_);
    } while (++perf.nIter < maxIter && !checkConvergence(perf, tol, relTol, perf.nIter, minIter));
    return perf;
}

// src/solvers/PBiCGStab/PBiCGStab.cpp:45-459 (x86 branches)
Perf PBiCGStab::solve(double* x, const Ldu& A, const double* b) const {
    Perf perf;
    const label n = A.n;
    std::vector<double> p(n, 0.0), y(n, 0.0), r(n, 0.0), r0(n, 0.0);
    A.spmv(y.data(), x);
    for (label i = 0; i < n; ++i) {
        r[i] = b[i] - y[i];
        r0[i] = r[i];
    }
    perf.initial = normFactor(r.data(), n);
    perf.final_ = perf.initial;
    perf.previous = perf.final_;
    if (history) history->assign(1, perf.initial);
    std::vector<double> Ay(n, 0.0), s(n, 0.0), z(n, 0.0), t(n, 0.0);
    double rho = 0, alpha = 0, omega = 0;
    // flexible variant (shadowRestart > 0): |r0| and the iterations since r0
    // was (re)set
    double r0norm = perf.initial - SMALL_;
    label k = 0;
    do {
        const double rhoOld = rho;
        rho = dot(r0.data(), r.data(), n);
        if (shadowRestart > 0 && k > 0 && std::fabs(rho) < shadowRestart * r0norm * perf.final_) {
            for (label i = 0; i < n; ++i) r0[i] = r[i];
            rho = dot(r.data(), r.data(), n);
            r0norm = std::sqrt(rho);
            k = 0;
        }
        if (checkSingularity(perf, std::fabs(rho))) break;
        if ((shadowRestart > 0 ? k : perf.nIter) == 0) {
            for (label i = 0; i < n; ++i) p[i] = r[i];
        } else {
            const double beta = (rho / rhoOld) * (alpha / omega);
            for (label i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * Ay[i]);
        }
        M->apply(y.data(), p.data(), n);
        A.spmv(Ay.data(), y.data());
        const double r0Ay = dot(r0.data(), Ay.data(), n);
        alpha = rho / r0Ay;
        for (label i = 0; i < n; ++i) s[i] = r[i] - alpha * Ay[i];
        perf.final_ = normFactor(s.data(), n);
        if (checkConvergence(perf, tol, relTol, perf.nIter + 1, minIter)) {
            for (label i = 0; i < n; ++i) x[i] += alpha * y[i];
            perf.nIter++;
            if (history) history->push_back(perf.final_);
            return perf;
        }
        M->apply(z.data(), s.data(), n);
        A.spmv(t.data(), z.data());
        const double tt = sumSqr(t.data(), n);
        omega = dot(t.data(), s.data(), n);
        omega /= tt;
        for (label i = 0; i < n; ++i) {
            x[i] += alpha * y[i] + omega * z[i];
            r[i] = s[i] - omega * t[i];
        }
        perf.final_ = normFactor(r.data(), n);
        if (history) history->push_back(perf.final_);
        ++k;
    } while (++perf.nIter < maxIter && !checkConvergence(perf, tol, relTol, perf.nIter, minIter));
    return perf;
}

// Flexible CG (no reference counterpart): PCG's loop shape and stopping test,
// with the direction A-orthogonalised against the previous one (Notay 2000,
// FCG(1); Golub and Ye 1999):
//   beta_k = -(z_k . q_k-1) / (p_k-1 . q_k-1),   q_k-1 = A p_k-1,
// which in exact arithmetic equals z_k . (r_k - r_k-1) / (r_k-1 . z_k-1) and
// makes p_k . A p_k-1 = 0 for any preconditioner (non-symmetric or
// iteration-dependent).  For a symmetric linear preconditioner it reduces to
// PCG's beta = rho_k / rho_k-1 (PCG.cpp:120-134).
Perf FCG::solve(double* x, const Ldu& A, const double* b) const {
    Perf perf;
    const label n = A.n;
    std::vector<double> p(n, 0.0), q(n, 0.0), r(n, 0.0), z(n, 0.0);
    A.spmv(q.data(), x);
    for (label i = 0; i < n; ++i) r[i] = b[i] - q[i];
    perf.initial = normFactor(r.data(), n);
    perf.final_ = perf.initial;
    if (history) history->assign(1, perf.initial);
    double pq = 0.0;
    do {
        M->apply(z.data(), r.data(), n);
        const double rz = dot(r.data(), z.data(), n);
        if (checkSingularity(perf, std::fabs(rz))) break;
        if (perf.nIter == 0) {
            for (label i = 0; i < n; ++i) p[i] = z[i];
        } else {
            const double zq = dot(z.data(), q.data(), n);  // q = A p of the last iteration
            const double beta = -zq / pq;
            for (label i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        }
        A.spmv(q.data(), p.data());
        pq = dot(p.data(), q.data(), n);
        const double alpha = rz / pq;
        for (label i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
        }
        perf.final_ = normFactor(r.data(), n);
        if (history) history->push_back(perf.final_);
    } while (++perf.nIter < maxIter && !checkConvergence(perf, tol, relTol, perf.nIter, minIter));
    return perf;
}

// src/solvers/MG/scalingFactor.cpp:11-65
double MGSolver::scalingFactor(double* Acf, double* field, const double* source,
                               const Ldu& A) const {
    A.spmv(Acf, field);
    double s0 = 0.0, s1 = 0.0;
    for (label i = 0; i < A.n; ++i) {
        s0 += source[i] * field[i];
        s1 += Acf[i] * field[i];
        field[i] += (source[i] - Acf[i]) / A.d[i];
    }
    return s0 / (s1 + SMALL_);
}

// src/solvers/MG/solveCoarsestLevel.cpp:10-55
void MGSolver::solveCoarsest(double* corr, const double* src) const {
    const Ldu& C = ag->coarse[ag->size() - 1];
    for (label i = 0; i < C.n; ++i) corr[i] = 0;
    if (C.symm) {
        DICPrecond pre(C);
        PCG s(pre);
        s.tol = 0.0;
        s.relTol = 1.0e-2;
        s.maxIter = 100;
        s.solve(corr, C, src);
    } else {
        DILUPrecond pre(C);
        PBiCGStab s(pre);
        s.tol = 0.0;
        s.relTol = 1.0e-2;
        s.maxIter = 100;
        s.solve(corr, C, src);
    }
}

// src/solvers/MG/Vcycle.cpp:27-332
void MGSolver::vcycle(double* psi, const double* source, double* Apsi, double* finestCorr,
                      double* finestRes, std::vector<std::vector<double>>& corr,
                      std::vector<std::vector<double>>& src) const {
    const label coarsest = ag->size() - 1;
    if (coarsest >= 0) ag->restrictField(src[0].data(), finestRes, 0);
    for (label lv = 0; lv < coarsest; ++lv) {
        const Ldu& C = ag->coarse[lv];
        const label nc = C.n;
        if (nPre) {
            for (label i = 0; i < nc; ++i) corr[lv][i] = 0.0;
            sm[lv + 1]->smooth(corr[lv].data(), C, src[lv].data(), nPre + lv);
            std::vector<double> ACf(nc, 0.0);
            if (scale && lv < coarsest - 1) {
                const double sf = scalingFactor(ACf.data(), corr[lv].data(), src[lv].data(), C);
                for (label i = 0; i < nc; ++i) corr[lv][i] *= sf;
            }
            C.spmv(ACf.data(), corr[lv].data());
            for (label i = 0; i < nc; ++i) src[lv][i] -= ACf[i];
        }
        ag->restrictField(src[lv + 1].data(), src[lv].data(), lv + 1);
    }
    if (coarsest >= 0) solveCoarsest(corr[coarsest].data(), src[coarsest].data());
    for (label lv = coarsest - 1; lv >= 0; --lv) {
        const Ldu& C = ag->coarse[lv];
        const label nc = C.n;
        std::vector<double> pre(nc, 0.0);
        if (nPre)
            for (label i = 0; i < nc; ++i) pre[i] = corr[lv][i];
        ag->prolongField(corr[lv].data(), corr[lv + 1].data(), lv + 1);
        if (scale && lv < coarsest - 1) {
            std::vector<double> ACf(nc, 0.0);
            const double sf = scalingFactor(ACf.data(), corr[lv].data(), src[lv].data(), C);
            for (label i = 0; i < nc; ++i) corr[lv][i] *= sf;
        }
        if (nPre)
            for (label i = 0; i < nc; ++i) corr[lv][i] += pre[i];
        sm[lv + 1]->smooth(corr[lv].data(), C, src[lv].data(), nPost + lv);
    }
    const label n = finest->n;
    if (coarsest >= 0) ag->prolongField(finestCorr, corr[0].data(), 0);
    if (scale) {
        const double fsf = scalingFactor(Apsi, finestCorr, finestRes, *finest);
        for (label i = 0; i < n; ++i) psi[i] += fsf * finestCorr[i];
    } else {
        for (label i = 0; i < n; ++i) psi[i] += finestCorr[i];
    }
    sm[0]->smooth(psi, *finest, source, nFinest);
}

// src/solvers/MG/MGSolve.cpp:17-187, including the Galerkin coefficients of
// every level recomputed from the current fine coefficients on each call that
// cycles (:108-111)
Perf MGSolver::solve(double* x, const Ldu& A, const double* b) const {
    Perf perf;
    const label n = A.n;
    std::vector<double> Apsi(n, 0.0), fCorr(n, 0.0), fRes(n, 0.0);
    A.spmv(Apsi.data(), x);
    for (label i = 0; i < n; ++i) fRes[i] = b[i] - Apsi[i];
    perf.initial = normFactor(fRes.data(), n);
    perf.final_ = perf.initial;
    perf.previous = perf.final_;
    if (history) history->assign(1, perf.initial);
    if (!checkConvergence(perf, tol, relTol, perf.nIter, minIter)) {
        const label L = ag->size();
        ag->agglomerateAll();
        std::vector<std::vector<double>> corr(L), src(L);
        for (label lv = 0; lv < L; ++lv) {
            corr[lv].assign(ag->coarse[lv].n, 0.0);
            src[lv].assign(ag->coarse[lv].n, 0.0);
        }
        do {
            vcycle(x, b, Apsi.data(), fCorr.data(), fRes.data(), corr, src);
            A.spmv(Apsi.data(), x);
            for (label i = 0; i < n; ++i) fRes[i] = b[i] - Apsi[i];
            perf.final_ = normFactor(fRes.data(), n);
            if (history) history->push_back(perf.final_);
        } while (++perf.nIter < maxIter &&
                 !checkConvergence(perf, tol, relTol, perf.nIter, minIter));
    }
    return perf;
}

// Agglomeration-AMG preconditioner (src/preconditioners/MGPreconditioner/
// MGPreconditioner.cpp:6-42 sketches it; the reference's precondition is empty):
// weights |upper| (:14-19), one smoother per level (:23-29), and
// precondition(w, r) = { w = 0; MGSolver::solve(w, A, r) } with
// maxIter = minIter = nVcycles and tol = relTol = 0.
MGPrecond::MGPrecond(const Ldu& M, int smootherType, label nCoarsest, label nV, label nPre,
                     label nPost, label nFinest, const double* weights)
    : A(&M), ag(M), mg(nullptr), nVcycles(nV) {
    std::vector<double> w(M.nFaces());
    if (weights) {
        for (label f = 0; f < M.nFaces(); ++f) w[f] = weights[f];
    } else {
        for (label f = 0; f < M.nFaces(); ++f) w[f] = std::fabs(M.upper()[f]);
    }
    ag.nCoarsest = nCoarsest;
    ag.agglomerate(w.data());
    ag.agglomerateAll();
    std::vector<const Smoother*> s;
    for (label i = 0; i < ag.size(); ++i) {
        Smoother* p = (smootherType == 2) ? (Smoother*)new ChebySmoother
                                          : (Smoother*)new GSSmoother;
        owned.push_back(p);
        s.push_back(p);
    }
    mg = new MGSolver(M, ag, s);
    mg->nPre = nPre;
    mg->nPost = nPost;
    mg->nFinest = nFinest;
    mg->maxIter = nV;
    mg->minIter = nV;
    mg->tol = 0;
    mg->relTol = 0;
}

MGPrecond::~MGPrecond() {
    delete mg;
    for (auto* s : owned) delete s;
}

void MGPrecond::apply(double* w, const double* r, label n) const {
    for (label i = 0; i < n; ++i) w[i] = 0;
    mg->solve(w, *A, r);
}

}  // namespace unap_oracle
