// formats.cpp -- matrix formats and files around the LDU solve path, over the
// C-ABI of include/unap_hip.h: COO / CSR to LDU conversion
// (src/tools/matrixConversion/coo2ldu.cpp:3-148, matrixConversion.cpp:38-99),
// the OpenFOAM-style text dump and its reader (printUNAP.cpp:8-59,
// printUNAP.hpp:28-40, readFromOpenFOAM.cpp:9-192) and the Hypre IJ reader
// (readFromHypre.cpp:11-254, 454-523) for one domain.  Host code: these run
// once, before any solve.
#include <algorithm>
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "host_setup.hpp"
#include "unap_hip.h"

namespace unap_hip {
void setLastError(const std::string& m);  // unap_hip_last_error() of this thread
}

namespace {

struct ArgError {
    int code;
    std::string msg;
};
[[noreturn]] void fail(const std::string& m, int code = UNAP_HIP_ERR_ARG) { throw ArgError{code, m}; }

// The converted matrix in external (row) numbering.
struct Ldu {
    int64_t n = 0;
    std::vector<int64_t> l, u;
    std::vector<double> lo, d, up;
    bool symm = true;
};

// COO (any entry order) to LDU, the result coo2ldu gives on its row-sorted
// input (reorderCOO, then coo2ldu):
//  * the entries in row-major order, columns ascending within a row;
//  * diagonal entries to diag; entries above it to faces in that order, so
//    faces come sorted by owner (row) and neighbour (column): owner < neighbour;
//  * symmetric: entries below the diagonal are not read (coo2ldu.cpp:75);
//    otherwise entry (r, c), c < r, is lower of the face (c, r).  coo2ldu
//    places the k-th below-diagonal entry of column c at the k-th face of
//    owner c (sortData, matrixConversion.cpp:4-35), the same face when the
//    structure is symmetric; here every such entry is matched to its face
//    and a structure that is not symmetric is refused.
Ldu cooToLdu(int64_t n, int64_t nnz, const int64_t* rows, const int64_t* cols, const double* vals, bool symmetric,
             int64_t base) {
    if (n <= 0 || n >= (1LL << 31) || nnz < 0) fail("sizes out of range");
    if (nnz > 0 && !(rows && cols && vals)) fail("null COO array");
    std::vector<int64_t> start(n + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) {
        const int64_t r = rows[k] - base, c = cols[k] - base;
        if (r < 0 || r >= n || c < 0 || c >= n) fail("COO index out of range");
        start[r + 1]++;
    }
    for (int64_t i = 0; i < n; ++i) start[i + 1] += start[i];
    std::vector<int64_t> at(start.begin(), start.end() - 1), ord(nnz);
    for (int64_t k = 0; k < nnz; ++k) ord[at[rows[k] - base]++] = k;  // stable by row
    for (int64_t i = 0; i < n; ++i) {
        auto b = ord.begin() + start[i], e = ord.begin() + start[i + 1];
        std::stable_sort(b, e, [&](int64_t x, int64_t y) { return cols[x] < cols[y]; });
        for (auto p = b; p + 1 < e; ++p)
            if (cols[*p] == cols[*(p + 1)]) fail("duplicate COO entry");
    }
    Ldu A;
    A.n = n;
    A.symm = symmetric;
    A.d.assign(n, 0.0);
    std::vector<int64_t> ownStart(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        ownStart[i] = (int64_t)A.u.size();
        for (int64_t q = start[i]; q < start[i + 1]; ++q) {
            const int64_t k = ord[q], c = cols[k] - base;
            if (c == i) {
                A.d[i] = vals[k];
            } else if (c > i) {
                A.l.push_back(i);
                A.u.push_back(c);
                A.up.push_back(vals[k]);
            }
        }
    }
    ownStart[n] = (int64_t)A.u.size();
    if (symmetric) return A;
    A.lo.assign(A.u.size(), 0.0);
    std::vector<char> seen(A.u.size(), 0);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t q = start[i]; q < start[i + 1]; ++q) {
            const int64_t k = ord[q], c = cols[k] - base;
            if (c >= i) continue;
            // face (owner c, neighbour i): faces of owner c ascend in neighbour
            auto b = A.u.begin() + ownStart[c], e = A.u.begin() + ownStart[c + 1];
            auto it = std::lower_bound(b, e, i);
            if (it == e || *it != i) fail("COO structure is not symmetric (a lower entry has no upper one)");
            const int64_t f = it - A.u.begin();
            A.lo[f] = vals[k];
            seen[f] = 1;
        }
    for (char s : seen)
        if (!s) fail("COO structure is not symmetric (an upper entry has no lower one)");
    return A;
}

// The matrix builders' conversion: on the device for matrices of at least
// deviceSetupMin() cells where a GPU is present (setup_dev.hip, the same
// result), on the host otherwise.
Ldu cooToLduBuild(int64_t n, int64_t nnz, const int64_t* rows, const int64_t* cols, const double* vals,
                  bool symmetric, int64_t base) {
    const int mn = unap_hip::deviceSetupMin();
    if (mn >= 0 && n >= mn && n < (1LL << 31) && nnz > 0 && rows && cols && vals) {
        Ldu A;
        std::string err;
        if (unap_hip::cooToLduDevice(n, nnz, rows, cols, vals, symmetric, base, A.l, A.u, A.lo, A.d, A.up, err)) {
            if (!err.empty()) fail(err);
            A.n = n;
            A.symm = symmetric;
            return A;
        }
    }
    return cooToLdu(n, nnz, rows, cols, vals, symmetric, base);
}

int create(const Ldu& A, unap_hip_ldu* out) {
    const int64_t nf = (int64_t)A.u.size();
    return unap_hip_ldu_create(A.n, nf, A.l.data(), A.u.data(), A.symm ? nullptr : A.lo.data(), A.d.data(),
                               A.up.data(), out);
}

// Lines of a text file (getline semantics: no empty last line for a final
// newline).
std::vector<std::string> readLines(const char* path) {
    if (!path) fail("null path");
    FILE* f = std::fopen(path, "r");
    if (!f) fail(std::string("cannot open ") + path + ": " + std::strerror(errno));
    std::vector<std::string> lines;
    std::string cur;
    int ch;
    while ((ch = std::fgetc(f)) != EOF) {
        if (ch == '\n') {
            lines.push_back(cur);
            cur.clear();
        } else {
            cur.push_back((char)ch);
        }
    }
    if (!cur.empty()) lines.push_back(cur);
    std::fclose(f);
    return lines;
}

// Whitespace-separated tokens of one line.
std::vector<std::string> tokens(const std::string& s) {
    std::vector<std::string> t;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\r')) ++i;
        size_t j = i;
        while (j < s.size() && !(s[j] == ' ' || s[j] == '\t' || s[j] == '\r')) ++j;
        if (j > i) t.push_back(s.substr(i, j - i));
        i = j;
    }
    return t;
}
int64_t toLabel(const std::string& s) { return std::strtoll(s.c_str(), nullptr, 10); }
double toScalar(const std::string& s) { return std::strtod(s.c_str(), nullptr); }
const std::string& tok(const std::vector<std::string>& t, size_t k, const char* what) {
    if (k >= t.size()) fail(std::string("malformed line: missing ") + what);
    return t[k];
}

template <class F>
int guard(F&& f) {
    try {
        f();
        return UNAP_HIP_OK;
    } catch (const ArgError& e) {
        unap_hip::setLastError(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        unap_hip::setLastError("host out of memory");
        return UNAP_HIP_ERR_INTERNAL;
    } catch (const std::exception& e) {  // the device conversion's HIP errors
        unap_hip::setLastError(e.what());
        return UNAP_HIP_ERR_HIP;
    }
}

// Faces of every cell in ascending face id (the order the reference's
// doubled COO lists a row's entries in, cooRCM.cpp:484-524).
struct CellFaces {
    std::vector<int64_t> start, face;
};
CellFaces cellFaces(int64_t n, int64_t nf, const int64_t* l, const int64_t* u) {
    CellFaces C;
    C.start.assign(n + 1, 0);
    for (int64_t f = 0; f < nf; ++f) {
        if (l[f] < 0 || l[f] >= n || u[f] < 0 || u[f] >= n || l[f] == u[f]) fail("face address out of range");
        C.start[l[f] + 1]++;
        C.start[u[f] + 1]++;
    }
    for (int64_t c = 0; c < n; ++c) C.start[c + 1] += C.start[c];
    C.face.resize(2 * nf);
    std::vector<int64_t> at(C.start.begin(), C.start.end() - 1);
    for (int64_t f = 0; f < nf; ++f) {
        C.face[at[l[f]]++] = f;
        C.face[at[u[f]]++] = f;
    }
    return C;
}

// Reverse Cuthill-McKee (ReorderingSSM, cooRCM.cpp:79-235): each connected
// component starts at its unvisited cell of least degree (the lowest id among
// equals); a visited cell's unvisited neighbours join the queue by ascending
// degree, equal degrees in the cell's face order; the sequence is reversed.
// Returns new_of_old (the new label of every cell).  The reference sorts the
// neighbours with a comparator that answers "degrees differ", which is no
// ordering, and looks the component seed up by cell id in a shrunken list:
// its order is defined only where neither matters; here the ordering is the
// one its comments describe.
std::vector<int64_t> rcmOrder(int64_t n, int64_t nf, const int64_t* l, const int64_t* u) {
    const CellFaces C = cellFaces(n, nf, l, u);
    auto deg = [&](int64_t c) { return C.start[c + 1] - C.start[c]; };
    // cells by (degree, id): the seed of each component is the first unvisited one
    std::vector<int64_t> byDeg(n);
    for (int64_t c = 0; c < n; ++c) byDeg[c] = c;
    std::stable_sort(byDeg.begin(), byDeg.end(), [&](int64_t a, int64_t b) { return deg(a) < deg(b); });
    std::vector<char> seen(n, 0);
    std::vector<int64_t> cm;
    cm.reserve(n);
    std::vector<int64_t> nb;
    size_t seedAt = 0;
    while ((int64_t)cm.size() < n) {
        while (seen[byDeg[seedAt]]) ++seedAt;
        const int64_t s0 = byDeg[seedAt];
        seen[s0] = 1;
        size_t head = cm.size();
        cm.push_back(s0);
        for (; head < cm.size(); ++head) {
            const int64_t c = cm[head];
            nb.clear();
            for (int64_t k = C.start[c]; k < C.start[c + 1]; ++k) {
                const int64_t f = C.face[k];
                const int64_t o = l[f] == c ? u[f] : l[f];
                if (!seen[o]) {
                    seen[o] = 1;
                    nb.push_back(o);
                }
            }
            std::stable_sort(nb.begin(), nb.end(), [&](int64_t a, int64_t b) { return deg(a) < deg(b); });
            cm.insert(cm.end(), nb.begin(), nb.end());
        }
    }
    std::vector<int64_t> newOfOld(n);
    for (int64_t k = 0; k < n; ++k) newOfOld[cm[n - 1 - k]] = k;
    return newOfOld;
}

}  // namespace

extern "C" {

int unap_hip_coo_to_ldu(int64_t n, int64_t nnz, const int64_t* rows, const int64_t* cols, const double* vals,
                        int symmetric, int64_t base, int64_t* n_faces, int64_t* lower_addr, int64_t* upper_addr,
                        double* lower, double* diag, double* upper) {
    return guard([&] {
        const Ldu A = cooToLdu(n, nnz, rows, cols, vals, symmetric != 0, base);
        const int64_t nf = (int64_t)A.u.size();
        if (n_faces) *n_faces = nf;
        if (lower_addr) std::memcpy(lower_addr, A.l.data(), sizeof(int64_t) * nf);
        if (upper_addr) std::memcpy(upper_addr, A.u.data(), sizeof(int64_t) * nf);
        if (upper) std::memcpy(upper, A.up.data(), sizeof(double) * nf);
        if (lower) std::memcpy(lower, A.symm ? A.up.data() : A.lo.data(), sizeof(double) * nf);
        if (diag) std::memcpy(diag, A.d.data(), sizeof(double) * n);
    });
}

int unap_hip_ldu_create_coo(int64_t n, int64_t nnz, const int64_t* rows, const int64_t* cols, const double* vals,
                            int symmetric, int64_t base, unap_hip_ldu* out) {
    int rc = UNAP_HIP_OK;
    const int g = guard([&] {
        if (!out) fail("out is null");
        rc = create(cooToLduBuild(n, nnz, rows, cols, vals, symmetric != 0, base), out);
    });
    return g != UNAP_HIP_OK ? g : rc;
}

// csr2ldu (coo2ldu.cpp:117-148): the row of every entry from the row starts,
// then the COO conversion.
int unap_hip_ldu_create_csr(int64_t n, const int64_t* row_start, const int64_t* cols, const double* vals,
                            int symmetric, int64_t base, unap_hip_ldu* out) {
    int rc = UNAP_HIP_OK;
    const int g = guard([&] {
        if (!out || !row_start) fail("null argument");
        if (n <= 0 || n >= (1LL << 31)) fail("sizes out of range");
        const int64_t nnz = row_start[n] - row_start[0];
        if (nnz < 0) fail("row starts decrease");
        std::vector<int64_t> rows(nnz);
        for (int64_t i = 0; i < n; ++i) {
            if (row_start[i + 1] < row_start[i]) fail("row starts decrease");
            for (int64_t k = row_start[i]; k < row_start[i + 1]; ++k) rows[k - row_start[0]] = i + base;
        }
        rc = create(cooToLduBuild(n, nnz, rows.data(), cols, vals, symmetric != 0, base), out);
    });
    return g != UNAP_HIP_OK ? g : rc;
}

// constructLDUMatrixFromOpenFOAM (readFromOpenFOAM.cpp:9-139): a header
// "nCells: N nnz: M symm: S", N lines "i i diag", then one line per face
// "owner neighbour upper [lower]".
int unap_hip_ldu_read_openfoam(const char* path, unap_hip_ldu* out) {
    int rc = UNAP_HIP_OK;
    const int g = guard([&] {
        if (!out) fail("out is null");
        const std::vector<std::string> lines = readLines(path);
        if (lines.empty()) fail("empty matrix file");
        const std::vector<std::string> h = tokens(lines[0]);
        Ldu A;
        A.n = toLabel(tok(h, 1, "nCells"));
        const int64_t nnz = toLabel(tok(h, 3, "nnz"));
        A.symm = toLabel(tok(h, 5, "symm")) != 0;
        const int64_t nf = A.symm ? nnz - A.n : (nnz - A.n) / 2;
        if (A.n <= 0 || nf < 0 || (int64_t)lines.size() != A.n + nf + 1)
            fail("matrix file: " + std::to_string(lines.size() - 1) + " lines for nCells " + std::to_string(A.n) +
                 " and nFaces " + std::to_string(nf));
        A.d.resize(A.n);
        A.l.resize(nf);
        A.u.resize(nf);
        A.up.resize(nf);
        if (!A.symm) A.lo.resize(nf);
        for (int64_t i = 0; i < A.n; ++i) A.d[i] = toScalar(tok(tokens(lines[1 + i]), 2, "diag"));
        for (int64_t f = 0; f < nf; ++f) {
            const std::vector<std::string> t = tokens(lines[1 + A.n + f]);
            A.l[f] = toLabel(tok(t, 0, "owner"));
            A.u[f] = toLabel(tok(t, 1, "neighbour"));
            A.up[f] = toScalar(tok(t, 2, "upper"));
            if (!A.symm) A.lo[f] = toScalar(tok(t, 3, "lower"));
        }
        rc = create(A, out);
    });
    return g != UNAP_HIP_OK ? g : rc;
}

// printLDUMatrix (printUNAP.cpp:8-59), written to `path` as given (the
// reference appends "_<rank>.txt" to its name argument): values in
// scientific notation with 15 digits after the point.
int unap_hip_ldu_write_openfoam(unap_hip_ldu A, const char* path) {
    int rc = UNAP_HIP_OK;
    const int g = guard([&] {
        if (!A || !path) fail("null argument");
        int64_t n = 0, nf = 0;
        int symm = 0;
        if ((rc = unap_hip_ldu_info(A, &n, &nf, &symm, nullptr, nullptr)) != UNAP_HIP_OK) return;
        std::vector<int64_t> l(nf), u(nf);
        std::vector<double> lo(nf), d(n), up(nf);
        if ((rc = unap_hip_ldu_get(A, l.data(), u.data(), lo.data(), d.data(), up.data())) != UNAP_HIP_OK) return;
        FILE* f = std::fopen(path, "w");
        if (!f) fail(std::string("cannot open ") + path + ": " + std::strerror(errno));
        const int64_t nnz = n + nf + (symm ? 0 : nf);
        std::fprintf(f, "nCells: %" PRId64 " nnz: %" PRId64 " symm: %d\n", n, nnz, symm ? 1 : 0);
        for (int64_t i = 0; i < n; ++i) std::fprintf(f, "%" PRId64 " %" PRId64 " %.15e\n", i, i, d[i]);
        for (int64_t k = 0; k < nf; ++k) {
            std::fprintf(f, "%" PRId64 " %" PRId64 " %.15e ", l[k], u[k], up[k]);
            if (!symm) std::fprintf(f, "%.15e", lo[k]);
            std::fprintf(f, "\n");
        }
        if (std::fclose(f) != 0) fail(std::string("cannot write ") + path);
    });
    return g != UNAP_HIP_OK ? g : rc;
}

// constructVectorFromOpenFOAM (readFromOpenFOAM.cpp:140-192): "nCells: N"
// then one value a line.
int unap_hip_vector_read_openfoam(const char* path, int64_t n, double* out) {
    return guard([&] {
        if (!out) fail("null output");
        const std::vector<std::string> lines = readLines(path);
        if (lines.empty()) fail("empty vector file");
        const int64_t m = toLabel(tok(tokens(lines[0]), 1, "nCells"));
        if ((int64_t)lines.size() != m + 1) fail("vector file: line count does not match nCells");
        if (m != n) fail("vector file holds " + std::to_string(m) + " values, " + std::to_string(n) + " expected");
        for (int64_t i = 0; i < n; ++i) out[i] = toScalar(tok(tokens(lines[1 + i]), 0, "value"));
    });
}

// printVector (printUNAP.hpp:28-40), written to `path` as given.
int unap_hip_vector_write_openfoam(const char* path, int64_t n, const double* v) {
    return guard([&] {
        if (!path || (n > 0 && !v) || n < 0) fail("bad argument");
        FILE* f = std::fopen(path, "w");
        if (!f) fail(std::string("cannot open ") + path + ": " + std::strerror(errno));
        std::fprintf(f, "nCells: %" PRId64 "\n", n);
        for (int64_t i = 0; i < n; ++i) std::fprintf(f, "%.15e\n", v[i]);
        if (std::fclose(f) != 0) fail(std::string("cannot write ") + path);
    });
}

// constructLDUMatrixFromHypre (readFromHypre.cpp:11-254) for one domain: a
// header "rStart rEnd cStart cEnd", then one "row col value" line an entry
// (global indices); the rows' entries go through the COO conversion with
// asymmetric storage (readFromHypre.cpp:215-223 passes symm = 0).  Columns
// outside [cStart, cEnd] couple to other processes' cells, which one domain
// does not have: refused.
int unap_hip_ldu_read_hypre(const char* path, unap_hip_ldu* out) {
    int rc = UNAP_HIP_OK;
    const int g = guard([&] {
        if (!out) fail("out is null");
        const std::vector<std::string> lines = readLines(path);
        if (lines.empty()) fail("empty matrix file");
        const std::vector<std::string> h = tokens(lines[0]);
        const int64_t rs = toLabel(tok(h, 0, "rStart")), re = toLabel(tok(h, 1, "rEnd"));
        const int64_t cs = toLabel(tok(h, 2, "cStart")), ce = toLabel(tok(h, 3, "cEnd"));
        const int64_t n = re - rs + 1, nnz = (int64_t)lines.size() - 1;
        std::vector<int64_t> r(nnz), c(nnz);
        std::vector<double> v(nnz);
        for (int64_t k = 0; k < nnz; ++k) {
            const std::vector<std::string> t = tokens(lines[1 + k]);
            const int64_t col = toLabel(tok(t, 1, "column"));
            if (col < cs || col > ce) fail("the matrix couples to cells of other processes (one domain only)");
            r[k] = toLabel(tok(t, 0, "row")) - rs;
            c[k] = col - cs;
            v[k] = toScalar(tok(t, 2, "value"));
        }
        rc = create(cooToLduBuild(n, nnz, r.data(), c.data(), v.data(), false, 0), out);
    });
    return g != UNAP_HIP_OK ? g : rc;
}

// constructVectorFromHypre (readFromHypre.cpp:454-523): "rStart rEnd", then
// "row value" a line.
int unap_hip_vector_read_hypre(const char* path, int64_t n, double* out) {
    return guard([&] {
        if (!out) fail("null output");
        const std::vector<std::string> lines = readLines(path);
        if (lines.empty()) fail("empty vector file");
        const std::vector<std::string> h = tokens(lines[0]);
        const int64_t m = toLabel(tok(h, 1, "rEnd")) - toLabel(tok(h, 0, "rStart")) + 1;
        if ((int64_t)lines.size() != m + 1) fail("vector file: line count does not match the row range");
        if (m != n) fail("vector file holds " + std::to_string(m) + " values, " + std::to_string(n) + " expected");
        for (int64_t i = 0; i < n; ++i) out[i] = toScalar(tok(tokens(lines[1 + i]), 1, "value"));
    });
}

// ReverseCuthillMckee over the LDU graph (rcmLDU_nowrite, cooRCM.cpp:477-637).
int unap_hip_rcm_order(int64_t n, int64_t n_faces, const int64_t* lower_addr, const int64_t* upper_addr,
                       int64_t* new_of_old) {
    return guard([&] {
        if (n <= 0 || n >= (1LL << 31) || n_faces < 0) fail("sizes out of range");
        if (!new_of_old || (n_faces > 0 && !(lower_addr && upper_addr))) fail("null argument");
        const std::vector<int64_t> r = rcmOrder(n, n_faces, lower_addr, upper_addr);
        std::memcpy(new_of_old, r.data(), sizeof(int64_t) * n);
    });
}

// rcmLDU_rewrite (cooRCM.cpp:639-785) for a given cell order: cells
// renumbered, each face turned so that its owner has the lower new label
// (upper and lower coefficients swap), faces stably sorted by owner, then
// neighbour.  new_face_of_old: rcmLDU_nowrite's postEdgeOrder, each old
// face's new position, ~position when it turned around.
int unap_hip_ldu_renumber(int64_t n, int64_t n_faces, const int64_t* lower_addr, const int64_t* upper_addr,
                          const int64_t* new_of_old, const double* lower, const double* diag, const double* upper,
                          int64_t* lower_addr_out, int64_t* upper_addr_out, int64_t* new_face_of_old,
                          double* lower_out, double* diag_out, double* upper_out) {
    return guard([&] {
        if (n <= 0 || n >= (1LL << 31) || n_faces < 0) fail("sizes out of range");
        if (!new_of_old || (n_faces > 0 && !(lower_addr && upper_addr))) fail("null argument");
        std::vector<char> hit(n, 0);
        for (int64_t c = 0; c < n; ++c) {
            const int64_t v = new_of_old[c];
            if (v < 0 || v >= n || hit[v]) fail("new_of_old is not a permutation of the cells");
            hit[v] = 1;
        }
        std::vector<int64_t> ol(n_faces), ou(n_faces), order(n_faces);
        std::vector<char> turned(n_faces, 0);
        for (int64_t f = 0; f < n_faces; ++f) {
            if (lower_addr[f] < 0 || lower_addr[f] >= n || upper_addr[f] < 0 || upper_addr[f] >= n ||
                lower_addr[f] == upper_addr[f])
                fail("face address out of range");
            int64_t a = new_of_old[lower_addr[f]], b = new_of_old[upper_addr[f]];
            if (a > b) {
                std::swap(a, b);
                turned[f] = 1;
            }
            ol[f] = a;
            ou[f] = b;
            order[f] = f;
        }
        std::stable_sort(order.begin(), order.end(),
                         [&](int64_t x, int64_t y) { return ol[x] != ol[y] ? ol[x] < ol[y] : ou[x] < ou[y]; });
        if ((upper_out || lower_out) && n_faces > 0 && !upper) fail("coefficient output without upper");
        const double* lw = lower ? lower : upper;  // symmetric: lower is upper
        for (int64_t k = 0; k < n_faces; ++k) {
            const int64_t f = order[k];
            if (lower_addr_out) lower_addr_out[k] = ol[f];
            if (upper_addr_out) upper_addr_out[k] = ou[f];
            if (new_face_of_old) new_face_of_old[f] = turned[f] ? ~k : k;
            if (upper_out) upper_out[k] = turned[f] ? lw[f] : upper[f];
            if (lower_out) lower_out[k] = turned[f] ? upper[f] : lw[f];
        }
        if (diag_out) {
            if (!diag) fail("diag_out without diag");
            for (int64_t c = 0; c < n; ++c) diag_out[new_of_old[c]] = diag[c];
        }
    });
}

}  // extern "C"
