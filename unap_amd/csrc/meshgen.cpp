// Seeded synthetic LDU box systems (host code).  The construction is specified
// in include/unap_meshgen.h; this file is its only implementation and is shared
// by the GPU library's tests, the benchmark and the CPU oracle.
#include "unap_meshgen.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    // Lemire multiply-shift: uniform integer in [0, m).
    uint64_t below(uint64_t m) {
        return (uint64_t)(((unsigned __int128)next() * (unsigned __int128)m) >> 64);
    }
};

// Visit the faces of the box in original order: for each cell c ascending,
// its +x, +y, +z neighbour when it exists.
template <class F>
void forEachFace(int64_t nx, int64_t ny, int64_t nz, F&& f) {
    int64_t face = 0;
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                const int64_t c = i + nx * (j + ny * k);
                if (i + 1 < nx) f(face++, c, c + 1);
                if (j + 1 < ny) f(face++, c, c + nx);
                if (k + 1 < nz) f(face++, c, c + nx * ny);
            }
}

}  // namespace

extern "C" {

int64_t unap_meshgen_ncells(int64_t nx, int64_t ny, int64_t nz) { return nx * ny * nz; }

int64_t unap_meshgen_nfaces(int64_t nx, int64_t ny, int64_t nz) {
    return (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1);
}

uint64_t unap_meshgen_splitmix64(uint64_t* state) {
    SplitMix64 g{*state};
    uint64_t v = g.next();
    *state = g.s;
    return v;
}

double unap_meshgen_uniform(uint64_t* state) {
    SplitMix64 g{*state};
    double v = g.uniform();
    *state = g.s;
    return v;
}

int unap_meshgen_box(int64_t nx, int64_t ny, int64_t nz, int kind, double p0, double p1,
                     double shift, uint64_t seed, int64_t* lowerAddr, int64_t* upperAddr,
                     double* lower, double* diag, double* upper, double* rhs, int64_t* permOut) {
    if (nx < 1 || ny < 1 || nz < 1) return -1;
    if (kind < UNAP_MESHGEN_POISSON_CONST || kind > UNAP_MESHGEN_CONVDIFF) return -2;
    if (!lowerAddr || !upperAddr || !upper || !diag) return -3;
    if (kind == UNAP_MESHGEN_CONVDIFF && !lower) return -4;
    const int64_t n = nx * ny * nz;
    const int64_t nf = unap_meshgen_nfaces(nx, ny, nz);
    if (n >= (int64_t(1) << 31) || nf >= (int64_t(1) << 31)) return -5;

    // Relabelling: Fisher-Yates over the identity, stream `seed`.
    std::vector<int64_t> perm(n);
    for (int64_t c = 0; c < n; ++c) perm[c] = c;
    {
        SplitMix64 g{seed};
        for (int64_t i = n - 1; i >= 1; --i) {
            int64_t j = (int64_t)g.below((uint64_t)(i + 1));
            std::swap(perm[i], perm[j]);
        }
    }
    if (permOut) std::memcpy(permOut, perm.data(), sizeof(int64_t) * n);

    // Pass 1: faces per new owner -> owner-major slots.
    std::vector<int64_t> start(n + 1, 0);
    forEachFace(nx, ny, nz, [&](int64_t, int64_t a, int64_t b) {
        const int64_t pa = perm[a], pb = perm[b];
        start[(pa < pb ? pa : pb) + 1]++;
    });
    for (int64_t c = 0; c < n; ++c) start[c + 1] += start[c];
    std::vector<int64_t> fill(start.begin(), start.end() - 1);

    // Pass 2: coefficients in original face order, diagonal in original ids.
    std::vector<double> dOrig(n, 0.0);
    SplitMix64 gc{seed + 1};
    SplitMix64 gf{seed + 3};
    forEachFace(nx, ny, nz, [&](int64_t, int64_t a, int64_t b) {
        double aab, aba;  // A(a,b) and A(b,a) in original labels
        if (kind == UNAP_MESHGEN_POISSON_CONST) {
            aab = -1.0;
            aba = -1.0;
        } else if (kind == UNAP_MESHGEN_POISSON_LOGUNIF) {
            const double e = p0 + (p1 - p0) * gc.uniform();
            const double w = std::pow(10.0, e);
            aab = -w;
            aba = -w;
            dOrig[a] += w;
            dOrig[b] += w;
        } else {
            const double F = -p0 + 2.0 * p0 * gf.uniform();
            const double outA = F > 0.0 ? F : 0.0;    // flux leaving a
            const double outB = -F > 0.0 ? -F : 0.0;  // flux leaving b
            aab = -(1.0 + outB);
            aba = -(1.0 + outA);
            dOrig[a] += 1.0 + outA;
            dOrig[b] += 1.0 + outB;
        }
        const int64_t pa = perm[a], pb = perm[b];
        const int64_t own = pa < pb ? pa : pb;
        const int64_t nbr = pa < pb ? pb : pa;
        const int64_t slot = fill[own]++;
        lowerAddr[slot] = own;
        upperAddr[slot] = nbr;
        // upper = A(owner, neighbour); lower = A(neighbour, owner)
        upper[slot] = pa < pb ? aab : aba;
        if (lower) lower[slot] = pa < pb ? aba : aab;
    });

    // Sort each owner's faces by neighbour (segments hold at most 6 faces).
    for (int64_t c = 0; c < n; ++c) {
        for (int64_t s = start[c] + 1; s < start[c + 1]; ++s) {
            const int64_t key = upperAddr[s];
            const double up = upper[s];
            const double lo = lower ? lower[s] : 0.0;
            int64_t t = s - 1;
            while (t >= start[c] && upperAddr[t] > key) {
                upperAddr[t + 1] = upperAddr[t];
                upper[t + 1] = upper[t];
                if (lower) lower[t + 1] = lower[t];
                --t;
            }
            upperAddr[t + 1] = key;
            upper[t + 1] = up;
            if (lower) lower[t + 1] = lo;
        }
    }

    for (int64_t c = 0; c < n; ++c) {
        const double d = (kind == UNAP_MESHGEN_POISSON_CONST) ? 6.0 + shift : dOrig[c] + shift;
        diag[perm[c]] = d;
    }

    if (rhs) {
        SplitMix64 gr{seed + 2};
        for (int64_t i = 0; i < n; ++i) rhs[i] = -1.0 + 2.0 * gr.uniform();
    }
    return 0;
}

}  // extern "C"
