// galerkin.hpp -- Galerkin coarse coefficients on the GPU
// (lduAgglomerateMatrix.cpp:3-104), as segmented sums over the agglomeration
// maps, one thread per coarse cell or coarse face.
//
// The reference accumulates, for coarse level k from fine level k:
//   diag_c = 0 + sum of the children's diag in ascending fine cell order
//            (restrictField, unapMatrix.hpp:392-407), then, faces ascending,
//            + 2 upper_f (symmetric) or + (upper_f + lower_f) for every fine
//            face f inside cluster c (:86-99, :69-75);
//   upper_cf, lower_cf = 0 + the coefficients of the fine faces mapped to the
//            coarse face cf, ascending, swapped when the fine owner lies in
//            the coarse neighbour (:46-66).
// Each per-coarse-entry sum keeps that order, so every coefficient is
// bit-identical to the serial loops (built with -ffp-contract=off).
#pragma once

#include "kernels.hpp"

namespace unap_hip {

// Coarse cell c: children child[childStart[c] .. childStart[c+1]) (fine cell
// ids ascending), internal faces inner[innerStart[c] .. innerStart[c+1])
// (fine face ids ascending).
static __global__ void kGalerkinDiag(int nc, const int* __restrict__ childStart, const int* __restrict__ child,
                                     const int* __restrict__ innerStart, const int* __restrict__ inner,
                                     const double* __restrict__ fd, const double* __restrict__ fup,
                                     const double* __restrict__ flo, double* __restrict__ cd) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double acc = 0.0;
    for (int k = childStart[c]; k < childStart[c + 1]; ++k) acc = acc + fd[child[k]];
    for (int k = innerStart[c]; k < innerStart[c + 1]; ++k) {
        const int f = inner[k];
        acc = acc + (flo ? fup[f] + flo[f] : 2 * fup[f]);
    }
    cd[c] = acc;
}

// Coarse face cf: fine faces fine[start[cf] .. start[cf+1]) ascending; an
// entry ~f (negative) marks a fine face whose owner lies in the coarse
// neighbour, whose upper and lower swap sides.  flo == null: symmetric, only
// upper is formed.
static __global__ void kGalerkinFaces(int ncf, const int* __restrict__ start, const int* __restrict__ fine,
                                      const double* __restrict__ fup, const double* __restrict__ flo,
                                      double* __restrict__ cup, double* __restrict__ clo) {
    const int cf = blockIdx.x * blockDim.x + threadIdx.x;
    if (cf >= ncf) return;
    double u = 0.0, l = 0.0;
    for (int k = start[cf]; k < start[cf + 1]; ++k) {
        const int e = fine[k];
        const int f = e >= 0 ? e : ~e;
        if (!flo) {
            u = u + fup[f];
        } else if (e >= 0) {
            u = u + fup[f];
            l = l + flo[f];
        } else {
            u = u + flo[f];
            l = l + fup[f];
        }
    }
    cup[cf] = u;
    if (clo) clo[cf] = l;
}

// Values of a SELL layout from coefficients in external order: slot q holds
// src[q] = 2 face + side (side 1: upper, the owner's row; 0: lower), or -1
// for padding; useUpper: upper on both sides (DIC on asymmetric storage).
template <typename T>
__global__ void kFillSell(long long nslots, const int* __restrict__ src, const double* __restrict__ up,
                          const double* __restrict__ lo, int useUpper, T* __restrict__ val) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < nslots;
         q += (long long)gridDim.x * blockDim.x) {
        const int s = src[q];
        double v = 0.0;
        if (s >= 0) {
            const int f = s >> 1;
            v = ((s & 1) || useUpper || !lo) ? up[f] : lo[f];
        }
        val[q] = (T)v;
    }
}
template <typename T>
__global__ void kFillDiag(int n, const int* __restrict__ perm, const double* __restrict__ d, T* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (T)d[perm[i]];
}

}  // namespace unap_hip
