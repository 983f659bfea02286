/*
 * unap_meshgen.h -- seeded synthetic LDU systems on hexahedral boxes.
 *
 * One generator feeds the GPU library, its parity tests, the CPU oracle and the
 * benchmark, so every consumer sees bit-identical systems for the same seed.
 *
 * Layout of a generated system (the LDU convention of the reference,
 * src/matrix/lduMatrix/lduMatrixSpMV.cpp:87-91):
 *   face f joins lowerAddr[f] ("owner") < upperAddr[f] ("neighbour");
 *   upper[f] = A(owner, neighbour)   (coefficient in the owner's row)
 *   lower[f] = A(neighbour, owner)   (coefficient in the neighbour's row)
 *   faces are sorted by (owner, neighbour): "upper-triangular order".
 *
 * Construction:
 *   - original cell id c = i + nx*(j + ny*k); faces +x, +y, +z per cell in
 *     ascending c give the original face ids;
 *   - cells are relabelled by a Fisher-Yates permutation drawn from the
 *     SplitMix64 stream `seed`; coefficients come from stream seed+1 (one draw
 *     per original face, in original face order), the right-hand side from
 *     stream seed+2 (one draw per cell, in new-label order) and convection
 *     fluxes from stream seed+3;
 *   - the diagonal is accumulated per original face in original face order and
 *     the shift is added last.
 */
#ifndef UNAP_MESHGEN_H
#define UNAP_MESHGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum unap_meshgen_kind {
    /* off-diagonal -1, diagonal 6 + shift for every cell */
    UNAP_MESHGEN_POISSON_CONST = 1,
    /* w_f = 10^(p0 + (p1 - p0) U); upper = lower = -w_f;
       diag_i = sum_{f touches i} w_f + shift */
    UNAP_MESHGEN_POISSON_LOGUNIF = 2,
    /* diffusion 1 per face, upwind convection with flux F = -p0 + 2 p0 U on the
       original orientation a->b (a < b); A(a,b) = -(1 + max(-F,0)),
       A(b,a) = -(1 + max(F,0)); diag_i = sum of (1 + outflow) + shift */
    UNAP_MESHGEN_CONVDIFF = 3
};

/* Number of cells and faces of an nx*ny*nz box. */
int64_t unap_meshgen_ncells(int64_t nx, int64_t ny, int64_t nz);
int64_t unap_meshgen_nfaces(int64_t nx, int64_t ny, int64_t nz);

/*
 * Fill one system.  Arrays are caller-owned: lowerAddr/upperAddr/lower/upper of
 * length nfaces, diag/rhs of length ncells.  `lower` may be NULL for the
 * symmetric kinds (lower == upper there).  `perm` (length ncells, may be NULL)
 * receives the new label of every original cell.  Returns 0 on success,
 * negative on bad arguments.
 */
int unap_meshgen_box(int64_t nx, int64_t ny, int64_t nz, int kind,
                     double p0, double p1, double shift, uint64_t seed,
                     int64_t* lowerAddr, int64_t* upperAddr,
                     double* lower, double* diag, double* upper,
                     double* rhs, int64_t* perm);

/* The SplitMix64 stream used above: next 64-bit word, and a uniform double in
   [0,1) from its top 53 bits.  Exposed so host code can draw batch sizes from
   the same generator. */
uint64_t unap_meshgen_splitmix64(uint64_t* state);
double   unap_meshgen_uniform(uint64_t* state);

#ifdef __cplusplus
}
#endif

#endif
