/*
 * unap_hip_fortran.h -- the reference's Fortran / C entry points
 * (src/tools/fortranInterface/fortranInterface.hpp:13-340,
 * fortranInterface.cpp:31-1102) over the MI355X solve path, exported by
 * unap_amd/lib/libunap_hip_fortran.so with the same symbol names, so that
 * Fortran codes written against UNAP (test/ex9f.f90, ex12f.f90) link
 * unchanged.
 *
 * Conventions, as in the reference:
 *  - every integer is a `label`, int64 (SConstruct:52; the reference's own C
 *    header cInterface/fortranInterface.h declares int*, which does not match
 *    its library: here the header matches);
 *  - handles are `long int` holding the object pointer (PTR2OBJ,
 *    fortranInterface.cpp:25-28);
 *  - vectors are host arrays in the caller's cell numbering;
 *  - COO input to coo2ldumatrixcreat_ is 1-based (fortranInterface.cpp:92-93);
 *    ldumatrixcreat_, csr2ldumatrixcreat_ and contruct_sw_matrix__ take
 *    0-based addresses as given;
 *  - the one-shot solvers (pcgsolversolve_, pbicgstabsolversolve_,
 *    mgsolversolve_) consume the matrix handle (fortranInterface.cpp:267,
 *    345, 463): it is destroyed and zeroed;
 *  - final_res_norm / res_norm are the absolute final residual ||r||_2 + 1e-37.
 *
 * Differences a caller can see:
 *  - errors do not exit the process: the call reports the message on stderr,
 *    returns num_iterations = -1 and a NaN residual where it has them, and
 *    unap_hip_fortran_last_status() / unap_hip_last_error() tell what failed;
 *  - ldumatrixcreat_ of the reference aliases `lower` to `upper` and builds a
 *    symmetric matrix holding the lower coefficients when lower != upper
 *    (fortranInterface.cpp:51-58); here it builds the asymmetric matrix,
 *    unless unap_hip_fortran_set_reference_aliasing(1) asks for the
 *    reference's result;
 *  - precondition type 0 (no preconditioner) is accepted (the reference
 *    dereferences a null preconditioner);
 *  - processor interfaces (matrixinterfacescreat_, contruct_sw_matrix_
 *    interfaces__) are accepted only with zero neighbours: one domain;
 *  - contruct_solver_mg__ returns, as its agglomeration handle, the solver
 *    handle itself (the agglomeration lives inside the solver).
 */
#ifndef UNAP_HIP_FORTRAN_H
#define UNAP_HIP_FORTRAN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t unap_label;

/* Status of the last entry point of this file called on this thread
   (UNAP_HIP_OK or a UNAP_HIP_ERR_* code of unap_hip.h). */
int unap_hip_fortran_last_status(void);
/* 1: ldumatrixcreat_ reproduces the reference's aliasing of lower to upper. */
void unap_hip_fortran_set_reference_aliasing(int on);

/* ------------------------------------------------ one-shot (fortranInterface.cpp:31-470) */
void ldumatrixcreat_(long* APtrPtr, unap_label* nCells, unap_label* upperSize, unap_label* lowerAddr,
                     unap_label* upperAddr, double* lower, double* diag, double* upper);
void coo2ldumatrixcreat_(long* APtrPtr, const double* data, const unap_label* rows, const unap_label* cols,
                         const unap_label* nCells, const unap_label* size, const unap_label* symm);
void csr2ldumatrixcreat_(long* APtrPtr, const double* data, const unap_label* compRows, const unap_label* cols,
                         const unap_label* nCells, const unap_label* size, const unap_label* symm);
void matrixinterfacescreat_(long* APtrPtr, const unap_label* nNeiProcs, const unap_label* destRank,
                            const unap_label* locPosition, const unap_label* faceCells, const double* data);
/* precond: 1 diagonal, 2 DIC, 3 DILU (0: none) */
void pcgsolversolve_(double* x, long* APtrPtr, double* b, unap_label* nCells, unap_label* precond, double* tol,
                     double* relTol, unap_label* maxIter, unap_label* minIter, unap_label* num_iterations,
                     double* final_res_norm);
void pbicgstabsolversolve_(double* x, long* APtrPtr, double* b, unap_label* nCells, unap_label* precond,
                           double* tol, double* relTol, unap_label* maxIter, unap_label* minIter,
                           unap_label* num_iterations, double* final_res_norm);
/* aggl: 1 weights |upper| / ||upper||, 2 faceArea; smoother: 1 Gauss-Seidel, 2 Chebyshev */
void mgsolversolve_(double* x, long* APtrPtr, double* b, unap_label* nCells, unap_label* aggl, unap_label* smoother,
                    double* tol, double* relTol, unap_label* maxIter, unap_label* minIter,
                    unap_label* num_iterations, double* final_res_norm, double* faceArea);

/* --------------------------------------- COO utilities (fortranInterface.cpp:473-534) */
void reordercoo_(double* val, unap_label* row, unap_label* col, unap_label* nCells, unap_label* size);
void reorderuface__(unap_label* row, unap_label* col, unap_label* nCells, unap_label* size, unap_label* newOrder);
void reorderlface__(unap_label* row, unap_label* col, unap_label* nCells, unap_label* size, unap_label* newOrder);
void reordervalue__(double* val, unap_label* newOrder, unap_label* size);

/* --------------------------- persistent handles (fortranInterface.cpp:537-1102) */
void contruct_sw_matrix__(long* APtrPtr, const unap_label* nCells, const unap_label* rows, const unap_label* cols,
                          const unap_label* size);
void contruct_sw_matrix_interfaces__(long* APtrPtr, const unap_label* nNeiProcs, const unap_label* destRank,
                                     const unap_label* offDiagRows, const unap_label* offDiagStarts);
void fill_sw_matrix_coefficients__(long* APtrPtr, const double* diag, const double* upper, const double* lower);
void fill_sw_matrix_interfaces_coefficients__(long* APtrPtr, const unap_label* offDiagStarts,
                                              const double* offDiagCoeffs);
void contruct_solver_mg__(long* mgPtrPtr, long* APtrPtr, long* AgglPtr, const double* weights,
                          const unap_label* smootherType, const unap_label* maxLevels,
                          const unap_label* nCellsCoarsest);
void sw_solver_mg_set_maxiter__(long* solverPtrPtr, const unap_label* maxIter);
void sw_solver_mg_set_miniter__(long* solverPtrPtr, const unap_label* minIter);
void sw_solver_mg_set_tol__(long* solverPtrPtr, const double* tol);
void sw_solver_mg_set_reltol__(long* solverPtrPtr, const double* relTol);
void sw_solver_mg_set_npresweeps__(long* solverPtrPtr, const unap_label* num);
void sw_solver_mg_set_npostsweeps__(long* solverPtrPtr, const unap_label* num);
void sw_solver_mg_set_nfinestsweeps__(long* solverPtrPtr, const unap_label* num);
void sw_solve_mg__(long* mgPtrPtr, long* APtrPtr, double* x, double* b, double* res_norm);
void contruct_solver_pbicgstab__(long* solverPtrPtr);
void sw_solver_pbicgstab_set_maxiter__(long* solverPtrPtr, const unap_label* maxIter);
void sw_solver_pbicgstab_set_miniter__(long* solverPtrPtr, const unap_label* minIter);
void sw_solver_pbicgstab_set_tol__(long* solverPtrPtr, const double* tol);
void sw_solver_pbicgstab_set_reltol__(long* solverPtrPtr, const double* relTol);
void sw_solver_pbicgstab_set_precond__(long* solverPtrPtr, long* APtrPtr, const unap_label* precondType);
void sw_solve_pbicgstab__(long* solverPtrPtr, long* APtrPtr, double* x, double* b, double* res_norm);
void sw_matrix_destroy__(long* APtrPtr);
void sw_solver_destroy_mg__(long* solverPtrPtr);
void sw_solver_destroy_pbicgstab__(long* solverPtrPtr);

#ifdef __cplusplus
}
#endif
#endif
