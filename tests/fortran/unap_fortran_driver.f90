! unap_fortran_driver.f90 -- a Fortran caller of libunap_hip_fortran.so, written
! the way the reference's Fortran users call UNAP (test/ex9f.f90,
! src/tools/fortranInterface/fortranInterface.hpp): implicit external calls,
! 8-byte integers (the library's label) and handles held in integer(8).
!
! usage: unap_fortran_driver <system.bin> <results.bin> [reference]
! ("reference": the library's sums in the reference's order, through
! unap_hip_set_reduction_order of include/unap_hip.h)
!
! system.bin (stream, native byte order), written by tests/test_gpu_fortran.py:
!   int64 n, nnz, nf, symm
!   int64 row(nnz), col(nnz)    1-based COO, grouped by row
!   real64 val(nnz), b(n)
!   int64 l(nf), u(nf)           0-based LDU face addresses
!   real64 diag(n), upper(nf), lower(nf)           coefficients, set 1
!   real64 diag2(n), upper2(nf), lower2(nf)        coefficients, set 2
!   real64 w(nf)                 agglomeration face weights
! (lower and lower2 are present only when symm = 0)
!
! results.bin: for each case, in the order below,
!   int64 num_iterations (-2 where the entry point returns none),
!   int64 matrix handle after the call, real64 residual, real64 x(n)
!
! symmetric system:
!   1 coo2ldumatrixcreat + pcgsolversolve, DIC
!   2 reordercoo + csr2ldumatrixcreat + pcgsolversolve, diagonal
!   3 ldumatrixcreat + pcgsolversolve, DIC
!   4 coo2ldumatrixcreat + mgsolversolve, |upper| weights, Gauss-Seidel
!   5 coo2ldumatrixcreat + mgsolversolve, face weights w, Chebyshev
!   6 contruct_sw_matrix + fill + contruct_solver_mg (w, Gauss-Seidel) + sw_solve_mg
!   7 the same solver after fill_sw_matrix_coefficients with set 2
! asymmetric system:
!   1 coo2ldumatrixcreat + pbicgstabsolversolve, DILU
!   2 ldumatrixcreat (lower, diag, upper) + pbicgstabsolversolve, DILU
!   3 contruct_sw_matrix + fill + contruct_solver_pbicgstab + set_precond DILU + sw_solve_pbicgstab
!   4 the same solver after fill_sw_matrix_coefficients with set 2
program unap_fortran_driver
  use iso_c_binding, only: c_int
  implicit none
  interface
    integer(c_int) function unap_hip_set_reduction_order(order) bind(C, name="unap_hip_set_reduction_order")
      import :: c_int
      integer(c_int), value :: order
    end function
  end interface
  integer(8) :: n, nnz, nf, symm, i, k
  integer(8), allocatable :: row(:), col(:), l(:), u(:), r0(:), c0(:), rs(:)
  real(8), allocatable :: val(:), b(:), d(:), up(:), lo(:), d2(:), up2(:), lo2(:), w(:), x(:), v0(:)
  integer(8) :: A, S, G, precond, aggl, smoother, maxiter, miniter, niter, symmflag, levels, coarsest
  real(8) :: tol, reltol, res
  character(len=1024) :: inpath, outpath, mode
  integer :: fin, fout

  call get_command_argument(1, inpath)
  call get_command_argument(2, outpath)
  call get_command_argument(3, mode)
  if (trim(mode) == 'reference') then
    if (unap_hip_set_reduction_order(1_c_int) /= 0) stop 2
  end if
  open(newunit=fin, file=trim(inpath), access='stream', form='unformatted', status='old')
  read(fin) n, nnz, nf, symm
  allocate(row(nnz), col(nnz), val(nnz), b(n), l(nf), u(nf), d(n), up(nf), lo(nf), d2(n), up2(nf), lo2(nf))
  allocate(w(nf), x(n), r0(nnz), c0(nnz), v0(nnz), rs(n + 1))
  read(fin) row, col, val, b, l, u, d, up
  if (symm == 0) read(fin) lo
  read(fin) d2, up2
  if (symm == 0) read(fin) lo2
  read(fin) w
  close(fin)
  open(newunit=fout, file=trim(outpath), access='stream', form='unformatted', status='replace')

  A = 0
  S = 0
  G = 0
  tol = 0.0d0
  reltol = 1.0d-8
  maxiter = 1000
  miniter = 1
  symmflag = symm

  if (symm /= 0) then
    ! 1: COO, PCG + DIC
    x = 0.0d0
    precond = 2
    call coo2ldumatrixcreat(A, val, row, col, n, nnz, symmflag)
    call pcgsolversolve(x, A, b, n, precond, tol, reltol, maxiter, miniter, niter, res)
    write(fout) niter, A, res, x
    ! 2: sorted COO (reordercoo makes it 0-based) as CSR, PCG + diagonal
    r0 = row
    c0 = col
    v0 = val
    call reordercoo(v0, r0, c0, n, nnz)
    rs = 0
    do k = 1, nnz
      rs(r0(k) + 2) = rs(r0(k) + 2) + 1
    end do
    do i = 2, n + 1
      rs(i) = rs(i) + rs(i - 1)
    end do
    x = 0.0d0
    precond = 1
    call csr2ldumatrixcreat(A, v0, rs, c0, n, nnz, symmflag)
    call pcgsolversolve(x, A, b, n, precond, tol, reltol, maxiter, miniter, niter, res)
    write(fout) niter, A, res, x
    ! 3: LDU arrays, upper passed as lower too, PCG + DIC
    x = 0.0d0
    precond = 2
    call ldumatrixcreat(A, n, nf, l, u, up, d, up)
    call pcgsolversolve(x, A, b, n, precond, tol, reltol, maxiter, miniter, niter, res)
    write(fout) niter, A, res, x
    ! 4: MG, |upper| weights, Gauss-Seidel
    x = 0.0d0
    aggl = 1
    smoother = 1
    maxiter = 100
    call coo2ldumatrixcreat(A, val, row, col, n, nnz, symmflag)
    call mgsolversolve(x, A, b, n, aggl, smoother, tol, reltol, maxiter, miniter, niter, res, w)
    write(fout) niter, A, res, x
    ! 5: MG, face weights, Chebyshev
    x = 0.0d0
    aggl = 2
    smoother = 2
    call coo2ldumatrixcreat(A, val, row, col, n, nnz, symmflag)
    call mgsolversolve(x, A, b, n, aggl, smoother, tol, reltol, maxiter, miniter, niter, res, w)
    write(fout) niter, A, res, x
    ! 6, 7: persistent matrix and MG solver
    call contruct_sw_matrix_(A, n, l, u, nf)
    call fill_sw_matrix_coefficients_(A, d, up, up)
    smoother = 1
    levels = 50
    coarsest = 50
    call contruct_solver_mg_(S, A, G, w, smoother, levels, coarsest)
    call sw_solver_mg_set_maxiter_(S, maxiter)
    call sw_solver_mg_set_reltol_(S, reltol)
    call sw_solver_mg_set_tol_(S, tol)
    x = 0.0d0
    call sw_solve_mg_(S, A, x, b, res)
    write(fout) -2_8, A, res, x
    call fill_sw_matrix_coefficients_(A, d2, up2, up2)
    x = 0.0d0
    call sw_solve_mg_(S, A, x, b, res)
    write(fout) -2_8, A, res, x
    call sw_solver_destroy_mg_(S)
    call sw_matrix_destroy_(A)
  else
    ! 1: COO, PBiCGStab + DILU
    x = 0.0d0
    precond = 3
    call coo2ldumatrixcreat(A, val, row, col, n, nnz, symmflag)
    call pbicgstabsolversolve(x, A, b, n, precond, tol, reltol, maxiter, miniter, niter, res)
    write(fout) niter, A, res, x
    ! 2: LDU arrays with a separate lower
    x = 0.0d0
    call ldumatrixcreat(A, n, nf, l, u, lo, d, up)
    call pbicgstabsolversolve(x, A, b, n, precond, tol, reltol, maxiter, miniter, niter, res)
    write(fout) niter, A, res, x
    ! 3, 4: persistent matrix and PBiCGStab solver
    call contruct_sw_matrix_(A, n, l, u, nf)
    call fill_sw_matrix_coefficients_(A, d, up, lo)
    call contruct_solver_pbicgstab_(S)
    call sw_solver_pbicgstab_set_maxiter_(S, maxiter)
    call sw_solver_pbicgstab_set_miniter_(S, miniter)
    call sw_solver_pbicgstab_set_reltol_(S, reltol)
    call sw_solver_pbicgstab_set_tol_(S, tol)
    call sw_solver_pbicgstab_set_precond_(S, A, precond)
    x = 0.0d0
    call sw_solve_pbicgstab_(S, A, x, b, res)
    write(fout) -2_8, A, res, x
    call fill_sw_matrix_coefficients_(A, d2, up2, lo2)
    x = 0.0d0
    call sw_solve_pbicgstab_(S, A, x, b, res)
    write(fout) -2_8, A, res, x
    call sw_solver_destroy_pbicgstab_(S)
    call sw_matrix_destroy_(A)
  end if
  close(fout)
  write(*, '(a)') 'unap_fortran_driver: done'
end program unap_fortran_driver
