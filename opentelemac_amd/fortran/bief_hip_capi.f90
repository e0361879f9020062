! ISO_C_BINDING interfaces to libbiefhip (include/bief_hip.h).
! The Fortran host layer (bief_hip_shim.F90) calls the HIP implementation
! only through these; they carry plain C types, as the header does.
module bief_hip_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: BHIP_OK = 0, BHIP_MAXITER = 1, BHIP_EBADARG = 2, &
       BHIP_ZERO_RHS = 3, BHIP_EDEVICE = 4
  integer(c_int), parameter :: BHIP_PREC_FP64 = 0, BHIP_PREC_MIXED = 1

  type, bind(c) :: bhip_solver_cfg
    integer(c_int) :: slv
    integer(c_int) :: precon
    integer(c_int) :: krylov
    integer(c_int) :: nitmax
    real(c_double) :: eps
    integer(c_int) :: precision
    real(c_double) :: inner_eps
    integer(c_int) :: inner_nitmax
  end type

  type, bind(c) :: bhip_solve_info
    integer(c_int) :: niter
    integer(c_int) :: inner_iters
    integer(c_int) :: outer_iters
    integer(c_int) :: relat
    real(c_double) :: relres
    integer(c_int) :: converged
  end type

  interface
    integer(c_int) function bhip_ctx_create(npoin, nelem, ndp, ikle, x, y, nplan, &
         block_nodes, ctx) bind(c, name='bhip_ctx_create')
      import :: c_int, c_ptr
      integer(c_int), value :: npoin, nelem, ndp, nplan, block_nodes
      type(c_ptr), value :: ikle, x, y
      type(c_ptr) :: ctx
    end function

    integer(c_int) function bhip_ctx_destroy(ctx) bind(c, name='bhip_ctx_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function bhip_matrix_set(ctx, typdia, typext, da, xa) &
         bind(c, name='bhip_matrix_set')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), value :: typdia, typext
      type(c_ptr), value :: da, xa
    end function

    integer(c_int) function bhip_matrix_get(ctx, typdia, da, xa) bind(c, name='bhip_matrix_get')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char) :: typdia
      type(c_ptr), value :: da, xa
    end function

    integer(c_int) function bhip_matvec(ctx, op, c, y, x) bind(c, name='bhip_matvec')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: op
      real(c_double), value :: c
      type(c_ptr), value :: y, x
    end function

    integer(c_int) function bhip_solve(ctx, cfg, x, b, info) bind(c, name='bhip_solve')
      import :: c_int, c_ptr, bhip_solver_cfg, bhip_solve_info
      type(c_ptr), value :: ctx
      type(bhip_solver_cfg) :: cfg
      type(c_ptr), value :: x, b
      type(bhip_solve_info) :: info
    end function
  end interface
end module bief_hip_capi
