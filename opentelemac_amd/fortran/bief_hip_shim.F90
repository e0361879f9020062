! Host layer with BIEF's own signatures, over the HIP C ABI.
!
! SUBROUTINE SOLVE(X, A, B, TB, CFG, INFOGR, MESH, AUX)
!     sources/utils/bief/solve.f:2-5, interface bief.f:5073
! SUBROUTINE MATVEC(OP, X, A, Y, C, MESH, LEGO)
!     sources/utils/bief/matvec.f:98-103, interface bief.f:2392
! SUBROUTINE MATRBL(OP, X, A, Y, C, MESH)
!     sources/utils/bief/matrbl.f:85, interface bief.f:2309
!
! Linked in place of solve.o / matvec.o / matrbl.o, these route the
! linear-system work of every caller (telemac2d PROPAG, telemac3d
! WAVE_EQUATION / DIFF3D / PREDIV, ...) to libbiefhip. Storage, ownership
! and side effects follow the reference: X in/out, B overwritten by D*B, A
! left preconditioned (D*A*D, TYPDIA 'I' for PRECON 2) when the diagonal
! preconditioner is used. Errors that the reference stops on (PLANTE) stop
! here too; non-convergence is reported in the listing, as GRACJG does.
!
! Supported: EBE storage (A%STO = 1) on P1 triangles (11) and P1 prisms
! (41), segment storage (A%STO = 3) on P1 triangles, a single matrix or a
! block of one matrix, SLV 1 (CG), 6 (BiCGSTAB), 7 (GMRES), PRECON 0/1, 2, 5.
!
! Compiled with -DBHIP_OWN_TYPES it uses the stand-alone type subset in
! bief_hip_types.f90; otherwise the real BIEF_DEF and DECLARATIONS_SPECIAL.

module bief_hip_state
  use iso_c_binding
  use bief_hip_capi
  implicit none
  type(c_ptr), save :: bhip_ctx_h = c_null_ptr
  integer, save :: ctx_npoin = -1, ctx_nelem = -1, ctx_ndp = -1
  integer(c_intptr_t), save :: ctx_key = 0
  ! precision of SOLVE: BHIP_PREC_FP64 (default, the reference arithmetic) or
  ! BHIP_PREC_MIXED; -1 = take BHIP_PRECISION from the environment
  integer, save :: bhip_precision = -1
  real(c_double), save :: bhip_inner_eps = 1.d-4
  integer, save :: bhip_block_nodes = 0
contains

  subroutine bhip_set_precision(mode, inner_eps)
    integer, intent(in) :: mode
    double precision, intent(in), optional :: inner_eps
    bhip_precision = mode
    if (present(inner_eps)) bhip_inner_eps = inner_eps
  end subroutine

  integer function bhip_precision_mode()
    character(len=16) :: v
    integer :: ln, st
    if (bhip_precision < 0) then
      bhip_precision = BHIP_PREC_FP64
      call get_environment_variable('BHIP_PRECISION', v, ln, st)
      if (st == 0 .and. ln > 0) then
        if (v(1:5) == 'mixed' .or. v(1:5) == 'MIXED') bhip_precision = BHIP_PREC_MIXED
      end if
    end if
    bhip_precision_mode = bhip_precision
  end function

  ! Per-mesh context, created on first use and kept while the mesh (its
  ! connectivity array and sizes) stays the same.
  function bhip_mesh_ctx(npoin, nelem, nelmax, ndp, ikle, x, y, have_xy, nplan) result(h)
    integer, intent(in) :: npoin, nelem, nelmax, ndp, nplan
    integer, intent(in), target :: ikle(nelmax, ndp)
    double precision, intent(in), target :: x(*), y(*)
    logical, intent(in) :: have_xy
    type(c_ptr) :: h
    integer(c_int), allocatable, target :: ik0(:,:)
    integer(c_intptr_t) :: key
    integer :: rc, k
    type(c_ptr) :: px, py
    key = transfer(c_loc(ikle(1, 1)), key)
    if (c_associated(bhip_ctx_h) .and. key == ctx_key .and. npoin == ctx_npoin .and. &
        nelem == ctx_nelem .and. ndp == ctx_ndp) then
      h = bhip_ctx_h
      return
    end if
    if (c_associated(bhip_ctx_h)) rc = bhip_ctx_destroy(bhip_ctx_h)
    bhip_ctx_h = c_null_ptr
    allocate(ik0(nelem, ndp))
    do k = 1, ndp
      ik0(:, k) = ikle(1:nelem, k) - 1
    end do
    px = c_null_ptr
    py = c_null_ptr
    if (have_xy) then
      px = c_loc(x(1))
      py = c_loc(y(1))
    end if
    rc = bhip_ctx_create(int(npoin, c_int), int(nelem, c_int), int(ndp, c_int), &
         c_loc(ik0(1, 1)), px, py, int(nplan, c_int), int(bhip_block_nodes, c_int), &
         bhip_ctx_h)
    deallocate(ik0)
    if (rc /= BHIP_OK) call bhip_fail('BHIP: MESH CONTEXT CREATION FAILED', rc)
    ctx_key = key
    ctx_npoin = npoin
    ctx_nelem = nelem
    ctx_ndp = ndp
    h = bhip_ctx_h
  end function

  subroutine bhip_fail(msg, rc)
#ifndef BHIP_OWN_TYPES
    use declarations_special
#endif
    character(len=*), intent(in) :: msg
    integer, intent(in) :: rc
#ifndef BHIP_OWN_TYPES
    write(lu, '(1X,A,A,I3)') msg, ' STATUS ', rc
    call plante(1)
    stop
#else
    write(6, '(1X,A,A,I3)') msg, ' STATUS ', rc
    error stop 2
#endif
  end subroutine

  integer function bhip_ndp(ielm)
    integer, intent(in) :: ielm
    select case (ielm)
    case (11)
      bhip_ndp = 3
    case (41)
      bhip_ndp = 6
    case default
      bhip_ndp = 0
    end select
  end function

  ! off-diagonal terms per element (ndp 2: per segment, XA1 and XA2)
  integer function bhip_ncoef(ndp, typext)
    integer, intent(in) :: ndp
    character(len=1), intent(in) :: typext
    bhip_ncoef = 0
    if (typext == 'S') bhip_ncoef = ndp * (ndp - 1) / 2
    if (typext == 'Q') bhip_ncoef = ndp * (ndp - 1)
  end function

  integer function bhip_opcode(op)
    character(len=8), intent(in) :: op
    select case (op)
    case ('X=AY    ')
      bhip_opcode = 0
    case ('X=CAY   ')
      bhip_opcode = 1
    case ('X=-AY   ')
      bhip_opcode = 2
    case ('X=X+AY  ')
      bhip_opcode = 3
    case ('X=X-AY  ')
      bhip_opcode = 4
    case ('X=X+CAY ')
      bhip_opcode = 5
    case ('X=TAY   ')
      bhip_opcode = 6
    case ('X=-TAY  ')
      bhip_opcode = 7
    case ('X=X+TAY ')
      bhip_opcode = 8
    case ('X=X-TAY ')
      bhip_opcode = 9
    case ('X=X+CTAY')
      bhip_opcode = 10
    case default
      bhip_opcode = -1
    end select
  end function

  ! A%X (NELMAX, NCOEF) -> contiguous (NELEM, NCOEF) for the C ABI and back
  subroutine bhip_pack_xa(xa, nelmax, nelem, ncoef, packed)
    integer, intent(in) :: nelmax, nelem, ncoef
    double precision, intent(in) :: xa(nelmax, *)
    double precision, intent(out) :: packed(nelem, *)
    integer :: c
    do c = 1, ncoef
      packed(:, c) = xa(1:nelem, c)
    end do
  end subroutine

  subroutine bhip_unpack_xa(packed, nelem, nelmax, ncoef, xa)
    integer, intent(in) :: nelmax, nelem, ncoef
    double precision, intent(in) :: packed(nelem, *)
    double precision, intent(inout) :: xa(nelmax, *)
    integer :: c
    do c = 1, ncoef
      xa(1:nelem, c) = packed(:, c)
    end do
  end subroutine
end module bief_hip_state


! ---------------------------------------------------------------------------
module bief_hip_load
#ifdef BHIP_OWN_TYPES
  use bief_hip_types
#else
  use bief_def
#endif
  use bief_hip_state
  implicit none
contains

! uploads a matrix (A%D, A%X, A%TYPDIA, A%TYPEXT) to the mesh context:
! element-by-element storage (A%STO = 1) over IKLE, or segment storage
! (A%STO = 3, P1 triangles) over GLOSEG, whose segments the context treats
! as two-node elements. ldxa: the leading dimension of A%X.
subroutine bhip_load_matrix(a, mesh, h, ndp, npoin, nelem, ldxa)
  type(bief_obj), intent(in), target :: a
  type(bief_mesh), intent(inout), target :: mesh
  type(c_ptr), intent(out) :: h
  integer, intent(out) :: ndp, npoin, nelem, ldxa
  integer :: nelmax, ncoef, rc
  double precision, allocatable, target :: xa(:,:)
  double precision, target :: dummy(1)
  logical :: have_xy
  if (a%type /= 3) call bhip_fail('BHIP: A IS NOT A MATRIX', BHIP_EBADARG)
  if (a%sto /= 1 .and. a%sto /= 3) call bhip_fail( &
       'BHIP: MATRIX STORAGE 1 (EBE) OR 3 (SEGMENTS) ONLY', BHIP_EBADARG)
  if (a%elmlin /= a%elmcol) call bhip_fail('BHIP: RECTANGULAR MATRICES NOT SUPPORTED', &
       BHIP_EBADARG)
  ndp = bhip_ndp(a%elmlin)
  if (ndp == 0) call bhip_fail('BHIP: ELEMENT NOT SUPPORTED (P1 TRIANGLES AND PRISMS)', &
       BHIP_EBADARG)
  npoin = a%d%dim1
  nelem = mesh%nelem
  nelmax = mesh%nelmax
  ldxa = nelmax
  if (a%sto == 3) then
    if (ndp /= 3) call bhip_fail('BHIP: SEGMENT STORAGE ON P1 TRIANGLES ONLY', BHIP_EBADARG)
    ndp = 2
    nelem = mesh%nseg
    nelmax = mesh%gloseg%maxdim1
    ! XA2 follows XA1 directly: XA(NSEG+1) (matvct.f:685-692)
    ldxa = nelem
  end if
  ! prisms: the nodes of one vertical column share (x, y), so the
  ! coordinate order keeps columns contiguous without the plane count
  have_xy = associated(mesh%x) .and. associated(mesh%y)
  if (have_xy) have_xy = associated(mesh%x%r) .and. associated(mesh%y%r)
  if (have_xy) then
    have_xy = size(mesh%x%r) >= npoin .and. size(mesh%y%r) >= npoin
  end if
  if (a%sto == 3) then
    if (have_xy) then
      h = bhip_mesh_ctx(npoin, nelem, nelmax, ndp, mesh%gloseg%i, mesh%x%r, mesh%y%r, .true., 1)
    else
      h = bhip_mesh_ctx(npoin, nelem, nelmax, ndp, mesh%gloseg%i, dummy, dummy, .false., 1)
    end if
  else if (have_xy) then
    h = bhip_mesh_ctx(npoin, nelem, nelmax, ndp, mesh%ikle%i, mesh%x%r, mesh%y%r, .true., 1)
  else
    h = bhip_mesh_ctx(npoin, nelem, nelmax, ndp, mesh%ikle%i, dummy, dummy, .false., 1)
  end if
  ncoef = bhip_ncoef(ndp, a%typext)
  allocate(xa(nelem, max(ncoef, 1)))
  if (ncoef > 0) call bhip_pack_xa(a%x%r, ldxa, nelem, ncoef, xa)
  rc = bhip_matrix_set(h, a%typdia, a%typext, c_loc(a%d%r(1)), c_loc(xa(1, 1)))
  deallocate(xa)
  if (rc /= BHIP_OK) call bhip_fail('BHIP: MATRIX UPLOAD FAILED', rc)
end subroutine
end module bief_hip_load


! ---------------------------------------------------------------------------
SUBROUTINE SOLVE(X, A, B, TB, CFG, INFOGR, MESH, AUX)
#ifdef BHIP_OWN_TYPES
  USE BIEF_HIP_TYPES
#else
  USE BIEF_DEF
  USE DECLARATIONS_SPECIAL
#endif
  USE BIEF_HIP_STATE
  USE BIEF_HIP_LOAD
  IMPLICIT NONE
  TYPE(SLVCFG), INTENT(INOUT)    :: CFG
  LOGICAL, INTENT(IN)            :: INFOGR
  TYPE(BIEF_OBJ), TARGET, INTENT(INOUT) :: X, B
  TYPE(BIEF_OBJ), INTENT(INOUT)  :: TB
  TYPE(BIEF_OBJ), INTENT(INOUT), TARGET :: A, AUX
  TYPE(BIEF_MESH), INTENT(INOUT), TARGET :: MESH
  TYPE(BIEF_OBJ), POINTER :: PA, PX, PB
  TYPE(BHIP_SOLVER_CFG) :: SC
  TYPE(BHIP_SOLVE_INFO) :: INFO
  TYPE(C_PTR) :: H
  INTEGER :: NDP, NPOIN, NELEM, NCOEF, RC, UNIT, LDXA
  CHARACTER(KIND=C_CHAR) :: TD
  CHARACTER(LEN=6) :: SNAME
  DOUBLE PRECISION, ALLOCATABLE, TARGET :: XA(:,:)
  DOUBLE PRECISION, TARGET :: DUMMY(1)
#ifdef BHIP_OWN_TYPES
  UNIT = 6
#else
  UNIT = LU
#endif
  ! single matrix (S = 0) or a block of one matrix (S = 1), solve.f:222-240
  IF (A%TYPE == 3) THEN
    PA => A
    PX => X
    PB => B
  ELSE IF (A%TYPE == 4 .AND. A%N == 1) THEN
    PA => A%ADR(1)%P
    PX => X%ADR(1)%P
    PB => B%ADR(1)%P
  ELSE
    CALL BHIP_FAIL('BHIP SOLVE: BLOCKS OF 4 OR 9 MATRICES NOT SUPPORTED', BHIP_EBADARG)
  END IF
  IF (CFG%SLV /= 1 .AND. CFG%SLV /= 6 .AND. CFG%SLV /= 7) THEN
    CALL BHIP_FAIL('BHIP SOLVE: SOLVER NOT AVAILABLE (1 CG, 6 CGSTAB, 7 GMRES)', BHIP_EBADARG)
  END IF
  CALL BHIP_LOAD_MATRIX(PA, MESH, H, NDP, NPOIN, NELEM, LDXA)
  SC%SLV = CFG%SLV
  SC%PRECON = CFG%PRECON
  SC%KRYLOV = CFG%KRYLOV
  SC%NITMAX = CFG%NITMAX
  SC%EPS = CFG%EPS
  SC%PRECISION = BHIP_PREC_FP64
  IF (CFG%SLV == 1 .OR. CFG%SLV == 7) SC%PRECISION = BHIP_PRECISION_MODE()
  SC%INNER_EPS = BHIP_INNER_EPS
  SC%INNER_NITMAX = 0
  RC = BHIP_SOLVE(H, SC, C_LOC(PX%R(1)), C_LOC(PB%R(1)), INFO)
  IF (RC /= BHIP_OK .AND. RC /= BHIP_MAXITER .AND. RC /= BHIP_ZERO_RHS) THEN
    CALL BHIP_FAIL('BHIP SOLVE: FAILED', RC)
  END IF
  ! A is left as the reference leaves it (preconditioned, precd1.f:136-142)
  NCOEF = BHIP_NCOEF(NDP, PA%TYPEXT)
  ALLOCATE(XA(NELEM, MAX(NCOEF, 1)))
  IF (NCOEF > 0) THEN
    RC = BHIP_MATRIX_GET(H, TD, C_LOC(PA%D%R(1)), C_LOC(XA(1, 1)))
    IF (RC /= BHIP_OK) CALL BHIP_FAIL('BHIP SOLVE: MATRIX READ-BACK FAILED', RC)
    CALL BHIP_UNPACK_XA(XA, NELEM, LDXA, NCOEF, PA%X%R)
  ELSE
    RC = BHIP_MATRIX_GET(H, TD, C_LOC(PA%D%R(1)), C_LOC(DUMMY(1)))
    IF (RC /= BHIP_OK) CALL BHIP_FAIL('BHIP SOLVE: MATRIX READ-BACK FAILED', RC)
  END IF
  DEALLOCATE(XA)
  PA%TYPDIA = TD
  ! listing lines in the reference's format (gracjg.F:426-435)
  SELECT CASE (CFG%SLV)
  CASE (1)
    SNAME = 'GRACJG'
  CASE (6)
    SNAME = 'CGSTAB'
  CASE DEFAULT
    SNAME = 'GMRES '
  END SELECT
  IF (RC == BHIP_ZERO_RHS) THEN
    IF (INFOGR) WRITE(UNIT, '(1X,A,A,G16.7)') TRIM(SNAME), &
         ' (BIEF) :  SOLUTION X=0 BECAUSE L2-NORM OF B VERY SMALL:', 0.D0
  ELSE IF (INFO%CONVERGED == 0) THEN
    IF (INFO%RELAT == 1) THEN
      WRITE(UNIT, '(1X,A,A,1I8,A,G16.7)') TRIM(SNAME), ' (BIEF) : EXCEEDING MAXIMUM ITERATIONS:', &
           INFO%NITER, ' RELATIVE PRECISION:', INFO%RELRES
    ELSE
      WRITE(UNIT, '(1X,A,A,1I8,A,G16.7)') TRIM(SNAME), ' (BIEF) : EXCEEDING MAXIMUM ITERATIONS:', &
           INFO%NITER, ' ABSOLUTE PRECISION:', INFO%RELRES
    END IF
  ELSE IF (INFOGR) THEN
    IF (INFO%RELAT == 1) THEN
      WRITE(UNIT, '(1X,A,A,1I8,A,G16.7)') TRIM(SNAME), ' (BIEF) : ', INFO%NITER, &
           ' ITERATIONS, RELATIVE PRECISION:', INFO%RELRES
    ELSE
      WRITE(UNIT, '(1X,A,A,1I8,A,G16.7)') TRIM(SNAME), ' (BIEF) : ', INFO%NITER, &
           ' ITERATIONS, ABSOLUTE PRECISION:', INFO%RELRES
    END IF
  END IF
END SUBROUTINE SOLVE


! ---------------------------------------------------------------------------
SUBROUTINE MATVEC(OP, X, A, Y, C, MESH, LEGO)
#ifdef BHIP_OWN_TYPES
  USE BIEF_HIP_TYPES
#else
  USE BIEF_DEF
#endif
  USE BIEF_HIP_STATE
  USE BIEF_HIP_LOAD
  IMPLICIT NONE
  CHARACTER(LEN=8), INTENT(IN)           :: OP
  TYPE(BIEF_OBJ), INTENT(INOUT), TARGET  :: X
  TYPE(BIEF_OBJ), INTENT(IN), TARGET     :: A, Y
  DOUBLE PRECISION, INTENT(IN)           :: C
  TYPE(BIEF_MESH), INTENT(INOUT), TARGET :: MESH
  LOGICAL, INTENT(IN), OPTIONAL          :: LEGO
  TYPE(C_PTR) :: H
  INTEGER :: NDP, NPOIN, NELEM, CODE, RC, LDXA
  IF (PRESENT(LEGO)) THEN
    ! the non-assembled W hand-over between calls (matvec.f:279-292)
    IF (.NOT. LEGO) CALL BHIP_FAIL('BHIP MATVEC: LEGO=.FALSE. NOT SUPPORTED', BHIP_EBADARG)
  END IF
  CODE = BHIP_OPCODE(OP)
  IF (CODE < 0) CALL BHIP_FAIL('BHIP MATVEC: UNKNOWN OPERATION '//OP, BHIP_EBADARG)
  CALL BHIP_LOAD_MATRIX(A, MESH, H, NDP, NPOIN, NELEM, LDXA)
  IF (Y%DIM1 < NPOIN .OR. SIZE(X%R) < NPOIN) THEN
    CALL BHIP_FAIL('BHIP MATVEC: VECTOR SIZES', BHIP_EBADARG)
  END IF
  X%ELM = A%ELMLIN
  X%DIM1 = NPOIN
  RC = BHIP_MATVEC(H, INT(CODE, C_INT), REAL(C, C_DOUBLE), C_LOC(Y%R(1)), C_LOC(X%R(1)))
  IF (RC /= BHIP_OK) CALL BHIP_FAIL('BHIP MATVEC: FAILED', RC)
END SUBROUTINE MATVEC


! ---------------------------------------------------------------------------
SUBROUTINE MATRBL(OP, X, A, Y, C, MESH)
#ifdef BHIP_OWN_TYPES
  USE BIEF_HIP_TYPES
#else
  USE BIEF_DEF
#endif
  USE BIEF_HIP_STATE
  IMPLICIT NONE
  CHARACTER(LEN=8), INTENT(IN)   :: OP
  TYPE(BIEF_OBJ), INTENT(INOUT)  :: X
  TYPE(BIEF_OBJ), INTENT(IN)     :: A, Y
  DOUBLE PRECISION, INTENT(IN)   :: C
  TYPE(BIEF_MESH), INTENT(INOUT) :: MESH
  INTERFACE
    SUBROUTINE MATVEC(OP, X, A, Y, C, MESH, LEGO)
#ifdef BHIP_OWN_TYPES
      USE BIEF_HIP_TYPES
#else
      USE BIEF_DEF
#endif
      CHARACTER(LEN=8), INTENT(IN)           :: OP
      TYPE(BIEF_OBJ), INTENT(INOUT), TARGET  :: X
      TYPE(BIEF_OBJ), INTENT(IN), TARGET     :: A, Y
      DOUBLE PRECISION, INTENT(IN)           :: C
      TYPE(BIEF_MESH), INTENT(INOUT), TARGET :: MESH
      LOGICAL, INTENT(IN), OPTIONAL          :: LEGO
    END SUBROUTINE
  END INTERFACE
  ! single matrix or a block of one (matrbl.f:194-202); one domain, so no
  ! PARCOM (matrbl.f:235-241)
  IF (A%TYPE == 3) THEN
    CALL MATVEC(OP, X, A, Y, C, MESH)
  ELSE IF (A%TYPE == 4 .AND. A%N == 1) THEN
    CALL MATVEC(OP, X%ADR(1)%P, A%ADR(1)%P, Y%ADR(1)%P, C, MESH)
  ELSE
    CALL BHIP_FAIL('BHIP MATRBL: BLOCKS OF 4 OR 9 MATRICES NOT SUPPORTED', BHIP_EBADARG)
  END IF
END SUBROUTINE MATRBL
