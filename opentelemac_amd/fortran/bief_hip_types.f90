! Stand-alone subset of BIEF's data model (sources/utils/bief/bief_def.f:
! BIEF_OBJ 74-196, BIEF_MESH 205-498, SLVCFG 506-532), with the same type
! and component names, for building the host layer without a TELEMAC tree.
! Inside a TELEMAC build the shim uses the real BIEF_DEF instead (the
! preprocessor switch BHIP_OWN_TYPES selects this module).
module bief_hip_types
  implicit none

  type pointer_to_bief_obj
    type(bief_obj), pointer :: p
  end type

  type bief_obj
    integer :: key = 123456
    integer :: type = 0            ! 2 vector, 3 matrix, 4 block
    character(len=6) :: name = ' '
    integer :: elm = 0             ! 11 P1 triangle, 41 P1 prism
    integer :: dim1 = 0, maxdim1 = 0, dim2 = 1, maxdim2 = 1
    double precision, pointer, dimension(:) :: r => null()
    integer, pointer, dimension(:) :: i => null()
    integer :: sto = 1             ! 1: element by element, 3: segments
    integer :: elmlin = 0, elmcol = 0
    character(len=1) :: typdia = '0', typext = '0'
    type(bief_obj), pointer :: d => null(), x => null()
    integer :: n = 0
    type(pointer_to_bief_obj), pointer, dimension(:) :: adr => null()
  end type

  type bief_mesh
    character(len=6) :: name = 'MESH  '
    integer, pointer :: nelem => null(), nelmax => null(), npoin => null()
    integer, pointer :: dim1 => null(), typelm => null(), nseg => null()
    type(bief_obj), pointer :: ikle => null(), x => null(), y => null(), gloseg => null()
  end type

  type slvcfg
    integer :: slv = 1, nitmax = 100, precon = 2, krylov = 5
    double precision :: eps = 1.d-8, zero = 1.d-10
    logical :: ok = .true.
    integer :: nit = 0
  end type
end module bief_hip_types
