! The generated network plugin's procedures (src/odevec_commons.F90:32
! GetJac, :55 GetLU, :78 GetRHS) for the library's network. Hosts pass them
! to SolveODE as external actual arguments (SolveODE does not call them:
! the network's functions run inside libodevec); a host may also call
! GetRHS and GetJac itself. They evaluate on the GPU through the C ABI
! (odevec_get_rhs, odevec_get_jac), with the network's default parameters,
! in the reference's layouts: rhs(nvector, neq); jac the packaging's
! storage, DENSE jac(nvector, neq, neq) = P_order, CSC jac(nvector, nnz)
! over the LU pattern. GetLU (LUmethod 2, closed-form LU) is not generated
! and stops if called.
module odevec_commons
  use, intrinsic :: iso_c_binding
  use odevec_main
  implicit none
  private
  public :: GetRHS, GetJac, GetLU

  interface
    function c_get_rhs(h, ncells, y, params, rhs) bind(C, name='odevec_get_rhs') result(rc)
      import :: c_ptr, c_int64_t, c_double, c_int32_t
      type(c_ptr), value :: h
      integer(c_int64_t), value :: ncells
      real(c_double), dimension(*), intent(in) :: y
      type(c_ptr), value :: params
      real(c_double), dimension(*), intent(out) :: rhs
      integer(c_int32_t) :: rc
    end function
    function c_get_jac(h, ncells, beta, dt, y, params, jac) bind(C, name='odevec_get_jac') result(rc)
      import :: c_ptr, c_int64_t, c_double, c_int32_t
      type(c_ptr), value :: h
      integer(c_int64_t), value :: ncells
      real(c_double), value :: beta, dt
      real(c_double), dimension(*), intent(in) :: y
      type(c_ptr), value :: params
      real(c_double), dimension(*), intent(out) :: jac
      integer(c_int32_t) :: rc
    end function
  end interface

contains

  !> f(y) of every cell (src/odevec_commons.F90:78-95)
  subroutine GetRHS(this, y, rhs)
    type(odevec) :: this
    double precision, dimension(this%nvector, this%neq) :: y, rhs
    integer(c_int32_t) :: rc
    if (.not. c_associated(this%handle)) error stop "odevec_commons: GetRHS before InitOdevec"
    rc = c_get_rhs(this%handle, int(this%nvector, c_int64_t), y, c_null_ptr, rhs)
    if (rc /= 0) error stop "odevec_commons: GetRHS failed"
  end subroutine

  !> P = 1 - dt*beta*J, permuted, in the packaging's storage (src/odevec_commons.F90:32-51)
  subroutine GetJac(this, beta, y, dt, jac)
    type(odevec) :: this
    double precision :: beta, dt
    double precision, dimension(this%nvector, this%neq) :: y
    double precision :: jac(*)
    integer(c_int32_t) :: rc
    if (.not. c_associated(this%handle)) error stop "odevec_commons: GetJac before InitOdevec"
    rc = c_get_jac(this%handle, int(this%nvector, c_int64_t), beta, dt, y, c_null_ptr, jac)
    if (rc /= 0) error stop "odevec_commons: GetJac failed"
  end subroutine

  !> LUmethod 2's closed-form LU (src/odevec_commons.F90:55-75) is not generated
  subroutine GetLU(this, beta, y, dt, LU)
    type(odevec) :: this
    double precision :: beta, dt
    double precision, dimension(this%nvector, this%neq) :: y
    double precision :: LU(*)
    error stop "odevec_commons: GetLU (LUmethod 2) is not generated"
  end subroutine
end module odevec_commons
