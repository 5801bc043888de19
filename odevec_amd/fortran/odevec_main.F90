! Fortran host interface of libodevec (C ABI: include/odevec.h).
!
! Same module name, derived type, component names and procedures as the
! reference's solver module (src/odevec.F90:33-74 type(odevec), :94
! InitOdevec, :158 SolveODE, :1058 CloseOdevec), so a host written against
! the reference (tests/rober.f90, tests/orego.f90, tests/primordial.f90,
! Fosite) compiles unchanged against this module plus odevec_commons.
!
! As in the reference, the network and nvector are fixed when the module is
! built (the reference's preprocessor writes them into pragmas; here they
! are cpp macros):
!   -DODEVEC_NETWORK='"ROBER/NONE/DENSE"'  generated variant in libodevec.so
!   -DODEVEC_NVECTOR=32  -DODEVEC_NEQ=3      vector length and species
!   -DODEVEC_PRECISION=0|1  -DODEVEC_MODE=0|1|2   (FP64|mixed, percell|lockstep|persistent)
! ODEVEC_MODE 1 (lockstep over the whole vector) reproduces the reference's
! shared step control; 0 gives every cell its own step control.
!
! Errors: the reference prints and STOPs on a failed solve
! (src/odevec.F90:380-389); so does this module (error stop with the
! library's message), after copying the per-cell status to this%status.
#ifndef ODEVEC_NETWORK
#define ODEVEC_NETWORK "ROBER/NONE/DENSE"
#endif
#ifndef ODEVEC_NVECTOR
#define ODEVEC_NVECTOR 32
#endif
#ifndef ODEVEC_NEQ
#define ODEVEC_NEQ 3
#endif
#ifndef ODEVEC_PRECISION
#define ODEVEC_PRECISION 0
#endif
#ifndef ODEVEC_MODE
#define ODEVEC_MODE 1
#endif
module odevec_main
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: odevec, InitOdevec, SolveODE, CloseOdevec, OdevecNetworkCount

  integer, parameter :: dp = kind(1.0d0)

  type, bind(C) :: odevec_opts_t
    integer(c_int32_t) :: precision, mode, group, block, grid, max_attempts, device_ptrs, abort_on_error, schedule
  end type

  type :: odevec
    integer :: nvector = ODEVEC_NVECTOR
    integer :: neq = ODEVEC_NEQ
    integer :: maxorder = 5
    real(dp) :: RelativeTolerance = 1e-6_dp
    real(dp) :: AbsoluteTolerance = 1e-10_dp
    real(dp) :: MinimalTimestep = real(1e-10, dp)
    logical :: CheckNegatives = .false.
    ! mixed form (ODEVEC_PRECISION=1): cells reaching this many convergence
    ! failures are solved again in FP64 (0: off); no reference counterpart
    integer :: FP64Fallback = 0
    ! mixed form: a corrector that leaves an element or charge invariant off
    ! by more than this many tolerances has not converged (0: off)
    real(dp) :: InvariantTolerance = 1.0_dp
    ! ... cells failing it this many times go to FP64 (FP64Fallback on) or are no
    ! longer tested (0: neither)
    integer :: InvariantFallbackFails = 32
    ! mixed form's corrector test: 0 every weighted component (max norm), 1 the
    ! RMS norm with the rate discount of CheckConvergence (as in FP64)
    integer :: MixedConvergenceTest = 0
    real(dp), allocatable :: y(:,:)
    integer(c_int32_t), allocatable :: status(:)
    type(c_ptr) :: handle = c_null_ptr
  end type

  interface
    function c_create(network, nvector, opts, out) bind(C, name='odevec_create') result(rc)
      import :: c_char, c_int64_t, c_int32_t, c_ptr, odevec_opts_t
      character(kind=c_char), dimension(*), intent(in) :: network
      integer(c_int64_t), value :: nvector
      type(odevec_opts_t), intent(in) :: opts
      type(c_ptr), intent(out) :: out
      integer(c_int32_t) :: rc
    end function
    subroutine c_default_opts(opts) bind(C, name='odevec_default_opts')
      import :: odevec_opts_t
      type(odevec_opts_t), intent(out) :: opts
    end subroutine
    function c_set_tolerances(h, rtol, atol) bind(C, name='odevec_set_tolerances') result(rc)
      import :: c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), value :: rtol, atol
      integer(c_int32_t) :: rc
    end function
    function c_set_dt_min(h, dt_min) bind(C, name='odevec_set_dt_min') result(rc)
      import :: c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), value :: dt_min
      integer(c_int32_t) :: rc
    end function
    function c_set_check_negatives(h, flag) bind(C, name='odevec_set_check_negatives') result(rc)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: flag
      integer(c_int32_t) :: rc
    end function
    function c_set_fp64_fallback(h, limit) bind(C, name='odevec_set_fp64_fallback') result(rc)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: limit
      integer(c_int32_t) :: rc
    end function
    function c_set_invariant_test(h, tol, fails) bind(C, name='odevec_set_invariant_test') result(rc)
      import :: c_ptr, c_int32_t, c_double
      type(c_ptr), value :: h
      real(c_double), value :: tol
      integer(c_int32_t), value :: fails
      integer(c_int32_t) :: rc
    end function
    function c_set_mixed_test(h, test) bind(C, name='odevec_set_mixed_convergence_test') result(rc)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t), value :: test
      integer(c_int32_t) :: rc
    end function
    function c_solve(h, time, dt, t_stop, y, mask, status) bind(C, name='odevec_solve') result(rc)
      import :: c_ptr, c_double, c_int32_t
      type(c_ptr), value :: h
      real(c_double), intent(inout) :: time, dt
      real(c_double), value :: t_stop
      real(c_double), dimension(*), intent(inout) :: y
      type(c_ptr), value :: mask
      integer(c_int32_t), dimension(*), intent(out) :: status
      integer(c_int32_t) :: rc
    end function
    function c_destroy(h) bind(C, name='odevec_destroy') result(rc)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: h
      integer(c_int32_t) :: rc
    end function
    function c_last_error() bind(C, name='odevec_last_error') result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function c_network_count() bind(C, name='odevec_network_count') result(n)
      import :: c_int32_t
      integer(c_int32_t) :: n
    end function
    function c_strlen(s) bind(C, name='strlen') result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
  end interface

contains

  !> Number of generated network variants in the library (no device needed).
  integer function OdevecNetworkCount()
    OdevecNetworkCount = int(c_network_count())
  end function

  subroutine fail(rc, what)
    integer(c_int32_t), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    p = c_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    print *, "odevec: ", what, " failed (", rc, "): ", msg(1:n)
    error stop 1
  end subroutine

  !> Replaces InitOdevec (src/odevec.F90:94-154).
  subroutine InitOdevec(this)
    type(odevec), intent(inout) :: this
    type(odevec_opts_t) :: opts
    integer(c_int32_t) :: rc
    call c_default_opts(opts)
    opts%precision = ODEVEC_PRECISION
    opts%mode = ODEVEC_MODE
    opts%group = min(this%nvector, 256)
    rc = c_create(ODEVEC_NETWORK // c_null_char, int(this%nvector, c_int64_t), opts, this%handle)
    if (rc /= 0) call fail(rc, "InitOdevec")
    if (.not. allocated(this%y)) allocate(this%y(this%nvector, this%neq))
    if (.not. allocated(this%status)) allocate(this%status(this%nvector))
    this%y = 0.0_dp
    this%status = 0
  end subroutine

  !> Replaces SolveODE(this,time,dt,t_stop,y,GetRHS,GetJac,GetLU,Mask)
  !! (src/odevec.F90:158-214). GetRHS/GetJac/GetLU are accepted for source
  !! compatibility; the network's functions are compiled into the library.
  subroutine SolveODE(this, time, dt, t_stop, y, GetRHS, GetJac, GetLU, Mask)
    type(odevec), intent(inout) :: this
    real(dp), intent(inout) :: time, dt
    real(dp), intent(in) :: t_stop
    real(dp), intent(inout) :: y(this%nvector, this%neq)
    external :: GetRHS, GetJac, GetLU
    logical, intent(in), optional :: Mask(this%nvector)
    integer(c_int8_t), allocatable, target :: m8(:)
    integer(c_int32_t) :: rc
    type(c_ptr) :: mp
    rc = c_set_tolerances(this%handle, this%RelativeTolerance, this%AbsoluteTolerance)
    if (rc == 0) rc = c_set_dt_min(this%handle, this%MinimalTimestep)
    if (rc == 0) rc = c_set_check_negatives(this%handle, merge(1_c_int32_t, 0_c_int32_t, this%CheckNegatives))
    if (rc == 0) rc = c_set_fp64_fallback(this%handle, int(this%FP64Fallback, c_int32_t))
    if (rc == 0) rc = c_set_invariant_test(this%handle, real(this%InvariantTolerance, c_double), &
                                           int(this%InvariantFallbackFails, c_int32_t))
    if (rc == 0) rc = c_set_mixed_test(this%handle, int(this%MixedConvergenceTest, c_int32_t))
    if (rc /= 0) call fail(rc, "SolveODE (settings)")
    mp = c_null_ptr
    if (present(Mask)) then
      allocate(m8(this%nvector))
      m8 = merge(1_c_int8_t, 0_c_int8_t, Mask)
      mp = c_loc(m8)
    end if
    rc = c_solve(this%handle, time, dt, t_stop, y, mp, this%status)
    if (rc /= 0) call fail(rc, "SolveODE")
  end subroutine

  !> Replaces CloseOdevec (src/odevec.F90:1058-1074).
  subroutine CloseOdevec(this)
    type(odevec), intent(inout) :: this
    integer(c_int32_t) :: rc
    if (c_associated(this%handle)) then
      rc = c_destroy(this%handle)
      this%handle = c_null_ptr
    end if
    if (allocated(this%y)) deallocate(this%y)
    if (allocated(this%status)) deallocate(this%status)
  end subroutine

end module odevec_main
