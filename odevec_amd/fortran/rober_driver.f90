! ROBER through the Fortran module, in the reference driver's calling
! pattern (tests/rober.f90:37-73): InitOdevec, set the tolerance fields, fill
! ode%y, repeated SolveODE calls to t_stop = 1e-10, 1e-9, ..., 1e10 with the
! dt hint reset to 1e-10 before each call, CloseOdevec. Prints t_stop and
! every cell's state after each call, full precision, for comparison with
! the Python binding on the same calls; then the plugin's GetRHS and GetJac
! at the final state (lines "RHS i f1 f2 f3", "JAC i P11 P21 ... P33").

program rober_driver
  use odevec_main
  use odevec_commons
  implicit none
  type(odevec) :: ode
  double precision :: t_start, t_stop, dt
  double precision, allocatable :: rhs(:, :), jac(:, :, :)
  integer :: i, j, k

  call InitOdevec(ode)
  ode%RelativeTolerance = 1d-6
  ode%AbsoluteTolerance = 1d-10
  ode%MinimalTimestep = 1d-20
  do i = 1, ode%nvector
    ode%y(i, :) = (/ 1d0, 0d0, 0d0 /)
  end do
  t_start = 0d0
  t_stop = 1d-10
  do while (t_stop < 1d11)
    dt = 1d-10
    call SolveODE(ode, t_start, dt, t_stop, ode%y, GetRHS, GetJac, GetLU)
    do i = 1, ode%nvector
      write (*, '(es25.17e3,i6,3es25.17e3,i3)') t_stop, i, (ode%y(i, j), j = 1, ode%neq), ode%status(i)
    end do
    t_stop = t_stop * 1d1
  end do
  allocate(rhs(ode%nvector, ode%neq), jac(ode%nvector, ode%neq, ode%neq))
  call GetRHS(ode, ode%y, rhs)
  call GetJac(ode, 0.5d0, ode%y, 1d-3, jac)
  do i = 1, ode%nvector
    write (*, '(a,i6,3es26.17e3)') 'RHS', i, (rhs(i, j), j = 1, ode%neq)
    write (*, '(a,i6,9es26.17e3)') 'JAC', i, ((jac(i, j, k), j = 1, ode%neq), k = 1, ode%neq)
  end do
  call CloseOdevec(ode)
end program rober_driver
