! Stand-ins for the generated network plugin's procedures
! (src/odevec_commons.F90:32 GetJac, :55 GetLU, :78 GetRHS): hosts pass them
! to SolveODE as external actual arguments, so they must resolve at link
! time. The network's RHS, Jacobian and LU run inside libodevec on the GPU;
! these are never called by SolveODE and stop if a host calls them.
module odevec_commons
  implicit none
contains
  subroutine GetRHS(this, y, rhs)
    class(*) :: this
    double precision :: y(:,:), rhs(:,:)
    error stop "odevec_commons: GetRHS runs on the device inside libodevec"
  end subroutine
  subroutine GetJac(this, beta, y, dt, jac)
    class(*) :: this
    double precision :: beta, y(:,:), dt
    double precision :: jac(*)
    error stop "odevec_commons: GetJac runs on the device inside libodevec"
  end subroutine
  subroutine GetLU(this, beta, y, dt, LU)
    class(*) :: this
    double precision :: beta, y(:,:), dt
    double precision :: LU(*)
    error stop "odevec_commons: GetLU runs on the device inside libodevec"
  end subroutine
end module odevec_commons
