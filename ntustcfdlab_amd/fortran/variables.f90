! variables.f90 -- state of the Fortran driver.
!
! The names, shapes and index ranges are those of the solver's module
! (variables_module.f90:9-23, 110-132, 159, 201-233, 277-283, 306-316,
! 331-343), so the hot-path shims in hotpath.f90 compile against either
! this module or the solver's own.  Sizes are run-time values here
! (allocatable arrays) instead of compile-time parameters.
module variables
   implicit none

   integer :: nx = 40, ny = 40, nz = 40
   real*8 :: lx = 1.0, ly = 1.0, lz = 1.0

   ! physics and solver settings; defaults as the solver sets them
   ! (single-precision literals widened, variables_module.f90:9-11,277;
   ! main.f90:44-60)
   real*8 :: Re = 10.0, dt = 1.0e-3, nu, Cs = 0.18
   real*8 :: omega = 1.5, zeta = 1.e-4, itmax = 3000
   integer :: LES = 0
   integer :: istart = 1, iend = 40, ik, istep, nstep
   real*8 :: pChangeMax_, cDrag, cLift, time = 0.0
   real*8 :: u_solid = 0, v_solid = 0, w_solid = 0

   character(len=20) :: WestWall_u = 'no-slip', WestWall_v = 'no-slip', WestWall_w = 'no-slip'
   character(len=20) :: EastWall_u = 'no-slip', EastWall_v = 'no-slip', EastWall_w = 'no-slip'
   character(len=20) :: SouthWall_u = 'no-slip', SouthWall_v = 'no-slip', SouthWall_w = 'no-slip'
   character(len=20) :: NorthWall_u = 'no-slip', NorthWall_v = 'no-slip', NorthWall_w = 'Dirichlet'
   character(len=20) :: BackWall_u = 'no-slip', BackWall_v = 'no-slip', BackWall_w = 'no-slip'
   character(len=20) :: FrontWall_u = 'no-slip', FrontWall_v = 'no-slip', FrontWall_w = 'no-slip'

   ! halo arrays (-1:n+2 on every axis)
   real*8, allocatable, dimension(:, :, :) :: u, v, w, p, u1, v1, w1, u2, v2, w2
   real*8, allocatable, dimension(:, :, :) :: u_star, v_star, w_star, FX, FY, FZ, ETA
   ! interior arrays (1:n)
   real*8, allocatable, dimension(:, :, :) :: uc, vc, wc, pre, Viseff
   real*8, allocatable, dimension(:, :, :, :) :: dudx, dvdx, dwdx
   ! grid: nodes (-1:n+3) and metrics (-1:n+2)
   real*8, allocatable, dimension(:) :: X, Y, Z, iDx, Dxs, iDy, Dys, iDz, Dzs

   ! airfoil (variables_module.f90:142-179): placement, pitching, polygon
   real*8, parameter :: xc = 1.0, yc = 4.0, zc = 4.0, reduce_frequency = 2.0
   real*8 :: AOA = 0, AOA1 = 5.D0, frequency, StartDynamic = 1000000
   integer :: poly = 5000
   integer, parameter :: nSubGrids = 50
   character(len=512) :: NACA_filename = 'NACA0012.DAT'
   real*8, allocatable, dimension(:) :: az, ay

   ! output (variables_module.f90:201-205,223,355-365): solution array,
   ! file counter, steady check
   integer, parameter :: nblocks = 1
   real, allocatable, dimension(:, :, :, :) :: Qout
   real :: temp = 1.0
   integer :: num = 0, isto = 100
   character(len=20) :: filename
   real*8 :: VelocityDifference
   real*8, allocatable, dimension(:, :, :) :: last_velocity

contains

   subroutine allocate_fields()
      allocate (u(-1:nx + 2, -1:ny + 2, -1:nz + 2), source=0.0d0)
      allocate (v, w, p, u1, v1, w1, u2, v2, w2, u_star, v_star, w_star, FX, FY, FZ, ETA, &
                mold=u)
      v = 0; w = 0; p = 0; u1 = 0; v1 = 0; w1 = 0; u2 = 0; v2 = 0; w2 = 0
      u_star = 0; v_star = 0; w_star = 0; FX = 0; FY = 0; FZ = 0; ETA = 0
      allocate (uc(nx, ny, nz), vc(nx, ny, nz), wc(nx, ny, nz), pre(nx, ny, nz), &
                Viseff(nx, ny, nz), source=0.0d0)
      allocate (dudx(nx, ny, nz, 3), dvdx(nx, ny, nz, 3), dwdx(nx, ny, nz, 3), source=0.0d0)
      allocate (Qout(nx, ny, nz, 5), source=0.0)
      allocate (last_velocity(-1:nx + 2, -1:ny + 2, -1:nz + 2), source=0.0d0)
      allocate (X(-1:nx + 3), Y(-1:ny + 3), Z(-1:nz + 3))
      allocate (iDx(-1:nx + 2), Dxs(-1:nx + 2), iDy(-1:ny + 2), Dys(-1:ny + 2), &
                iDz(-1:nz + 2), Dzs(-1:nz + 2))
      istart = 1
      iend = nz
   end subroutine allocate_fields

end module variables
