! hotpath.f90 -- drop-in replacements for the solver's hot-path subroutines.
!
! Each subroutine keeps the name and the empty argument list of the
! reference subroutine it replaces (GradientPhiGauss keeps its five
! arguments), works on the same module variables, and runs on the GPU
! through libntust_hip.so.  Linking this file instead of smagorinsky.f90,
! discretisation_QUICK_centre.f90 / discretisation_upwind.f90, GS.f90,
! calcul_new_velocity.f90, virtualForceIntegrator.f90 and
! boundary_conditions.f90 leaves the caller (main.f90:139-259) unchanged.
! ntust_setup must run once after the gridder (main.f90:82-86) and
! ntust_teardown before MPI_FINALIZE (main.f90:362).

subroutine ntust_setup()
   use variables
   use ntust_hip
   implicit none
   call ntust_check(ntust_init(nx, ny, nz, iDx, Dxs, iDy, Dys, iDz, Dzs), 'ntust_init')
   call ntust_push_bc()
end subroutine ntust_setup

subroutine ntust_teardown()
   use ntust_hip
   implicit none
   call ntust_check(ntust_finalize(), 'ntust_finalize')
end subroutine ntust_teardown

! BC strings of variables_module.f90:110-132 -> library codes
subroutine ntust_push_bc()
   use variables
   use ntust_hip
   implicit none
   integer(c_int) :: codes(18)
   codes = [code(WestWall_u), code(WestWall_v), code(WestWall_w), &
            code(EastWall_u), code(EastWall_v), code(EastWall_w), &
            code(SouthWall_u), code(SouthWall_v), code(SouthWall_w), &
            code(NorthWall_u), code(NorthWall_v), code(NorthWall_w), &
            code(BackWall_u), code(BackWall_v), code(BackWall_w), &
            code(FrontWall_u), code(FrontWall_v), code(FrontWall_w)]
   call ntust_check(ntust_set_bc(codes), 'ntust_set_bc')
contains
   integer(c_int) function code(s)
      character(len=*), intent(in) :: s
      if (s == 'no-slip') then
         code = 0
      else if (s == 'Neumann') then
         code = 1
      else if (s == 'Dirichlet') then
         code = 2
      else
         code = 99
      end if
   end function code
end subroutine ntust_push_bc

! smagorinsky.f90:42-87
subroutine GradientPhiGauss(Phi, dPhidX, Dxm, Dym, Dzm)
   use variables, only: istart, iend
   use ntust_hip
   implicit none
   real*8 :: Phi(*), dPhidX(*), Dxm(*), Dym(*), Dzm(*)
   call ntust_check(ntust_gradient_phi_gauss(Phi, c_null_ptr, dPhidX, Dxm, Dym, Dzm, &
                                             istart, iend), 'GradientPhiGauss')
end subroutine GradientPhiGauss

! smagorinsky.f90:1-37 (the three gradients of :6-8 fused in)
subroutine CalculateSmagorinskyViscosity()
   use variables
   use ntust_hip
   implicit none
   call ntust_check(ntust_smagorinsky(u, v, w, c_null_ptr, c_null_ptr, c_null_ptr, Cs, &
                                      istart, iend, dudx, dvdx, dwdx, Viseff), &
                    'CalculateSmagorinskyViscosity')
end subroutine CalculateSmagorinskyViscosity

! discretisation_QUICK_centre.f90:1-345
subroutine discretisation_QUICK_centre()
   use variables
   use ntust_hip
   implicit none
   call ntust_check(ntust_discretisation_quick_centre(u, v, w, Viseff, dt, nu, LES, istart, &
                                                      iend, u_star, v_star, w_star), &
                    'discretisation_QUICK_centre')
end subroutine discretisation_QUICK_centre

! discretisation_upwind.f90:1-275
subroutine discretisation_upwind()
   use variables
   use ntust_hip
   implicit none
   call ntust_check(ntust_discretisation_upwind(u, v, w, dt, nu, u_star, v_star, w_star), &
                    'discretisation_upwind')
end subroutine discretisation_upwind

! GS.f90:1-106
subroutine gauss_seidel()
   use variables
   use ntust_hip
   implicit none
   integer(c_int) :: n
   call ntust_check(ntust_gauss_seidel(u_star, v_star, w_star, p, pre, dt, omega, zeta, &
                                       itmax, n, pChangeMax_), 'gauss_seidel')
   ik = n
end subroutine gauss_seidel

! calcul_new_velocity.f90:1-184
subroutine calcul_new_velocity()
   use variables
   use ntust_hip
   implicit none
   call ntust_check(ntust_calcul_new_velocity(u_star, v_star, w_star, p, ETA, u_solid, &
                                              v_solid, w_solid, dt, istart, iend, u1, v1, &
                                              w1, u2, v2, w2, FX, FY, FZ, uc, vc, wc), &
                    'calcul_new_velocity')
end subroutine calcul_new_velocity

! calcul_new_velocity.f90:195-243
subroutine Updating_velocity()
   use variables
   use ntust_hip
   implicit none
   call ntust_check(ntust_updating_velocity(u2, v2, w2, istart, iend, u, v, w), &
                    'Updating_velocity')
end subroutine Updating_velocity

! virtualForceIntegrator.f90:3-85
subroutine virtualForceIntegrator()
   use variables
   use ntust_hip
   implicit none
   call ntust_check(ntust_virtual_force_integrator(FY, FZ, lx, cDrag, cLift), &
                    'virtualForceIntegrator')
end subroutine virtualForceIntegrator

! boundary_conditions.f90:1-376
subroutine boundary_conditions()
   use variables
   use ntust_hip
   implicit none
   call ntust_push_bc()
   call ntust_check(ntust_boundary_conditions(u, v, w, p, u1, v1, w1, u2, v2, w2, u_star, &
                                              v_star, w_star), 'boundary_conditions')
end subroutine boundary_conditions

! solver_1st_member.f90:1-79, with the reference's explicit arguments: the
! 7-diagonal matrix of the Krylov pressure path (mdim = 7), computed on the
! GPU from the metrics given to ntust_setup.
subroutine solver_1st_member(coef, jcoef, nx, ny, nz, ndim, mdim, iDx, Dxs, iDy, Dys, iDz, Dzs)
   use ntust_hip
   implicit none
   integer :: nx, ny, nz, ndim, mdim
   integer :: jcoef(1:mdim)
   double precision :: coef(1:ndim, 1:mdim)
   double precision :: iDx(-1:nx+2), Dxs(-1:nx+2), iDy(-1:nx+2), Dys(-1:nx+2)
   double precision :: iDz(-1:nx+2), Dzs(-1:nx+2)
   integer(c_int) :: jc(4)
   if (mdim /= 7 .or. ndim /= nx*ny*nz) then
      write (*, '(a)') 'solver_1st_member: need ndim = nx*ny*nz and mdim = 7'
      error stop 1
   end if
   call ntust_check(ntust_poisson_coeffs(coef, jc), 'solver_1st_member')
   jcoef(1:4) = jc
end subroutine solver_1st_member

! solver.f90:1-144 solver_BICG3 with the reference's arguments; the work
! vectors (p, r, r2, v, ss, t) and the MPI arguments are not used: the
! iteration runs on the GPU.  Prints what the reference prints on rank 0.
subroutine solver_BICG3(coef, jcoef, b, x, n, m, eps, p, r, r2, v, ss, t, itmax, myid, &
                        iDIMstart, iDIMend, DIMcount)
   use ntust_hip
   implicit none
   integer, intent(in) :: n, m, itmax
   double precision, intent(in) :: eps
   integer, dimension(m), intent(inout) :: jcoef
   double precision, dimension(n, *), intent(inout) :: coef
   double precision, dimension(n), intent(inout) :: b, x, p, r, r2, v, ss, t
   integer :: myid, iDIMstart, iDIMend
   integer, dimension(1:40) :: DIMcount
   integer(c_int) :: it
   real(c_double) :: nrm
   call ntust_check(ntust_solver_bicg3(coef, jcoef, b, x, n, m, eps, itmax, it, nrm), &
                    'solver_BICG3')
   if (myid == 0) then
      print *, '      '
      if (it >= itmax) then
         print *, 'non convergence =', it, nrm
      else
         print *, '  Iterations BICG3 =', it, nrm
      end if
   end if
end subroutine solver_BICG3

! solver.f90:146-190 matmul_ell: x = A y on the GPU, bit for bit
subroutine matmul_ell(x, coef, jcoef, y, n, m)
   use ntust_hip
   implicit none
   integer, intent(in) :: n, m
   double precision, dimension(n, *), intent(in) :: coef
   integer, dimension(m), intent(in) :: jcoef
   double precision, dimension(n), intent(in) :: y
   double precision, dimension(n), intent(out) :: x
   call ntust_check(ntust_matmul_ell(x, coef, jcoef, y, n, m), 'matmul_ell')
end subroutine matmul_ell
