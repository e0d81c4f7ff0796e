! ntust_hip.f90 -- iso_c_binding interface to libntust_hip.so (include/ntust.h).
!
! Arrays bind as assumed-size real(c_double) dummies, so passing a Fortran
! module array passes its base address without a copy; optional arrays
! bind as type(c_ptr), value (pass c_loc(a) or c_null_ptr).  Every function
! returns the library status: 0 ok, > 0 warning, < 0 error.
module ntust_hip
   use iso_c_binding
   implicit none

   integer(c_int), parameter :: NTUST_WARN_ITMAX = 1, NTUST_E_IO = -6
   integer(c_int), parameter :: NTUST_SOR_LEXICOGRAPHIC = 0, NTUST_SOR_REDBLACK = 1, &
                                NTUST_SOR_MIXED = 2
   integer(c_int), parameter :: NTUST_QUICK_CENTRE = 0, NTUST_UPWIND = 1
   integer(c_int), parameter :: NTUST_ARITH_EXACT = 0, NTUST_ARITH_FAST = 1, &
                                NTUST_ARITH_FP32 = 2

   interface
      integer(c_int) function ntust_init(nx, ny, nz, iDx, Dxs, iDy, Dys, iDz, Dzs) &
            bind(C, name="ntust_init")
         import :: c_int, c_double
         integer(c_int), value :: nx, ny, nz
         real(c_double), intent(in) :: iDx(*), Dxs(*), iDy(*), Dys(*), iDz(*), Dzs(*)
      end function

      integer(c_int) function ntust_finalize() bind(C, name="ntust_finalize")
         import :: c_int
      end function

      type(c_ptr) function ntust_last_error() bind(C, name="ntust_last_error")
         import :: c_ptr
      end function

      integer(c_int) function ntust_set_bc(codes) bind(C, name="ntust_set_bc")
         import :: c_int
         integer(c_int), intent(in) :: codes(18)
      end function

      integer(c_int) function ntust_gradient_phi_gauss(Phi, prebase, dPhidX, Dxm, Dym, Dzm, &
            istart, iend) bind(C, name="ntust_gradient_phi_gauss")
         import :: c_int, c_double, c_ptr
         real(c_double), intent(in) :: Phi(*)
         type(c_ptr), value :: prebase
         real(c_double), intent(inout) :: dPhidX(*)
         real(c_double), intent(in) :: Dxm(*), Dym(*), Dzm(*)
         integer(c_int), value :: istart, iend
      end function

      integer(c_int) function ntust_smagorinsky(u, v, w, pre_u, pre_v, pre_w, Cs, istart, iend, &
            dudx, dvdx, dwdx, Viseff) bind(C, name="ntust_smagorinsky")
         import :: c_int, c_double, c_ptr
         real(c_double), intent(in) :: u(*), v(*), w(*)
         type(c_ptr), value :: pre_u, pre_v, pre_w
         real(c_double), value :: Cs
         integer(c_int), value :: istart, iend
         real(c_double), intent(inout) :: dudx(*), dvdx(*), dwdx(*)
         real(c_double), intent(inout) :: Viseff(*)
      end function

      integer(c_int) function ntust_discretisation_quick_centre(u, v, w, Viseff, dt, nu, LES, &
            istart, iend, u_star, v_star, w_star) bind(C, name="ntust_discretisation_quick_centre")
         import :: c_int, c_double
         real(c_double), intent(in) :: u(*), v(*), w(*), Viseff(*)
         real(c_double), value :: dt, nu
         integer(c_int), value :: LES, istart, iend
         real(c_double), intent(inout) :: u_star(*), v_star(*), w_star(*)
      end function

      integer(c_int) function ntust_discretisation_upwind(u, v, w, dt, nu, u_star, v_star, &
            w_star) bind(C, name="ntust_discretisation_upwind")
         import :: c_int, c_double
         real(c_double), intent(in) :: u(*), v(*), w(*)
         real(c_double), value :: dt, nu
         real(c_double), intent(inout) :: u_star(*), v_star(*), w_star(*)
      end function

      integer(c_int) function ntust_gauss_seidel(u_star, v_star, w_star, p, pre, dt, omega, &
            zeta, itmax, ik, pChangeMax) bind(C, name="ntust_gauss_seidel")
         import :: c_int, c_double
         real(c_double), intent(in) :: u_star(*), v_star(*), w_star(*)
         real(c_double), intent(inout) :: p(*), pre(*)
         real(c_double), value :: dt, omega, zeta, itmax
         integer(c_int), intent(out) :: ik
         real(c_double), intent(out) :: pChangeMax
      end function

      integer(c_int) function ntust_calcul_new_velocity(u_star, v_star, w_star, p, ETA, &
            u_solid, v_solid, w_solid, dt, istart, iend, u1, v1, w1, u2, v2, w2, FX, FY, FZ, &
            uc, vc, wc) bind(C, name="ntust_calcul_new_velocity")
         import :: c_int, c_double
         real(c_double), intent(in) :: u_star(*), v_star(*), w_star(*), p(*), ETA(*)
         real(c_double), value :: u_solid, v_solid, w_solid, dt
         integer(c_int), value :: istart, iend
         real(c_double), intent(inout) :: u1(*), v1(*), w1(*), u2(*), v2(*), w2(*)
         real(c_double), intent(inout) :: FX(*), FY(*), FZ(*), uc(*), vc(*), wc(*)
      end function

      integer(c_int) function ntust_updating_velocity(u2, v2, w2, istart, iend, u, v, w) &
            bind(C, name="ntust_updating_velocity")
         import :: c_int, c_double
         real(c_double), intent(in) :: u2(*), v2(*), w2(*)
         integer(c_int), value :: istart, iend
         real(c_double), intent(inout) :: u(*), v(*), w(*)
      end function

      integer(c_int) function ntust_virtual_force_integrator(FY, FZ, lx, cDrag, cLift) &
            bind(C, name="ntust_virtual_force_integrator")
         import :: c_int, c_double
         real(c_double), intent(in) :: FY(*), FZ(*)
         real(c_double), value :: lx
         real(c_double), intent(out) :: cDrag, cLift
      end function

      integer(c_int) function ntust_boundary_conditions(u, v, w, p, u1, v1, w1, u2, v2, w2, &
            u_star, v_star, w_star) bind(C, name="ntust_boundary_conditions")
         import :: c_int, c_double
         real(c_double), intent(inout) :: u(*), v(*), w(*), p(*)
         real(c_double), intent(inout) :: u1(*), v1(*), w1(*), u2(*), v2(*), w2(*)
         real(c_double), intent(inout) :: u_star(*), v_star(*), w_star(*)
      end function

      integer(c_int) function ntust_configure(dt, nu, Cs, LES, scheme, omega, zeta, itmax, &
            sor_mode, lx, force_every) bind(C, name="ntust_configure")
         import :: c_int, c_double
         real(c_double), value :: dt, nu, Cs
         integer(c_int), value :: LES, scheme
         real(c_double), value :: omega, zeta, itmax
         integer(c_int), value :: sor_mode
         real(c_double), value :: lx
         integer(c_int), value :: force_every
      end function

      integer(c_int) function ntust_state_upload(u, v, w, p, ETA) &
            bind(C, name="ntust_state_upload")
         import :: c_int, c_double
         real(c_double), intent(in) :: u(*), v(*), w(*), p(*), ETA(*)
      end function

      integer(c_int) function ntust_state_download(u, v, w, p, pre, uc, vc, wc) &
            bind(C, name="ntust_state_download")
         import :: c_int, c_double
         real(c_double), intent(inout) :: u(*), v(*), w(*), p(*), pre(*), uc(*), vc(*), wc(*)
      end function

      integer(c_int) function ntust_step(nsteps, ik, pchange, cdrag, clift) &
            bind(C, name="ntust_step")
         import :: c_int, c_double
         integer(c_int), value :: nsteps
         integer(c_int), intent(out) :: ik(*)
         real(c_double), intent(out) :: pchange(*), cdrag(*), clift(*)
      end function

      ! airfoil geometry: ETA is type(c_ptr) (c_loc(ETA), or c_null_ptr for
      ! the resident ETA only)
      integer(c_int) function ntust_ray_casting(az0, ay0, poly, AOA, zc, yc, Yn, Zn, nsub, &
                                                ETA) bind(C, name="ntust_ray_casting")
         import :: c_int, c_double, c_ptr
         real(c_double), intent(in) :: az0(*), ay0(*), Yn(*), Zn(*)
         integer(c_int), value :: poly, nsub
         real(c_double), value :: AOA, zc, yc
         type(c_ptr), value :: ETA
      end function

      integer(c_int) function ntust_dynamic_aoa(AOA1, time, dt, StartDynamic, AOA) &
            bind(C, name="ntust_dynamic_aoa")
         import :: c_int, c_double
         real(c_double), value :: AOA1, time, dt, StartDynamic
         real(c_double), intent(out) :: AOA
      end function

      ! output, steady check and restart on the resident state; paths are
      ! C strings (trim(name)//c_null_char)
      integer(c_int) function ntust_output_q(path, qout) bind(C, name="ntust_output_q")
         import :: c_int, c_char, c_float
         character(kind=c_char), dimension(*), intent(in) :: path
         real(c_float) :: qout(*)
      end function

      integer(c_int) function ntust_read_q(path, qout) bind(C, name="ntust_read_q")
         import :: c_int, c_char, c_float
         character(kind=c_char), dimension(*), intent(in) :: path
         real(c_float) :: qout(*)
      end function

      integer(c_int) function ntust_reading_variables(qout) &
            bind(C, name="ntust_reading_variables")
         import :: c_int, c_float
         real(c_float), intent(in) :: qout(*)
      end function

      integer(c_int) function ntust_write_mesh_x(path, Xs, Ys, Zs) &
            bind(C, name="ntust_write_mesh_x")
         import :: c_int, c_char, c_double
         character(kind=c_char), dimension(*), intent(in) :: path
         real(c_double), intent(in) :: Xs(*), Ys(*), Zs(*)
      end function

      integer(c_int) function ntust_check_steady(istep, isto, vdiff) &
            bind(C, name="ntust_check_steady")
         import :: c_int, c_double
         integer(c_int), value :: istep, isto
         real(c_double), intent(inout) :: vdiff
      end function

      integer(c_int) function ntust_restart_write(path, istep, time) &
            bind(C, name="ntust_restart_write")
         import :: c_int, c_char, c_double
         character(kind=c_char), dimension(*), intent(in) :: path
         integer(c_int), value :: istep
         real(c_double), value :: time
      end function

      integer(c_int) function ntust_restart_read(path, istep, time) &
            bind(C, name="ntust_restart_read")
         import :: c_int, c_char, c_double
         character(kind=c_char), dimension(*), intent(in) :: path
         integer(c_int), intent(out) :: istep
         real(c_double), intent(out) :: time
      end function

      ! members of the Krylov pressure path (solver_1st_member.f90,
      ! solver_pressure_field.f90; the right-hand side in the form GS.f90 uses)
      integer(c_int) function ntust_poisson_coeffs(coef, jcoef) &
            bind(C, name="ntust_poisson_coeffs")
         import :: c_int, c_double
         real(c_double), intent(out) :: coef(*)
         integer(c_int), intent(out) :: jcoef(4)
      end function

      integer(c_int) function ntust_matmul_ell(x, coef, jcoef, y, n, m) &
            bind(C, name="ntust_matmul_ell")
         import :: c_int, c_double
         real(c_double), intent(out) :: x(*)
         real(c_double), intent(in) :: coef(*), y(*)
         integer(c_int), intent(in) :: jcoef(*)
         integer(c_int), value :: n, m
      end function

      integer(c_int) function ntust_solver_bicg3(coef, jcoef, b, x, n, m, eps, itmax, &
                                                 iterations, norm) &
            bind(C, name="ntust_solver_bicg3")
         import :: c_int, c_double
         real(c_double), intent(in) :: coef(*), b(*)
         integer(c_int), intent(in) :: jcoef(*)
         real(c_double), intent(inout) :: x(*)
         integer(c_int), value :: n, m, itmax
         real(c_double), value :: eps
         integer(c_int), intent(out) :: iterations
         real(c_double), intent(out) :: norm
      end function

      integer(c_int) function ntust_poisson_rhs(u_star, v_star, w_star, dt, b) &
            bind(C, name="ntust_poisson_rhs")
         import :: c_int, c_double
         real(c_double), intent(in) :: u_star(*), v_star(*), w_star(*)
         real(c_double), value :: dt
         real(c_double), intent(out) :: b(*)
      end function

      integer(c_int) function ntust_pressure_field(x, p, pre) &
            bind(C, name="ntust_pressure_field")
         import :: c_int, c_double
         real(c_double), intent(in) :: x(*)
         real(c_double), intent(inout) :: p(*), pre(*)
      end function

      ! discrete adjoint of the resident step (no counterpart in the reference)
      integer(c_int) function ntust_set_body_force(fu, fv, fw) &
            bind(C, name="ntust_set_body_force")
         import :: c_int, c_ptr
         type(c_ptr), value :: fu, fv, fw
      end function

      integer(c_int) function ntust_kinetic_energy(ke) bind(C, name="ntust_kinetic_energy")
         import :: c_int, c_double
         real(c_double), intent(out) :: ke
      end function

      integer(c_int) function ntust_checkpoint_reserve(nslots) &
            bind(C, name="ntust_checkpoint_reserve")
         import :: c_int
         integer(c_int), value :: nslots
      end function

      integer(c_int) function ntust_checkpoint_save(slot) bind(C, name="ntust_checkpoint_save")
         import :: c_int
         integer(c_int), value :: slot
      end function

      integer(c_int) function ntust_checkpoint_load(slot) bind(C, name="ntust_checkpoint_load")
         import :: c_int
         integer(c_int), value :: slot
      end function

      integer(c_int) function ntust_adjoint_set(lu, lv, lw, lp) bind(C, name="ntust_adjoint_set")
         import :: c_int, c_ptr
         type(c_ptr), value :: lu, lv, lw, lp
      end function

      integer(c_int) function ntust_adjoint_get(lu, lv, lw, lp) bind(C, name="ntust_adjoint_get")
         import :: c_int, c_ptr
         type(c_ptr), value :: lu, lv, lw, lp
      end function

      integer(c_int) function ntust_adjoint_add_ke(scale) bind(C, name="ntust_adjoint_add_ke")
         import :: c_int, c_double
         real(c_double), value :: scale
      end function

      integer(c_int) function ntust_adjoint_step(ik) bind(C, name="ntust_adjoint_step")
         import :: c_int
         integer(c_int), intent(out) :: ik
      end function

      integer(c_int) function ntust_adjoint_gradient(gfu, gfv, gfw) &
            bind(C, name="ntust_adjoint_gradient")
         import :: c_int, c_ptr
         type(c_ptr), value :: gfu, gfv, gfw
      end function

      integer(c_int) function ntust_adjoint_run(nsteps, checkpoint_every, navg, J, ik, ms) &
            bind(C, name="ntust_adjoint_run")
         import :: c_int, c_double
         integer(c_int), value :: nsteps, checkpoint_every, navg
         real(c_double), intent(out) :: J
         integer(c_int), intent(out) :: ik(*)
         real(c_double), intent(out) :: ms(4)
      end function

      integer(c_int) function ntust_adjoint_objective(w_ke, w_cd, w_cl, eta_gradient) &
            bind(C, name="ntust_adjoint_objective")
         import :: c_int, c_double
         real(c_double), value :: w_ke, w_cd, w_cl
         integer(c_int), value :: eta_gradient
      end function

      integer(c_int) function ntust_adjoint_eta_gradient(geta) &
            bind(C, name="ntust_adjoint_eta_gradient")
         import :: c_int, c_ptr
         type(c_ptr), value :: geta
      end function

      integer(c_int) function ntust_adjoint_checkpoint_plan(nsteps, reserve_bytes, every, &
            bytes) bind(C, name="ntust_adjoint_checkpoint_plan")
         import :: c_int, c_double, c_ptr
         integer(c_int), value :: nsteps
         real(c_double), value :: reserve_bytes
         integer(c_int), intent(out) :: every
         type(c_ptr), value :: bytes
      end function

      ! predictor and eddy-viscosity arithmetic of the resident step
      ! (NTUST_ARITH_*; no counterpart in the reference)
      integer(c_int) function ntust_set_arithmetic(mode) bind(C, name="ntust_set_arithmetic")
         import :: c_int
         integer(c_int), value :: mode
      end function

      ! Lagrangian markers of a body at rest (no counterpart in the
      ! reference): n markers at (zm, ym) with weights wm, kernel widths
      ! hx, hy, hz; Xn, Yn, Zn the node coordinates X(-1:nx+3) etc.
      integer(c_int) function ntust_set_markers(n, zm, ym, wm, hx, hy, hz, Xn, Yn, Zn) &
            bind(C, name="ntust_set_markers")
         import :: c_int, c_double
         integer(c_int), value :: n
         real(c_double), intent(in) :: zm(*), ym(*), wm(*), Xn(*), Yn(*), Zn(*)
         real(c_double), value :: hx, hy, hz
      end function

      integer(c_int) function ntust_marker_velocity(U) bind(C, name="ntust_marker_velocity")
         import :: c_int, c_double
         real(c_double), intent(out) :: U(*)
      end function

      integer(c_int) function ntust_adjoint_marker_gradient(gz, gy) &
            bind(C, name="ntust_adjoint_marker_gradient")
         import :: c_int, c_double
         real(c_double), intent(out) :: gz(*), gy(*)
      end function

      ! force integral on device arrays, allocating nothing (for callers
      ! that cannot allocate: graph capture, serving loops)
      integer(c_int) function ntust_virtual_force_integrator_noalloc(FY, FZ, lx, coef) &
            bind(C, name="ntust_virtual_force_integrator_noalloc")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: FY, FZ, coef
         real(c_double), value :: lx
      end function
   end interface

contains

   ! Stops the program on a library error, as the reference stops on a
   ! Fortran runtime error; warnings (status > 0) continue silently, as the
   ! reference continues after a non-converged pressure solve.
   subroutine ntust_check(rc, where)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      character(kind=c_char), pointer :: msg(:)
      integer :: n
      if (rc >= 0) return
      call c_f_pointer(ntust_last_error(), msg, [512])
      n = 0
      do while (n < 512)
         if (msg(n + 1) == c_null_char) exit
         n = n + 1
      end do
      write (*, '(a,a,a,i0,a,512a)') 'libntust_hip: ', where, ' failed (', rc, '): ', msg(1:n)
      error stop 1
   end subroutine

end module ntust_hip
