! driver.f90 -- Fortran driver of the GPU hot path.
!
! Follows the solver's initialisation and time loop (main.f90:77-110,
! 139-259): grid, initial state, boundary conditions, then per step
!   CalculateSmagorinskyViscosity, discretisation_QUICK_centre (or _upwind),
!   gauss_seidel, calcul_new_velocity, virtualForceIntegrator,
!   Updating_velocity, boundary_conditions
! through the drop-in shims of hotpath.f90 (mode 0), or the same step fused
! on the device with the state resident in HBM (mode 1).
!
! usage: ntust_driver OUT NX NY NZ RE DT NSTEP SCHEME LES CS OMEGA ZETA ITMAX SOR MODE [IC]
!   SCHEME 0 QUICK_centre, 1 upwind; SOR 0 lexicographic, 1 red-black;
!   IC: optional raw file with u, v, w (halo arrays, real*8) as initial state.
! OUT receives (stream, native endianness): nstep, ik(nstep) int32,
!   cDrag(nstep), cLift(nstep), u, v, w, p (halo arrays), pre, uc, vc, wc.
program ntust_driver
   use variables
   use ntust_hip
   implicit none
   character(len=512) :: out, icfile, arg
   integer :: scheme, sor, mode, s, lun
   integer(c_int), allocatable :: iks(:)
   real*8, allocatable :: cds(:), cls(:), pcs(:)
   real*8 :: t0, t1

   call get_command_argument(1, out)
   nx = iarg(2); ny = iarg(3); nz = iarg(4)
   Re = rarg(5); dt = rarg(6); nstep = iarg(7); scheme = iarg(8); LES = iarg(9)
   Cs = rarg(10); omega = rarg(11); zeta = rarg(12); itmax = rarg(13)
   sor = iarg(14); mode = iarg(15)
   icfile = ''
   if (command_argument_count() >= 16) call get_command_argument(16, icfile)

   nu = 1./Re                      ! reading_data.f90:38
   call allocate_fields()
   call uniform_grid(nx, lx, X, iDx, Dxs)
   call uniform_grid(ny, ly, Y, iDy, Dys)
   call uniform_grid(nz, lz, Z, iDz, Dzs)
   call ntust_setup()
   if (len_trim(icfile) > 0) then
      open (newunit=lun, file=trim(icfile), access='stream', form='unformatted', status='old')
      read (lun) u, v, w
      close (lun)
   end if
   call boundary_conditions()      ! main.f90:100

   allocate (iks(nstep), cds(nstep), cls(nstep), pcs(nstep))
   call cpu_time(t0)
   ! the SOR ordering the gauss_seidel shim uses
   call ntust_check(ntust_configure(dt, nu, Cs, LES, scheme, omega, zeta, itmax, sor, lx, 1), &
                    'ntust_configure')
   if (mode == 0) then
      do istep = 1, nstep
         call CalculateSmagorinskyViscosity()
         if (scheme == 0) then
            call discretisation_QUICK_centre()
         else
            call discretisation_upwind()
         end if
         call gauss_seidel()
         call calcul_new_velocity()
         call virtualForceIntegrator()
         call Updating_velocity()
         call boundary_conditions()
         time = time + dt
         iks(istep) = ik; cds(istep) = cDrag; cls(istep) = cLift
      end do
   else
      call ntust_check(ntust_state_upload(u, v, w, p, ETA), 'ntust_state_upload')
      call ntust_check(ntust_step(nstep, iks, pcs, cds, cls), 'ntust_step')
      call ntust_check(ntust_state_download(u, v, w, p, pre, uc, vc, wc), 'ntust_state_download')
   end if
   call cpu_time(t1)
   write (*, '(a,i0,a,f10.4,a)') 'ntust_driver: ', nstep, ' steps in ', t1 - t0, ' s'

   open (newunit=lun, file=trim(out), access='stream', form='unformatted', status='replace')
   write (lun) int(nstep, c_int)
   write (lun) iks, cds, cls, u, v, w, p, pre, uc, vc, wc
   close (lun)
   call ntust_teardown()

contains

   integer function iarg(n)
      integer, intent(in) :: n
      call get_command_argument(n, arg)
      read (arg, *) iarg
   end function iarg

   real*8 function rarg(n)
      integer, intent(in) :: n
      call get_command_argument(n, arg)
      read (arg, *) rarg
   end function rarg

   ! uniform grid as gridder.f90:235-327: nodes accumulated from 0 with
   ! d = l/n (reading_data.f90:25), cell widths and centre spacings with the
   ! ghost rules of the solver
   subroutine uniform_grid(n, l, Xn, iD, Ds)
      integer, intent(in) :: n
      real*8, intent(in) :: l
      real*8, intent(out) :: Xn(-1:n + 3), iD(-1:n + 2), Ds(-1:n + 2)
      real*8 :: d
      integer :: m
      d = l/n
      Xn(1) = 0.0
      Xn(0) = Xn(1) - d
      Xn(-1) = Xn(0) - d
      do m = 2, n + 3
         Xn(m) = Xn(m - 1) + d
      end do
      do m = 1, n - 1
         iD(m) = Xn(m + 1) - Xn(m)
         Ds(m) = (Xn(m + 2) - Xn(m))/2.0
      end do
      iD(0) = iD(1); iD(-1) = iD(1)
      iD(n) = Xn(n + 1) - Xn(n); iD(n + 1) = iD(n); iD(n + 2) = iD(n)
      Ds(0) = Ds(1); Ds(-1) = Ds(1)
      Ds(n) = Ds(n - 1); Ds(n + 1) = Ds(n - 1); Ds(n + 2) = Ds(n - 1)
   end subroutine uniform_grid

end program ntust_driver
