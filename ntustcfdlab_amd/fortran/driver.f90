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
! usage: ntust_driver OUT NX NY NZ RE DT NSTEP SCHEME LES CS OMEGA ZETA ITMAX SOR MODE
!                     [IC [ISTO [NACA AOA1 START L POLY]]]
!   SCHEME 0 QUICK_centre, 1 upwind; SOR 0 lexicographic, 1 red-black;
!   IC: optional raw file with u, v, w (halo arrays, real*8) as initial state
!   ('-' for none).
!   ISTO > 0: the output of main.f90:113-118,315-328 -- the solution file
!   NNNN.q at the start and every ISTO steps, and check_steady after each
!   but the first -- written into the working directory, in mode 0 by
!   Fortran I/O from the host arrays as the solver's own routines do, in
!   mode 1 by the library from the resident state; the VelocityDifference
!   values go to vdiff.bin (stream: count int32, values real*8).
!   NACA: the airfoil table ('-' for none): RayCasting at start-up with
!   AOA = AOA1 (main.f90:109-110), and from step START on dynamic_AOA and
!   RayCasting before each step (main.f90:144-155), through the shims of
!   airfoil_gpu.f90 in mode 0 and on the resident ETA in mode 1; L: the
!   y and z extent of the grid (the airfoil sits at y = z = 4); POLY: the
!   table's vertex count.
! OUT receives (stream, native endianness): nstep, ik(nstep) int32,
!   cDrag(nstep), cLift(nstep), u, v, w, p (halo arrays), pre, uc, vc, wc.
! MODE 2: the Krylov path instead of time steps -- solver_1st_member,
!   solver_BICG3 (tolerance ZETA, cap ITMAX) from x = 0 for the right-hand
!   side b(nx*ny*nz) read from IC, then matmul_ell of the solution, all
!   through the shims of hotpath.f90 (solver_BICG3 prints the iteration
!   count and norm as the reference does); OUT receives x and A x.
program ntust_driver
   use variables
   use ntust_hip
   implicit none
   character(len=512) :: out, icfile, arg
   integer :: scheme, sor, mode, s, n, lun, nout, q
   logical :: airfoil
   integer(c_int), allocatable :: iks(:)
   real*8, allocatable :: cds(:), cls(:), pcs(:), vds(:)
   real*8 :: t0, t1

   call get_command_argument(1, out)
   nx = iarg(2); ny = iarg(3); nz = iarg(4)
   Re = rarg(5); dt = rarg(6); nstep = iarg(7); scheme = iarg(8); LES = iarg(9)
   Cs = rarg(10); omega = rarg(11); zeta = rarg(12); itmax = rarg(13)
   sor = iarg(14); mode = iarg(15)
   icfile = ''
   if (command_argument_count() >= 16) call get_command_argument(16, icfile)
   if (icfile == '-') icfile = ''
   isto = 0
   if (command_argument_count() >= 17) isto = iarg(17)
   NACA_filename = '-'
   if (command_argument_count() >= 22) then
      call get_command_argument(18, NACA_filename)
      AOA1 = rarg(19); StartDynamic = rarg(20)
      ly = rarg(21); lz = ly; poly = iarg(22)
   end if
   airfoil = NACA_filename /= '-'

   nu = 1./Re                      ! reading_data.f90:38
   call allocate_fields()
   call uniform_grid(nx, lx, X, iDx, Dxs)
   call uniform_grid(ny, ly, Y, iDy, Dys)
   call uniform_grid(nz, lz, Z, iDz, Dzs)
   call ntust_setup()
   if (mode == 2) then
      call krylov()
      call ntust_teardown()
      stop
   end if
   if (len_trim(icfile) > 0) then
      open (newunit=lun, file=trim(icfile), access='stream', form='unformatted', status='old')
      read (lun) u, v, w
      close (lun)
   end if
   call boundary_conditions()      ! main.f90:100
   if (airfoil) then               ! main.f90:109-110
      allocate (az(poly), ay(poly))
      AOA = AOA1
      call RayCasting()
   end if

   allocate (iks(nstep), cds(nstep), cls(nstep), pcs(nstep), vds(nstep + 1))
   nout = 0
   call cpu_time(t0)
   ! the SOR ordering the gauss_seidel shim uses
   call ntust_check(ntust_configure(dt, nu, Cs, LES, scheme, omega, zeta, itmax, sor, lx, 1), &
                    'ntust_configure')
   if (mode == 0) then
      if (isto > 0) call output_host()   ! main.f90:117
      do istep = 1, nstep
         if (airfoil .and. istep >= StartDynamic) then   ! main.f90:144-155
            call dynamic_AOA()
            call RayCasting()
         end if
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
         if (isto > 0) then
            if (mod(istep, isto) == 0) then   ! main.f90:317-328
               VelocityDifference = 1024
               call output_host()
               call steady_host()
               nout = nout + 1; vds(nout) = VelocityDifference
            end if
         end if
      end do
   else
      call ntust_check(ntust_state_upload(u, v, w, p, ETA), 'ntust_state_upload')
      if (isto > 0) call output_resident()
      ! runs of steps end at an output step, and from StartDynamic on every
      ! step re-casts the airfoil at the pitching angle
      s = 1
      do while (s <= nstep)
         n = nstep - s + 1
         if (isto > 0) n = min(n, isto - mod(s - 1, isto))
         if (airfoil) then
            if (s >= StartDynamic) then
               n = 1
               call ntust_check(ntust_dynamic_aoa(AOA1, time, dt, StartDynamic, AOA), &
                                'ntust_dynamic_aoa')
               call ntust_check(ntust_ray_casting(az, ay, poly, AOA, zc, yc, Y, Z, nSubGrids, &
                                                  c_null_ptr), 'ntust_ray_casting')
            else if (StartDynamic <= nstep) then
               n = min(n, int(StartDynamic) - s)
            end if
         end if
         call ntust_check(ntust_step(n, iks(s:), pcs(s:), cds(s:), cls(s:)), 'ntust_step')
         do q = 1, n
            time = time + dt
         end do
         istep = s + n - 1
         if (isto > 0) then
            if (mod(istep, isto) == 0) then
               VelocityDifference = 1024
               call output_resident()
               call ntust_check(ntust_check_steady(istep, isto, VelocityDifference), &
                                'ntust_check_steady')
               nout = nout + 1; vds(nout) = VelocityDifference
            end if
         end if
         s = s + n
      end do
      call ntust_check(ntust_state_download(u, v, w, p, pre, uc, vc, wc), 'ntust_state_download')
   end if
   call cpu_time(t1)
   write (*, '(a,i0,a,f10.4,a)') 'ntust_driver: ', nstep, ' steps in ', t1 - t0, ' s'

   open (newunit=lun, file=trim(out), access='stream', form='unformatted', status='replace')
   write (lun) int(nstep, c_int)
   write (lun) iks, cds, cls, u, v, w, p, pre, uc, vc, wc
   close (lun)
   if (isto > 0) then
      open (newunit=lun, file='vdiff.bin', access='stream', form='unformatted', &
            status='replace')
      write (lun) int(nout, c_int), vds(1:nout)
      close (lun)
   end if
   call ntust_teardown()

contains

   ! the solution file from the host arrays, by Fortran I/O in the form of
   ! filer.f90:161-187 (appended to NNNN.q, num counts the files)
   subroutine output_host()
      Qout(:, :, :, 1) = pre
      Qout(:, :, :, 2) = uc
      Qout(:, :, :, 3) = vc
      Qout(:, :, :, 4) = wc
      Qout(:, :, :, 5) = ETA(1:nx, 1:ny, 1:nz)
      write (filename, '(I4.4)') num
      open (newunit=lun, file=trim(filename)//'.q', position='append', form='unformatted')
      write (lun) nblocks
      write (lun) nx, ny, nz
      write (lun) temp, temp, temp, temp
      write (lun) Qout
      close (lun)
      num = num + 1
   end subroutine output_host

   ! the same file from the resident state: Qout packed on the device
   subroutine output_resident()
      write (filename, '(I4.4)') num
      call ntust_check(ntust_output_q(trim(filename)//'.q'//c_null_char, Qout), &
                       'ntust_output_q')
      num = num + 1
   end subroutine output_resident

   ! the Krylov path through the solver's own call sequence
   subroutine krylov()
      integer :: nd, dc(40)
      integer :: jc(7)
      real*8, allocatable :: cf(:, :), b(:), x(:), ax(:), wk(:, :)
      nd = nx*ny*nz
      allocate (cf(nd, 7), b(nd), x(nd), ax(nd), wk(nd, 6))
      call solver_1st_member(cf, jc, nx, ny, nz, nd, 7, iDx, Dxs, iDy, Dys, iDz, Dzs)
      open (newunit=lun, file=trim(icfile), access='stream', form='unformatted', status='old')
      read (lun) b
      close (lun)
      x = 0
      dc = 0
      call solver_BICG3(cf, jc, b, x, nd, 4, zeta, wk(:, 1), wk(:, 2), wk(:, 3), wk(:, 4), &
                        wk(:, 5), wk(:, 6), int(itmax), 0, 1, nz, dc)
      call matmul_ell(ax, cf, jc, x, nd, 4)
      open (newunit=lun, file=trim(out), access='stream', form='unformatted', status='replace')
      write (lun) x, ax
      close (lun)
   end subroutine krylov

   ! check_steady.f90:1-30 on the host arrays
   subroutine steady_host()
      integer :: i, j, k
      if (istep /= isto) then
         VelocityDifference = 0
         do k = 1, nz
            do j = 1, ny
               do i = 1, nx
                  if (abs(wc(i, j, k) - last_velocity(i, j, k)) > VelocityDifference) &
                     VelocityDifference = abs(wc(i, j, k) - last_velocity(i, j, k))
               end do
            end do
         end do
      end if
      last_velocity(1:nx, 1:ny, 1:nz) = wc
   end subroutine steady_host

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
