! qio_probe.f90 -- test fixture: the Fortran runtime's own output of the
! solver's PLOT3D files, and flang's evaluation of the .q restart
! arithmetic, for comparison with the C restatement in oracle/.
!
! Writes with unformatted sequential I/O (the form of filer.f90:173-182 and
! gridder.f90:368-378): a solution file (nblocks; nx ny nz; four reals;
! Q(nx,ny,nz,5)) and a grid file (nblocks; nx ny nz; X, Y, Z), and
! evaluates the reading_variables expressions of reading_data.f90:60-80
! (single-precision face averages, double-precision centring) on the given
! Q, writing u, v, w (halo) and p, uc, vc, wc (interior) as a stream.
!
! usage: qio_probe NX NY NZ IN OUTQ OUTX OUTR
!   IN (stream): Q (nx,ny,nz,5) real*4, Xs(nx), Ys(ny), Zs(nz) real*8,
!                u, v, w (-1:n+2 halo arrays) real*8, the state before
program qio_probe
   implicit none
   integer :: nx, ny, nz, i, j, k, lun
   character(len=512) :: arg, fin, fq, fx, fr
   real, allocatable :: q(:,:,:,:), gx(:,:,:), gy(:,:,:), gz(:,:,:)
   real*8, allocatable :: xs(:), ys(:), zs(:), u(:,:,:), v(:,:,:), w(:,:,:)
   real*8, allocatable :: p(:,:,:), uc(:,:,:), vc(:,:,:), wc(:,:,:)
   real :: hdr = 1.0
   integer, parameter :: nb = 1

   call get_command_argument(1, arg); read (arg, *) nx
   call get_command_argument(2, arg); read (arg, *) ny
   call get_command_argument(3, arg); read (arg, *) nz
   call get_command_argument(4, fin)
   call get_command_argument(5, fq)
   call get_command_argument(6, fx)
   call get_command_argument(7, fr)
   allocate (q(nx, ny, nz, 5), gx(nx, ny, nz), gy(nx, ny, nz), gz(nx, ny, nz))
   allocate (xs(nx), ys(ny), zs(nz))
   allocate (u(-1:nx+2, -1:ny+2, -1:nz+2), v(-1:nx+2, -1:ny+2, -1:nz+2), &
             w(-1:nx+2, -1:ny+2, -1:nz+2))
   allocate (p(nx, ny, nz), uc(nx, ny, nz), vc(nx, ny, nz), wc(nx, ny, nz))

   open (newunit=lun, file=trim(fin), access='stream', form='unformatted', status='old')
   read (lun) q, xs, ys, zs, u, v, w
   close (lun)

   ! solution file, appended twice: position='append' keeps the first copy
   do i = 1, 2
      open (newunit=lun, file=trim(fq), form='unformatted', position='append')
      write (lun) nb
      write (lun) nx, ny, nz
      write (lun) hdr, hdr, hdr, hdr
      write (lun) q
      close (lun)
   end do

   forall (i = 1:nx, j = 1:ny, k = 1:nz)
      gx(i, j, k) = xs(i)
      gy(i, j, k) = ys(j)
      gz(i, j, k) = zs(k)
   end forall
   open (newunit=lun, file=trim(fx), form='unformatted')
   write (lun) nb
   write (lun) nx, ny, nz
   write (lun) gx, gy, gz
   close (lun)

   do k = 1, nz - 1
      do j = 1, ny - 1
         do i = 1, nx - 1
            u(i, j, k) = 0.5*(q(i, j, k, 2) + q(i + 1, j, k, 2))
            v(i, j, k) = 0.5*(q(i, j, k, 3) + q(i, j + 1, k, 3))
            w(i, j, k) = 0.5*(q(i, j, k, 4) + q(i, j, k + 1, 4))
         end do
      end do
   end do
   p = q(:, :, :, 1)
   do k = 1, nz
      do j = 1, ny
         do i = 1, nx
            uc(i, j, k) = 0.5*(u(i, j, k) + u(i - 1, j, k))
            vc(i, j, k) = 0.5*(v(i, j, k) + v(i, j - 1, k))
            wc(i, j, k) = 0.5*(w(i, j, k) + w(i, j, k - 1))
         end do
      end do
   end do
   open (newunit=lun, file=trim(fr), access='stream', form='unformatted', status='replace')
   write (lun) u, v, w, p, uc, vc, wc
   close (lun)
end program qio_probe
