! airfoil_gpu.f90 -- drop-in replacements for RayCasting and dynamic_AOA
! (airfoil.f90:311-321, 580-892) with the reference's names and empty
! argument lists, on the module variables.  Linking this file instead of
! airfoil.f90 leaves main.f90:109-110 and 144-155 unchanged (the other
! routines of airfoil.f90, MATRIX_PARSEC and RayCasting_HSUAN, are not
! called by main.f90).
!
! RayCasting reads the polygon table as the reference does and computes ETA
! on the GPU into the module array ETA (and the library's resident ETA).
! Unlike the reference, az and ay keep the table's values afterwards (the
! rotated polygon stays on the device).

subroutine RayCasting()
   use variables
   use ntust_hip, only: ntust_check
   use iso_c_binding
   implicit none
   ! ETA bound as an array here (the module array needs no TARGET
   ! attribute); ntust_hip's interface binds it as a pointer for the
   ! resident-only form
   interface
      integer(c_int) function cast_into(az0, ay0, poly, AOA, zc, yc, Yn, Zn, nsub, ETA) &
            bind(C, name="ntust_ray_casting")
         import :: c_int, c_double
         real(c_double), intent(in) :: az0(*), ay0(*), Yn(*), Zn(*)
         integer(c_int), value :: poly, nsub
         real(c_double), value :: AOA, zc, yc
         real(c_double) :: ETA(*)
      end function
   end interface
   integer :: lun, i
   open (newunit=lun, file=trim(NACA_filename), form='formatted', status='old')
   do i = 1, poly
      read (lun, *) az(i), ay(i)
   end do
   close (lun)
   call ntust_check(cast_into(az, ay, poly, AOA, zc, yc, Y, Z, nSubGrids, ETA), &
                    'ntust_ray_casting')
end subroutine RayCasting

subroutine dynamic_AOA()
   use variables
   use ntust_hip
   implicit none
   call ntust_check(ntust_dynamic_aoa(AOA1, time, dt, StartDynamic, AOA), 'ntust_dynamic_aoa')
end subroutine dynamic_AOA
