! setup_weibel.F90 -- InitSetup of the 2-D Weibel run (setup_weibel.F90:17-26 of the reference).
!
! The reference calls InitPrtl twice with opposite drifts, placing both
! flavours at every random position (init_prtl.F90:138-221).  The same load is
! produced by the library's seeded counter-based generator, so that a run is
! reproducible independently of the Fortran runtime's random_number.
module setup
   use parameters
   use vars
   use pictor_cabi
   implicit none
contains

   subroutine InitSetup
      integer(c_int64_t) :: n, nout
      integer(c_int) :: rc
      n = 2_c_int64_t*epc*nx*ny
      allocate (qp(n), xp(n), yp(n), zp(n), up(n), vp(n), wp(n), flvp(n))
      rc = pictor_setup_weibel(2_c_int32_t, int(nx, c_int32_t), int(ny, c_int32_t), 1_c_int32_t, &
                               int(epc, c_int32_t), real(drift, c_double), real(uth, c_double), &
                               int(seed, c_int64_t), n, qp, xp, yp, zp, up, vp, wp, flvp, nout)
      call pictor_check(rc, 'InitSetup')
      used_prtl_arr_size = nout
   end subroutine InitSetup

end module setup
