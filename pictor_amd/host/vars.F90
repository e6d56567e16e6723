! vars.F90 -- host state in the reference layout (src/vars.F90:23-45).
!
! Fields Ex..Jz(mx,my,mz) with mx = nx+5 (2 low and 3 high ghost layers,
! subdomains.F90:42-50) and the particle SoA qp,xp,yp,zp,up,vp,wp + flvp.
! The HDF5 and MPI modules of the reference are not used: the device holds the
! state during the run and these arrays are its upload/download image.
module vars
   use parameters
   use iso_c_binding
   implicit none
   integer :: mx, my, mz
   real(psn), dimension(:, :, :), allocatable :: Ex, Ey, Ez, Bx, By, Bz
   real(psn), dimension(:, :, :), allocatable :: Jx, Jy, Jz
   integer(c_int32_t), dimension(:), allocatable :: flvp
   real(psn), dimension(:), allocatable :: qp, xp, yp, zp, up, vp, wp
   integer(c_int64_t) :: used_prtl_arr_size = 0
   integer :: t
   type(c_ptr) :: gpu = c_null_ptr                     ! the library context (var_gpu.F90's role)
end module vars
