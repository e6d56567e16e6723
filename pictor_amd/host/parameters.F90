! parameters.F90 -- run parameters of the Fortran host (parameters.F90:5-87 of the reference).
!
! The reference fixes these at compile time; the same names are kept here.
! The defaults are the 2-D Weibel configuration of the benchmark (256 x 256
! cells, 16 particles per cell per species, Courant 0.45, 10 cells per skin
! depth, e-/e+ pair plasma).  Every value can be overridden on the command
! line as name=value (see MAIN.F90).
module parameters
   implicit none
   integer, parameter :: psn = kind(1.0d0)          ! computation precision of the host arrays
   integer :: nx = 256                               ! no. of grid cells along x
   integer :: ny = 256                               ! no. of grid cells along y
   integer :: nz = 1                                 ! 2-D
   integer :: tfinish = 100                          ! number of time steps
   real(psn) :: g0 = 1.0_psn                         ! Lorentz factor scale of the skin depth
   integer :: epc = 16                               ! particles per cell per species
   real(psn) :: mi_me = 1.0_psn                      ! pair plasma
   real(psn) :: c_ompe = 10.0_psn                    ! cells per skin depth
   real(psn) :: c = 0.45_psn                         ! Courant number
   integer :: curr_filter = 0                        ! passes of the 1-2-1 current filter per step
   integer :: prtl_reorder_period = 20               ! counting sort into tiles every 20 steps
   integer :: energy_save_period = 10                ! print CalcEnergy every this many steps
   integer :: energy_file = 0                        ! 1: also write data_folder/energy_<t> (save_total_energy)
   integer :: fld_save_period = 0                    ! write data_folder/fld_<t> every this many steps (0: never)
   integer :: fsave_ratio = 1                        ! sampling of the field files (CollectFld)
   integer :: restart_save_period = 0                ! write data_folder/restart_<t> every this many steps
   character(len=200) :: data_folder = 'data'        ! output folder (must exist)
   character(len=200) :: restart_file = ''           ! start from this restart file (empty: from the setup)
   ! run-time choices of the GPU path
   integer :: form = 1                               ! 0 = FP64 reference SoA, 1 = FP32 split
   integer :: fused = 0                              ! 1 = pictor_step per step, 0 = one call per phase
   integer :: fdtd = 0                               ! 0 = Yee (em_update.F90), 1 = Lehe (em_update_lehe.F90)
   real(psn) :: drift = 0.5_psn                      ! beam drift (units of c)
   real(psn) :: uth = 0.01_psn                       ! thermal spread of gamma*beta
   integer :: seed = 12345
end module parameters
