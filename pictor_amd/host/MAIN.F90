! MAIN.F90 -- the reference's time loop (src/MAIN.F90:42-284) driving the gfx950 particle step.
!
! Each phase subroutine below keeps the name and position of the call it
! replaces in the reference loop and forwards to the C ABI.  Phases the
! one-GPU run does not need (load balancing, particle exchange between ranks)
! are left out.  Usage:  pictor_main [name=value ...]  with names from
! parameters.F90 (nx, ny, epc, tfinish, form, fused, fdtd, curr_filter, seed, c,
! c_ompe, energy_save_period, energy_file, fld_save_period, fsave_ratio,
! restart_save_period, data_folder, restart_file, prtl_reorder_period).
program MAIN
   use iso_c_binding
   use parameters
   use vars
   use setup
   use pictor_cabi
   implicit none
   integer(c_int64_t) :: c0, c1, crate
   real(c_double) :: secs

   integer :: tstart
   integer(c_int64_t) :: t_restart

   call ReadArgs
   call Initialisations
   tstart = 1
   if (len_trim(restart_file) > 0) then   ! the restart branch, MAIN.F90:35-39
      call check(pictor_load_state(gpu, trim(restart_file)//c_null_char, t_restart), 'pictor_load_state')
      tstart = int(t_restart) + 1
   end if
   call SaveEnergy(tstart - 1)
   call system_clock(c0, crate)
   do t = tstart, tfinish
      if (fused == 1) then
         call check(pictor_step(gpu, 1_c_int32_t), 'pictor_step')
      else
         call UpdateBfieldHalf          ! MAIN.F90:51/54, SendRecvFlds(B) :58
         call MoveDepositPrtl           ! MAIN.F90:65/68
         call ExchangePrtl              ! MAIN.F90:88/92 (+ AppendParticles :147)
         call UpdateBfieldHalf          ! MAIN.F90:103/114
         call UpdateMagFld_PML          ! MAIN.F90:104
         call UpdateEfield              ! MAIN.F90:108/115
         call UpdateElcFld_PML          ! MAIN.F90:109
         call UpdateCurrentAllFaces     ! MAIN.F90:124/129
         call smoothen_current          ! MAIN.F90:158/162
         call AddCurrent                ! MAIN.F90:171/174, SendRecvFlds(E) :184/190
         call EnforceBC_Final           ! MAIN.F90:209
         call ResetCurrent              ! MAIN.F90:248/251
         call ReorderPrtl               ! MAIN.F90:261/264
      end if
      if (energy_save_period > 0) then
         if (mod(t, energy_save_period) == 0) call SaveEnergy(t)
      end if
      call Savedata(t)                  ! MAIN.F90:229
   end do
   call check(pictor_synchronize(gpu), 'pictor_synchronize')
   call system_clock(c1)
   secs = real(c1 - c0, c_double)/real(crate, c_double)
   write (*, '(a,i0,a,es12.4,a,es12.4)') 'steps ', tfinish, '  seconds ', secs, &
      '  particle-updates/s ', real(used_prtl_arr_size, c_double)*tfinish/max(secs, 1d-30)
   call check(pictor_finalize(gpu), 'pictor_finalize')
   write (*, '(a)') 'Simulation Finished!'

contains

   subroutine check(rc, where)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      call pictor_check(rc, where)
   end subroutine check

   subroutine ReadArgs
      integer :: i, k, ios
      character(len=256) :: arg, key, val
      do i = 1, command_argument_count()
         call get_command_argument(i, arg)
         k = index(arg, '=')
         if (k == 0) cycle
         key = arg(1:k - 1)
         val = arg(k + 1:)
         select case (trim(key))
         case ('nx'); read (val, *, iostat=ios) nx
         case ('ny'); read (val, *, iostat=ios) ny
         case ('epc'); read (val, *, iostat=ios) epc
         case ('tfinish'); read (val, *, iostat=ios) tfinish
         case ('form'); read (val, *, iostat=ios) form
         case ('fused'); read (val, *, iostat=ios) fused
         case ('fdtd'); read (val, *, iostat=ios) fdtd
         case ('curr_filter'); read (val, *, iostat=ios) curr_filter
         case ('seed'); read (val, *, iostat=ios) seed
         case ('c'); read (val, *, iostat=ios) c
         case ('c_ompe'); read (val, *, iostat=ios) c_ompe
         case ('energy_save_period'); read (val, *, iostat=ios) energy_save_period
         case ('prtl_reorder_period'); read (val, *, iostat=ios) prtl_reorder_period
         case ('energy_file'); read (val, *, iostat=ios) energy_file
         case ('fld_save_period'); read (val, *, iostat=ios) fld_save_period
         case ('fsave_ratio'); read (val, *, iostat=ios) fsave_ratio
         case ('restart_save_period'); read (val, *, iostat=ios) restart_save_period
         case ('data_folder'); data_folder = val
         case ('restart_file'); restart_file = val
         case default
            write (*, '(a,a)') 'unknown parameter ', trim(key)
            error stop 2
         end select
      end do
   end subroutine ReadArgs

   ! Initialisations, MAIN.F90:338-395: InitScale and the allocations happen in
   ! pictor_init; InitSetup loads the particles; InitAll_gpu is the upload.
   subroutine Initialisations
      type(pictor_config) :: cfg
      mx = nx + 5
      my = ny + 5
      mz = 1
      allocate (Ex(mx, my, mz), Ey(mx, my, mz), Ez(mx, my, mz))
      allocate (Bx(mx, my, mz), By(mx, my, mz), Bz(mx, my, mz))
      allocate (Jx(mx, my, mz), Jy(mx, my, mz), Jz(mx, my, mz))
      Ex = 0; Ey = 0; Ez = 0; Bx = 0; By = 0; Bz = 0; Jx = 0; Jy = 0; Jz = 0
      call InitSetup
      cfg = pictor_config(2, nx, ny, 1, epc, form, 0, fdtd, PICTOR_GEOM_CARTESIAN, curr_filter, c, c_ompe, g0, mi_me, &
                          0.0d0, 0.0d0)
      call check(pictor_init(cfg, gpu), 'pictor_init')
      call check(pictor_upload_fields(gpu, Ex, Ey, Ez, Bx, By, Bz), 'pictor_upload_fields')
      call check(pictor_upload_particles(gpu, used_prtl_arr_size, qp, xp, yp, zp, up, vp, wp, flvp), &
                 'pictor_upload_particles')
      write (*, '(a,i0,a,i0,a,i0,a,i0)') 'nx ', nx, ' ny ', ny, ' particles ', used_prtl_arr_size, ' form ', form
   end subroutine Initialisations

   subroutine UpdateBfieldHalf
      call check(pictor_update_b_half(gpu), 'UpdateBfieldHalf')
   end subroutine UpdateBfieldHalf

   subroutine MoveDepositPrtl
      call check(pictor_move_deposit(gpu), 'MoveDepositPrtl')
   end subroutine MoveDepositPrtl

   subroutine ExchangePrtl
      call check(pictor_exchange_particles(gpu), 'ExchangePrtl')
   end subroutine ExchangePrtl

   subroutine UpdateEfield
      call check(pictor_update_e(gpu), 'UpdateEfield')
   end subroutine UpdateEfield

   subroutine UpdateCurrentAllFaces
      call check(pictor_current_halo_add(gpu), 'UpdateCurrentAllFaces')
   end subroutine UpdateCurrentAllFaces

   subroutine smoothen_current
      call check(pictor_smoothen_current(gpu), 'smoothen_current')
   end subroutine smoothen_current

   subroutine AddCurrent
      call check(pictor_add_current(gpu), 'AddCurrent')
   end subroutine AddCurrent

   subroutine ResetCurrent
      call check(pictor_reset_current(gpu), 'ResetCurrent')
   end subroutine ResetCurrent

   subroutine UpdateMagFld_PML
      call check(pictor_pml(gpu, 1_c_int32_t), 'UpdateMagFld_PML')
   end subroutine UpdateMagFld_PML

   subroutine UpdateElcFld_PML
      call check(pictor_pml(gpu, 0_c_int32_t), 'UpdateElcFld_PML')
   end subroutine UpdateElcFld_PML

   subroutine EnforceBC_Final
      call check(pictor_enforce_bc_final(gpu), 'EnforceBC_Final')
   end subroutine EnforceBC_Final

   ! Savedata (savedata.F90:11-60): field files and restart files at their periods
   subroutine Savedata(step)
      integer, intent(in) :: step
      character(len=32) :: ts
      write (ts, '(i0)') step
      if (fld_save_period > 0) then
         if (mod(step, fld_save_period) == 0) &
            call check(pictor_save_fields(gpu, trim(data_folder)//'/fld_'//trim(ts)//c_null_char, &
                                          int(fsave_ratio, c_int32_t), 1_c_int32_t), 'save_field_collective')
      end if
      if (restart_save_period > 0) then
         if (mod(step, restart_save_period) == 0) &
            call check(pictor_save_state(gpu, trim(data_folder)//'/restart_'//trim(ts)//c_null_char, &
                                         int(step, c_int64_t)), 'SaveRestartData')
      end if
   end subroutine Savedata

   subroutine ReorderPrtl
      if (prtl_reorder_period > 0) then
         if (modulo(t, prtl_reorder_period) == 0) call check(pictor_sort_tiles(gpu), 'ReorderPrtl')
      end if
   end subroutine ReorderPrtl

   ! CalcEnergy (savedata_routines.F90:265-327) on the device, FP64 reductions
   subroutine SaveEnergy(step)
      integer, intent(in) :: step
      real(c_double) :: d(9)
      character(len=32) :: ts
      call check(pictor_diagnostics(gpu, d), 'pictor_diagnostics')
      write (*, '(a,i8,5es26.17)') 'energy ', step, d(1), d(2), d(3), d(4), d(8)
      if (energy_file == 1) then   ! save_total_energy, savedata.F90:856-890
         write (ts, '(i0)') step
         call check(pictor_save_energy(gpu, trim(data_folder)//'/energy_'//trim(ts)//c_null_char), &
                    'save_total_energy')
      end if
   end subroutine SaveEnergy

end program MAIN
