! pictor_cabi.F90 -- ISO_C_BINDING interface to libpictor_hip.so (include/pictor_hip.h).
!
! One interface per C entry point.  Arrays are passed as assumed-size
! real(c_double)/integer(c_int32_t) dummies, i.e. by address, in the layout the
! reference keeps them in (src/vars.F90:23-24, 43-45): the Fortran arrays are
! handed over unchanged.
module pictor_cabi
   use iso_c_binding
   implicit none

   integer(c_int), parameter :: PICTOR_OK = 0
   integer(c_int), parameter :: PICTOR_ERR_CAPACITY = 12
   integer(c_int32_t), parameter :: PICTOR_FORM_FP64 = 0
   integer(c_int32_t), parameter :: PICTOR_FORM_FP32_SPLIT = 1
   integer(c_int32_t), parameter :: PICTOR_FDTD_YEE = 0
   integer(c_int32_t), parameter :: PICTOR_FDTD_LEHE = 1
   integer(c_int32_t), parameter :: PICTOR_GEOM_CARTESIAN = 0
   integer(c_int32_t), parameter :: PICTOR_GEOM_CYL = 1
   integer(c_int32_t), parameter :: PICTOR_GEOM_CYL_MODAL = 2
   ! boundary face types (bc_face%type_fld / type_prtl: 'prdc', 'cond', 'open', 'pml', 'refl', 'remv')
   integer(c_int32_t), parameter :: PICTOR_BC_PERIODIC = 0
   integer(c_int32_t), parameter :: PICTOR_BC_COND = 1
   integer(c_int32_t), parameter :: PICTOR_BC_OPEN = 2
   integer(c_int32_t), parameter :: PICTOR_BC_PML = 3
   integer(c_int32_t), parameter :: PICTOR_BC_REFL = 4
   integer(c_int32_t), parameter :: PICTOR_BC_REMV = 5

   type, bind(C) :: pictor_config
      integer(c_int32_t) :: ndim, nx, ny, nz, epc, form, tile, fdtd, geometry, curr_filter
      real(c_double) :: c, c_ompe, g0, mi_me, capacity_factor, dr
   end type pictor_config

   interface
      integer(c_int) function pictor_init(cfg, ctx) bind(C, name="pictor_init")
         import :: c_int, c_ptr, pictor_config
         type(pictor_config), intent(in) :: cfg
         type(c_ptr), intent(out) :: ctx
      end function
      integer(c_int) function pictor_finalize(ctx) bind(C, name="pictor_finalize")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      type(c_ptr) function pictor_last_error() bind(C, name="pictor_last_error")
         import :: c_ptr
      end function
      integer(c_int) function pictor_get_scalars(ctx, out) bind(C, name="pictor_get_scalars")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double) :: out(*)
      end function
      integer(c_int) function pictor_upload_fields(ctx, ex, ey, ez, bx, by, bz) bind(C, name="pictor_upload_fields")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: ex(*), ey(*), ez(*), bx(*), by(*), bz(*)
      end function
      integer(c_int) function pictor_download_fields(ctx, ex, ey, ez, bx, by, bz, jx, jy, jz) &
            bind(C, name="pictor_download_fields")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double) :: ex(*), ey(*), ez(*), bx(*), by(*), bz(*), jx(*), jy(*), jz(*)
      end function
      integer(c_int) function pictor_upload_particles(ctx, n, q, x, y, z, u, v, w, flv) &
            bind(C, name="pictor_upload_particles")
         import :: c_int, c_ptr, c_double, c_int64_t, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: n
         real(c_double), intent(in) :: q(*), x(*), y(*), z(*), u(*), v(*), w(*)
         integer(c_int32_t), intent(in) :: flv(*)
      end function
      integer(c_int) function pictor_particle_count(ctx, n) bind(C, name="pictor_particle_count")
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), intent(out) :: n
      end function
      integer(c_int) function pictor_download_particles(ctx, cap, q, x, y, z, u, v, w, flv, tile, n_out) &
            bind(C, name="pictor_download_particles")
         import :: c_int, c_ptr, c_double, c_int64_t, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: cap
         real(c_double) :: q(*), x(*), y(*), z(*), u(*), v(*), w(*)
         integer(c_int32_t) :: flv(*)
         type(c_ptr), value :: tile
         integer(c_int64_t), intent(out) :: n_out
      end function
      ! the reference's tagp carried with every particle (Cartesian geometry)
      integer(c_int) function pictor_upload_particles_tagged(ctx, n, q, x, y, z, u, v, w, flv, tag) &
            bind(C, name="pictor_upload_particles_tagged")
         import :: c_int, c_ptr, c_double, c_int64_t, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: n
         real(c_double), intent(in) :: q(*), x(*), y(*), z(*), u(*), v(*), w(*)
         integer(c_int32_t), intent(in) :: flv(*), tag(*)
      end function
      integer(c_int) function pictor_download_particles_tagged(ctx, cap, q, x, y, z, u, v, w, flv, tile, tag, &
            n_out) bind(C, name="pictor_download_particles_tagged")
         import :: c_int, c_ptr, c_double, c_int64_t, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), value :: cap
         real(c_double) :: q(*), x(*), y(*), z(*), u(*), v(*), w(*)
         integer(c_int32_t) :: flv(*), tag(*)
         type(c_ptr), value :: tile
         integer(c_int64_t), intent(out) :: n_out
      end function
      integer(c_int) function pictor_update_b_half(ctx) bind(C, name="pictor_update_b_half")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function pictor_move_deposit(ctx) bind(C, name="pictor_move_deposit")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function pictor_exchange_particles(ctx) bind(C, name="pictor_exchange_particles")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function pictor_update_e(ctx) bind(C, name="pictor_update_e")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function pictor_current_halo_add(ctx) bind(C, name="pictor_current_halo_add")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function pictor_smoothen_current(ctx) bind(C, name="pictor_smoothen_current")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      ! push scheme of the cylindrical grid: 0 Boris (cyl_movdep.F90), 1 Strang (cyl_movdep_strang.F90)
      integer(c_int) function pictor_cyl_push_scheme(ctx, scheme) bind(C, name="pictor_cyl_push_scheme")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: scheme
      end function
      integer(c_int) function pictor_cyl_sort_period(ctx, period) bind(C, name="pictor_cyl_sort_period")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: period
      end function
      ! the azimuthal-mode cylinder (PICTOR_GEOM_CYL_MODAL): plasma, laser, moving window
      integer(c_int) function pictor_cyl_plasma(ctx, nr_plasma, k0, nzc, ppc, seed) bind(C, name="pictor_cyl_plasma")
         import :: c_int, c_ptr, c_int32_t, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: nr_plasma, k0, nzc, ppc
         integer(c_int64_t), value :: seed
      end function
      integer(c_int) function pictor_cyl_laser(ctx, a0, omega0, tau, w0, zc) bind(C, name="pictor_cyl_laser")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: a0, omega0, tau, w0, zc
      end function
      integer(c_int) function pictor_moving_window(ctx, on, nr_plasma, ppc, seed) bind(C, name="pictor_moving_window")
         import :: c_int, c_ptr, c_int32_t, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: on, nr_plasma, ppc
         integer(c_int64_t), value :: seed
      end function
      integer(c_int) function pictor_add_current(ctx) bind(C, name="pictor_add_current")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function pictor_field_halo(ctx, which) bind(C, name="pictor_field_halo")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: which
      end function
      integer(c_int) function pictor_reset_current(ctx) bind(C, name="pictor_reset_current")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function pictor_sort_tiles(ctx) bind(C, name="pictor_sort_tiles")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function pictor_step(ctx, nsteps) bind(C, name="pictor_step")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: nsteps
      end function
      integer(c_int) function pictor_synchronize(ctx) bind(C, name="pictor_synchronize")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function pictor_diagnostics(ctx, out) bind(C, name="pictor_diagnostics")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double) :: out(*)
      end function
      integer(c_int) function pictor_setup_weibel(ndim, nx, ny, nz, epc, drift, uth, seed, cap, &
            q, x, y, z, u, v, w, flv, n_out) bind(C, name="pictor_setup_weibel")
         import :: c_int, c_int32_t, c_int64_t, c_double
         integer(c_int32_t), value :: ndim, nx, ny, nz, epc
         real(c_double), value :: drift, uth
         integer(c_int64_t), value :: seed, cap
         real(c_double) :: q(*), x(*), y(*), z(*), u(*), v(*), w(*)
         integer(c_int32_t) :: flv(*)
         integer(c_int64_t), intent(out) :: n_out
      end function
      ! SetBC_Fld / SetBC_Prtl / SetBoundarySpeed (bc.F90:23-63, 210-217); axis 0..2, side 0 lower, 1 upper
      integer(c_int) function pictor_set_bc_fld(ctx, axis, side, type, position, thickness) &
            bind(C, name="pictor_set_bc_fld")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: axis, side, type
         real(c_double), value :: position, thickness
      end function
      integer(c_int) function pictor_set_bc_prtl(ctx, axis, side, type, position) bind(C, name="pictor_set_bc_prtl")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: axis, side, type
         real(c_double), value :: position
      end function
      integer(c_int) function pictor_set_boundary_speed(ctx, axis, side, speed) &
            bind(C, name="pictor_set_boundary_speed")
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: axis, side
         real(c_double), value :: speed
      end function
      ! UpdateMagFld_PML (which = 1) / UpdateElcFld_PML (which = 0), bc_pml.F90
      integer(c_int) function pictor_pml(ctx, which) bind(C, name="pictor_pml")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: which
      end function
      ! EnforceBC_Final, bc.F90:289-340
      integer(c_int) function pictor_enforce_bc_final(ctx) bind(C, name="pictor_enforce_bc_final")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      ! save_total_energy / save_field_collective (savedata.F90:856-890, 356-497), .npz files
      integer(c_int) function pictor_save_energy(ctx, path) bind(C, name="pictor_save_energy")
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: ctx
         character(kind=c_char) :: path(*)
      end function
      integer(c_int) function pictor_save_fields(ctx, path, res, save_current) bind(C, name="pictor_save_fields")
         import :: c_int, c_ptr, c_char, c_int32_t
         type(c_ptr), value :: ctx
         character(kind=c_char) :: path(*)
         integer(c_int32_t), value :: res, save_current
      end function
      ! SaveRestartData / the restart reload (savedata.F90:73-141, reload.F90:17-338)
      integer(c_int) function pictor_save_state(ctx, path, t) bind(C, name="pictor_save_state")
         import :: c_int, c_ptr, c_char, c_int64_t
         type(c_ptr), value :: ctx
         character(kind=c_char) :: path(*)
         integer(c_int64_t), value :: t
      end function
      integer(c_int) function pictor_load_state(ctx, path, t) bind(C, name="pictor_load_state")
         import :: c_int, c_ptr, c_char, c_int64_t
         type(c_ptr), value :: ctx
         character(kind=c_char) :: path(*)
         integer(c_int64_t), intent(out) :: t
      end function
   end interface

contains

   ! Abort(err) of the reference (communication.F90:96-102), with the library's message
   subroutine pictor_check(rc, where)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      character(kind=c_char), pointer :: msg(:)
      integer :: n
      if (rc == PICTOR_OK) return
      call c_f_pointer(pictor_last_error(), msg, [4096])
      n = 0
      do while (n < 4096)
         if (msg(n + 1) == c_null_char) exit
         n = n + 1
      end do
      write (*, '(a,a,a,i0,a)', advance='no') 'pictor error in ', where, ' (', rc, '): '
      write (*, '(4096a)') msg(1:n)
      error stop 1
   end subroutine pictor_check

end module pictor_cabi
