!! Fortran drop-in for the reference's ocean_redi call surface, backed by the
!! gfx950 kernels of libredi_hip.so.
!!
!! The modules below carry the reference's module, type, component and
!! procedure names -- constants, grid (hgrid_t), ocean_metrics
!! (ocean_metrics_t), multilayer_cgrid_state (multilayer_cgrid_state_t,
!! tracer_t), ocean_boundary_types (ocean_bc_state_t, OBC_WALL), ocean_eos
!! (ocean_eos_t) and ocean_redi (ocean_redi_t with init / destroy /
!! enter_data / exit_data, redi_calc_coeffs, redi_apply_flux,
!! redi_apply_flux_hoist: redi/ocean_redi.F90:71-128, 466-564, 572-607,
!! 927-1002, 1201-1262) -- with the same dummy-argument lists, so a caller
!! written against the reference modules compiles against these unchanged.
!! They hold no numerics: every call goes through the C ABI.
!!
!! The model's arrays are host arrays here, as in the reference's DATA=none
!! lane.  ocean_redi_t keeps device mirrors of h, the tracers, the metrics and
!! the face diffusivities, refreshes them from the host arrays on every call,
!! and after each call copies back what the reference leaves in host memory:
!! the tracers (redi_apply_flux), the public coefficient arrays uPoL..vhEff
!! (redi_calc_coeffs), khtr_u / khtr_v and tr_snap (redi_apply_flux).  A
!! caller that keeps its state on the device calls the C ABI directly.

module constants
   use, intrinsic :: iso_fortran_env, only: real64
   implicit none
   integer, parameter :: wp = real64
   integer, parameter :: NZ_STACK_MAX = 2*256 + 2   ! the library takes nz <= 256
   real(wp), parameter :: GRAVITY = 9.80665_wp
   real(wp), parameter :: H_DIV_EPS = 1.0e-20_wp
end module constants

module grid
   use constants, only: wp
   implicit none
   type :: hgrid_t
      integer :: nx_total = 0, ny_total = 0   ! incl. ghosts
      integer :: nx_phys = 0, ny_phys = 0
      integer :: nghost = 2
      real(wp) :: dx = 0.0_wp, dy = 0.0_wp
   end type hgrid_t
end module grid

module ocean_metrics
   use constants, only: wp
   implicit none
   type :: ocean_metrics_t
      real(wp), allocatable :: dy_cu(:, :), dx_cv(:, :)   ! (nx+1,ny), (nx,ny+1)
      real(wp), allocatable :: idxCu(:, :), idyCv(:, :)   ! (nx+1,ny), (nx,ny+1)
      real(wp), allocatable :: areaT(:, :)                ! (nx,ny)
      real(wp), allocatable :: wet_u(:, :), wet_v(:, :)   ! (nx+1,ny), (nx,ny+1)
   end type ocean_metrics_t
end module ocean_metrics

module multilayer_cgrid_state
   use constants, only: wp
   implicit none
   type :: tracer_t
      real(wp), allocatable :: hTr(:, :, :)   ! (nx, ny, nz), k = 1 at the bed
   end type tracer_t
   type :: multilayer_cgrid_state_t
      integer :: nz_ml = 0
      integer :: idx_temperature = 0, idx_salinity = 0
      real(wp), allocatable :: h_layer(:, :, :)
      type(tracer_t), allocatable :: tracers(:)
   end type multilayer_cgrid_state_t
end module multilayer_cgrid_state

module ocean_boundary_types
   implicit none
   integer, parameter :: OBC_WALL = 1
   type :: ocean_bc_face_tag_t
      integer :: bc_type = OBC_WALL
   end type ocean_bc_face_tag_t
   type :: ocean_bc_state_t
      type(ocean_bc_face_tag_t) :: west, east, south, north
   end type ocean_bc_state_t
end module ocean_boundary_types

module ocean_eos
   use constants, only: wp
   implicit none
   integer, parameter :: EOS_VARIANT_LINEAR = 1, EOS_VARIANT_WRIGHT_97 = 2, &
                         EOS_VARIANT_ROQUET_SPV = 4
   integer, parameter :: TS_POT_PRAC = 1, TS_CONS_ABS = 2
   type :: ocean_eos_t
      logical :: is_init = .false.
      integer :: variant = EOS_VARIANT_LINEAR
      real(wp) :: rho0 = 1035.0_wp, T_ref = 10.0_wp, S_ref = 35.0_wp
      real(wp) :: alpha_T = 1.7e-4_wp, beta_S = 7.6e-4_wp, p_ref = 0.0_wp
      integer :: ts_convention = TS_POT_PRAC
   end type ocean_eos_t
end module ocean_eos

module ocean_redi
   use, intrinsic :: iso_c_binding
   use constants, only: wp, NZ_STACK_MAX
   use grid, only: hgrid_t
   use ocean_metrics, only: ocean_metrics_t
   use multilayer_cgrid_state, only: multilayer_cgrid_state_t
   use ocean_eos, only: ocean_eos_t
   use ocean_boundary_types, only: ocean_bc_state_t, OBC_WALL
   use redi_hip_c
   implicit none
   private
   public :: ocean_redi_t, redi_calc_coeffs, redi_apply_flux, redi_apply_flux_hoist

   type :: ocean_redi_t
      logical :: is_init = .false.
      logical :: enable = .false.
      logical :: continuous = .true.
      real(wp) :: khtr = 0.0_wp
      real(wp) :: ref_pres = -1.0_wp
      integer :: nx_total = 0, ny_total = 0, nz_ml = 0, nsurf = 0
      real(wp), allocatable :: uPoL(:, :, :), uPoR(:, :, :), uhEff(:, :, :)
      integer, allocatable :: uKoL(:, :, :), uKoR(:, :, :)
      real(wp), allocatable :: vPoL(:, :, :), vPoR(:, :, :), vhEff(:, :, :)
      integer, allocatable :: vKoL(:, :, :), vKoR(:, :, :)
      real(wp), allocatable :: khtr_u(:, :), khtr_v(:, :)
      real(wp), allocatable :: tr_snap(:, :, :)
      ! the library's state object and the device mirrors
      type(c_ptr) :: handle = c_null_ptr
      type(c_ptr) :: d_h = c_null_ptr
      type(c_ptr), allocatable :: d_tr(:)
      type(c_ptr) :: d_met(7) = c_null_ptr        ! dy_cu dx_cv idxCu idyCv areaT wet_u wet_v
      type(c_ptr) :: d_ext(2) = c_null_ptr        ! khtr_u_ext, khtr_v_ext
      type(c_ptr) :: d_cf(12) = c_null_ptr        ! uPoL..vhEff, khtr_u, khtr_v exports
   contains
      procedure, non_overridable :: init => ocean_redi_init
      procedure, non_overridable :: destroy => ocean_redi_destroy
      procedure, non_overridable :: enter_data => ocean_redi_enter_data
      procedure, non_overridable :: exit_data => ocean_redi_exit_data
   end type ocean_redi_t

contains

   ! ---- device mirrors ------------------------------------------------------

   subroutine check(rc, what)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: what
      if (rc < 0) then
         write (*, '(3a)') 'ocean_redi (HIP): ', what, ' failed: '//redi_error_message()
         error stop 1
      end if
   end subroutine check

   subroutine hip_check(rc, what)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: what
      if (rc /= 0) then
         write (*, '(3a,i0)') 'ocean_redi (HIP): ', what, ' returned hipError ', rc
         error stop 1
      end if
   end subroutine hip_check

   subroutine dev_alloc(p, bytes)
      type(c_ptr), intent(inout) :: p
      integer(c_size_t), intent(in) :: bytes
      if (c_associated(p)) return
      call hip_check(hipMalloc(p, bytes), 'hipMalloc')
   end subroutine dev_alloc

   subroutine dev_free(p)
      type(c_ptr), intent(inout) :: p
      if (c_associated(p)) call hip_check(hipFree(p), 'hipFree')
      p = c_null_ptr
   end subroutine dev_free

   subroutine put2(d, a)
      type(c_ptr), intent(inout) :: d
      real(wp), intent(in), target, contiguous :: a(:, :)
      integer(c_size_t) :: nb
      nb = int(size(a), c_size_t)*8_c_size_t
      call dev_alloc(d, nb)
      call hip_check(hipMemcpy(d, c_loc(a), nb, HIP_MEMCPY_H2D), 'hipMemcpy H2D')
   end subroutine put2

   subroutine put3(d, a)
      type(c_ptr), intent(inout) :: d
      real(wp), intent(in), target, contiguous :: a(:, :, :)
      integer(c_size_t) :: nb
      nb = int(size(a), c_size_t)*8_c_size_t
      call dev_alloc(d, nb)
      call hip_check(hipMemcpy(d, c_loc(a), nb, HIP_MEMCPY_H2D), 'hipMemcpy H2D')
   end subroutine put3

   subroutine get_bytes(hostp, d, nb)
      type(c_ptr), intent(in) :: hostp, d
      integer(c_size_t), intent(in) :: nb
      call hip_check(hipMemcpy(hostp, d, nb, HIP_MEMCPY_D2H), 'hipMemcpy D2H')
   end subroutine get_bytes

   ! The C views of the model's types, with the device mirrors refreshed.
   subroutine stage(this, grid, metrics, ms, g, m, s)
      type(ocean_redi_t), intent(inout), target :: this
      type(hgrid_t), intent(in) :: grid
      type(ocean_metrics_t), intent(in) :: metrics
      type(multilayer_cgrid_state_t), intent(in) :: ms
      type(redi_hgrid_c), intent(out) :: g
      type(redi_metrics_c), intent(out) :: m
      type(redi_state_c), intent(out) :: s
      integer :: it, nt
      g = redi_hgrid_c(grid%nx_total, grid%ny_total, grid%nx_phys, grid%ny_phys, grid%nghost, 0, &
                       grid%dx, grid%dy)
      call put2(this%d_met(1), metrics%dy_cu)
      call put2(this%d_met(2), metrics%dx_cv)
      call put2(this%d_met(3), metrics%idxCu)
      call put2(this%d_met(4), metrics%idyCv)
      call put2(this%d_met(5), metrics%areaT)
      call put2(this%d_met(6), metrics%wet_u)
      call put2(this%d_met(7), metrics%wet_v)
      m = redi_metrics_c(this%d_met(1), this%d_met(2), this%d_met(3), this%d_met(4), &
                         this%d_met(5), this%d_met(6), this%d_met(7))
      nt = 0
      if (allocated(ms%tracers)) nt = size(ms%tracers)
      if (allocated(this%d_tr)) then
         if (size(this%d_tr) /= nt) then
            do it = 1, size(this%d_tr)
               call dev_free(this%d_tr(it))
            end do
            deallocate (this%d_tr)
         end if
      end if
      if (.not. allocated(this%d_tr)) then
         allocate (this%d_tr(nt))
         this%d_tr = c_null_ptr
      end if
      s%nz_ml = ms%nz_ml
      s%idx_temperature = ms%idx_temperature
      s%idx_salinity = ms%idx_salinity
      s%ntracers = nt
      s%h_layer = c_null_ptr
      if (allocated(ms%h_layer)) then
         call put3(this%d_h, ms%h_layer)
         s%h_layer = this%d_h
      end if
      do it = 1, nt
         call put3(this%d_tr(it), ms%tracers(it)%hTr)
      end do
      s%tracers = c_null_ptr
      if (nt > 0) s%tracers = c_loc(this%d_tr)
   end subroutine stage

   ! ---- ocean_redi_t procedures (redi/ocean_redi.F90:466-564) -----------------

   subroutine ocean_redi_init(this, grid, nz_ml)
      class(ocean_redi_t), intent(inout) :: this
      type(hgrid_t), intent(in) :: grid
      integer, intent(in) :: nz_ml
      type(redi_hgrid_c) :: g
      integer :: nx, ny, nz, ns
      if (this%is_init) call this%destroy()
      nx = grid%nx_total; ny = grid%ny_total; nz = max(nz_ml, 1); ns = 2*nz + 2
      if (ns > NZ_STACK_MAX) then
         write (*, '(a,i0,a,i0)') 'ocean_redi_init: 2*nz_ml+2 = ', ns, ' exceeds NZ_STACK_MAX = ', &
            NZ_STACK_MAX
         error stop 1
      end if
      g = redi_hgrid_c(nx, ny, grid%nx_phys, grid%ny_phys, grid%nghost, 0, grid%dx, grid%dy)
      call check(redi_init_c(this%handle, g, int(nz, c_int)), 'redi_init')
      this%nx_total = nx; this%ny_total = ny; this%nz_ml = nz; this%nsurf = ns
      allocate (this%uPoL(nx + 1, ny, ns), source=0.0_wp)
      allocate (this%uPoR(nx + 1, ny, ns), source=0.0_wp)
      allocate (this%uKoL(nx + 1, ny, ns), source=1)
      allocate (this%uKoR(nx + 1, ny, ns), source=1)
      allocate (this%uhEff(nx + 1, ny, ns - 1), source=0.0_wp)
      allocate (this%vPoL(nx, ny + 1, ns), source=0.0_wp)
      allocate (this%vPoR(nx, ny + 1, ns), source=0.0_wp)
      allocate (this%vKoL(nx, ny + 1, ns), source=1)
      allocate (this%vKoR(nx, ny + 1, ns), source=1)
      allocate (this%vhEff(nx, ny + 1, ns - 1), source=0.0_wp)
      allocate (this%khtr_u(nx + 1, ny), source=0.0_wp)
      allocate (this%khtr_v(nx, ny + 1), source=0.0_wp)
      allocate (this%tr_snap(nx, ny, nz), source=0.0_wp)
      this%is_init = .true.
   end subroutine ocean_redi_init

   subroutine ocean_redi_destroy(this)
      class(ocean_redi_t), intent(inout) :: this
      integer :: k
      if (c_associated(this%handle)) call check(redi_destroy_c(this%handle), 'redi_destroy')
      this%handle = c_null_ptr
      call dev_free(this%d_h)
      if (allocated(this%d_tr)) then
         do k = 1, size(this%d_tr)
            call dev_free(this%d_tr(k))
         end do
         deallocate (this%d_tr)
      end if
      do k = 1, 7
         call dev_free(this%d_met(k))
      end do
      do k = 1, 2
         call dev_free(this%d_ext(k))
      end do
      do k = 1, 12
         call dev_free(this%d_cf(k))
      end do
      if (allocated(this%uPoL)) deallocate (this%uPoL, this%uPoR, this%uKoL, this%uKoR, this%uhEff, &
                                            this%vPoL, this%vPoR, this%vKoL, this%vKoR, this%vhEff, &
                                            this%khtr_u, this%khtr_v, this%tr_snap)
      this%is_init = .false.
   end subroutine ocean_redi_destroy

   ! Device residency is the library's; these stay for source compatibility.
   subroutine ocean_redi_enter_data(this)
      class(ocean_redi_t), intent(inout) :: this
   end subroutine ocean_redi_enter_data

   subroutine ocean_redi_exit_data(this)
      class(ocean_redi_t), intent(inout) :: this
   end subroutine ocean_redi_exit_data

   subroutine configure(this)
      type(ocean_redi_t), intent(in) :: this
      call check(redi_configure_c(this%handle, merge(1_c_int, 0_c_int, this%enable), this%khtr, &
                                  merge(1_c_int, 0_c_int, this%continuous), this%ref_pres), &
                 'redi_configure')
   end subroutine configure

   ! ---- Phase A (redi/ocean_redi.F90:572-607) --------------------------------

   subroutine redi_calc_coeffs(grid, metrics, eos, this, ms)
      type(hgrid_t), intent(in) :: grid
      type(ocean_metrics_t), intent(in) :: metrics
      type(ocean_eos_t), intent(in) :: eos
      type(ocean_redi_t), intent(inout), target :: this
      type(multilayer_cgrid_state_t), intent(in) :: ms
      type(redi_hgrid_c) :: g
      type(redi_metrics_c) :: m
      type(redi_state_c) :: s
      type(redi_eos_c) :: e
      integer(c_int) :: rc
      integer(c_size_t) :: nu, nv, ns
      if (.not. this%is_init .or. .not. this%enable) return
      call configure(this)
      call stage(this, grid, metrics, ms, g, m, s)
      e = redi_eos_c(eos%variant, eos%ts_convention, eos%rho0, eos%T_ref, eos%S_ref, eos%alpha_T, &
                     eos%beta_S, eos%p_ref)
      rc = redi_calc_coeffs_c(g, m, e, this%handle, s, c_null_ptr)
      call check(rc, 'redi_calc_coeffs')
      if (rc == REDI_SKIPPED) return
      ! the public coefficient arrays, in the reference layout
      call export_state(this, nu, nv, ns)
      call get_bytes(c_loc(this%uPoL), this%d_cf(1), nu*ns*8)
      call get_bytes(c_loc(this%uPoR), this%d_cf(2), nu*ns*8)
      call get_bytes(c_loc(this%uKoL), this%d_cf(3), nu*ns*4)
      call get_bytes(c_loc(this%uKoR), this%d_cf(4), nu*ns*4)
      call get_bytes(c_loc(this%uhEff), this%d_cf(5), nu*(ns - 1)*8)
      call get_bytes(c_loc(this%vPoL), this%d_cf(6), nv*ns*8)
      call get_bytes(c_loc(this%vPoR), this%d_cf(7), nv*ns*8)
      call get_bytes(c_loc(this%vKoL), this%d_cf(8), nv*ns*4)
      call get_bytes(c_loc(this%vKoR), this%d_cf(9), nv*ns*4)
      call get_bytes(c_loc(this%vhEff), this%d_cf(10), nv*(ns - 1)*8)
   end subroutine redi_calc_coeffs

   ! ---- Phase B (redi/ocean_redi.F90:927-1002, 1201-1262) ----------------------

   subroutine apply(grid, metrics, this, ms, dt, khtr_u_ext, khtr_v_ext, bc, hoist)
      type(hgrid_t), intent(in) :: grid
      type(ocean_metrics_t), intent(in) :: metrics
      type(ocean_redi_t), intent(inout), target :: this
      type(multilayer_cgrid_state_t), intent(inout), target :: ms
      real(wp), intent(in) :: dt
      real(wp), intent(in), optional :: khtr_u_ext(:, :), khtr_v_ext(:, :)
      type(ocean_bc_state_t), intent(in), optional :: bc
      logical, intent(in) :: hoist
      type(redi_hgrid_c) :: g
      type(redi_metrics_c) :: m
      type(redi_state_c) :: s
      type(redi_bc_c), target :: b
      type(c_ptr) :: kue, kve, bp
      integer(c_int) :: rc
      integer :: it, nt
      integer(c_size_t) :: nb, nu, nv
      if (.not. this%is_init .or. .not. this%enable) return
      if (.not. allocated(ms%tracers)) return
      nt = size(ms%tracers)
      if (nt < 1) return
      call configure(this)
      call stage(this, grid, metrics, ms, g, m, s)
      kue = c_null_ptr; kve = c_null_ptr
      if (present(khtr_u_ext) .and. present(khtr_v_ext)) then
         call put2(this%d_ext(1), khtr_u_ext)
         call put2(this%d_ext(2), khtr_v_ext)
         kue = this%d_ext(1); kve = this%d_ext(2)
      end if
      bp = c_null_ptr
      if (present(bc)) then
         b = redi_bc_c(merge(1, 0, bc%west%bc_type == OBC_WALL), &
                       merge(1, 0, bc%east%bc_type == OBC_WALL), &
                       merge(1, 0, bc%south%bc_type == OBC_WALL), &
                       merge(1, 0, bc%north%bc_type == OBC_WALL))
         bp = c_loc(b)
      end if
      if (hoist) then
         rc = redi_apply_flux_hoist_c(g, m, this%handle, s, dt, kue, kve, bp, c_null_ptr)
      else
         rc = redi_apply_flux_c(g, m, this%handle, s, dt, kue, kve, bp, c_null_ptr)
      end if
      call check(rc, 'redi_apply_flux')
      if (rc == REDI_SKIPPED) return
      this%tr_snap = ms%tracers(nt)%hTr   ! the last tracer's pre-step state, as the reference leaves it
      do it = 1, nt
         nb = int(size(ms%tracers(it)%hTr), c_size_t)*8_c_size_t
         call get_bytes(c_loc(ms%tracers(it)%hTr), this%d_tr(it), nb)
      end do
      ! the resolved face diffusivities khtr_u / khtr_v
      call export_state(this, nu, nv, nb)
      call get_bytes(c_loc(this%khtr_u), this%d_cf(11), nu*8)
      call get_bytes(c_loc(this%khtr_v), this%d_cf(12), nv*8)
   end subroutine apply

   ! The library's Phase-A state and face diffusivities exported into device
   ! scratch of the reference shapes (d_cf(1:10): uPoL..vhEff, d_cf(11:12):
   ! khtr_u, khtr_v); nu / nv faces per surface, ns surfaces.
   subroutine export_state(this, nu, nv, ns)
      type(ocean_redi_t), intent(inout), target :: this
      integer(c_size_t), intent(out) :: nu, nv, ns
      type(redi_coeffs_c) :: cf
      ns = int(this%nsurf, c_size_t)
      nu = int(this%nx_total + 1, c_size_t)*int(this%ny_total, c_size_t)
      nv = int(this%nx_total, c_size_t)*int(this%ny_total + 1, c_size_t)
      call dev_alloc(this%d_cf(1), nu*ns*8); call dev_alloc(this%d_cf(2), nu*ns*8)
      call dev_alloc(this%d_cf(3), nu*ns*4); call dev_alloc(this%d_cf(4), nu*ns*4)
      call dev_alloc(this%d_cf(5), nu*(ns - 1)*8)
      call dev_alloc(this%d_cf(6), nv*ns*8); call dev_alloc(this%d_cf(7), nv*ns*8)
      call dev_alloc(this%d_cf(8), nv*ns*4); call dev_alloc(this%d_cf(9), nv*ns*4)
      call dev_alloc(this%d_cf(10), nv*(ns - 1)*8)
      call dev_alloc(this%d_cf(11), nu*8); call dev_alloc(this%d_cf(12), nv*8)
      cf = redi_coeffs_c(this%d_cf(1), this%d_cf(2), this%d_cf(3), this%d_cf(4), this%d_cf(5), &
                         this%d_cf(6), this%d_cf(7), this%d_cf(8), this%d_cf(9), this%d_cf(10), &
                         this%d_cf(11), this%d_cf(12))
      call check(redi_export_coeffs_c(this%handle, cf, c_null_ptr), 'redi_export_coeffs')
   end subroutine export_state

   subroutine redi_apply_flux(grid, metrics, this, ms, dt, khtr_u_ext, khtr_v_ext, bc)
      type(hgrid_t), intent(in) :: grid
      type(ocean_metrics_t), intent(in) :: metrics
      type(ocean_redi_t), intent(inout), target :: this
      type(multilayer_cgrid_state_t), intent(inout), target :: ms
      real(wp), intent(in) :: dt
      real(wp), intent(in), optional :: khtr_u_ext(:, :), khtr_v_ext(:, :)
      type(ocean_bc_state_t), intent(in), optional :: bc
      call apply(grid, metrics, this, ms, dt, khtr_u_ext, khtr_v_ext, bc, .false.)
   end subroutine redi_apply_flux

   subroutine redi_apply_flux_hoist(grid, metrics, this, ms, dt, khtr_u_ext, khtr_v_ext, bc)
      type(hgrid_t), intent(in) :: grid
      type(ocean_metrics_t), intent(in) :: metrics
      type(ocean_redi_t), intent(inout), target :: this
      type(multilayer_cgrid_state_t), intent(inout), target :: ms
      real(wp), intent(in) :: dt
      real(wp), intent(in), optional :: khtr_u_ext(:, :), khtr_v_ext(:, :)
      type(ocean_bc_state_t), intent(in), optional :: bc
      call apply(grid, metrics, this, ms, dt, khtr_u_ext, khtr_v_ext, bc, .true.)
   end subroutine redi_apply_flux_hoist

end module ocean_redi
