!! redi_hip_c -- iso_c_binding interfaces to libredi_hip.so (include/redi_hip.h)
!! and the few HIP runtime calls the Fortran drop-in modules need.
!!
!! The C structs are mirrored as bind(C) types with the same member order;
!! device arrays travel as type(c_ptr).  Every entry returns the library's
!! status code (REDI_OK = 0, REDI_SKIPPED = 1, negative = error).
module redi_hip_c
   use, intrinsic :: iso_c_binding
   implicit none
   public

   integer(c_int), parameter :: REDI_OK = 0, REDI_SKIPPED = 1
   integer(c_int), parameter :: HIP_MEMCPY_H2D = 1, HIP_MEMCPY_D2H = 2

   type, bind(C) :: redi_hgrid_c          ! redi_hgrid_t
      integer(c_int) :: nx_total, ny_total, nx_phys, ny_phys, nghost, pad_
      real(c_double) :: dx, dy
   end type

   type, bind(C) :: redi_metrics_c        ! redi_metrics_t (device pointers)
      type(c_ptr) :: dy_cu, dx_cv, idxCu, idyCv, areaT, wet_u, wet_v
   end type

   type, bind(C) :: redi_eos_c            ! redi_eos_t
      integer(c_int) :: variant, ts_convention
      real(c_double) :: rho0, T_ref, S_ref, alpha_T, beta_S, p_ref
   end type

   type, bind(C) :: redi_state_c          ! redi_state_t
      integer(c_int) :: nz_ml, idx_temperature, idx_salinity, ntracers
      type(c_ptr) :: h_layer               ! device
      type(c_ptr) :: tracers               ! host array of ntracers device pointers
   end type

   type, bind(C) :: redi_bc_c             ! redi_bc_t
      integer(c_int) :: west, east, south, north
   end type

   type, bind(C) :: redi_coeffs_c         ! redi_coeffs_t (device pointers)
      type(c_ptr) :: uPoL, uPoR, uKoL, uKoR, uhEff, vPoL, vPoR, vKoL, vKoR, vhEff, &
                     khtr_u, khtr_v
   end type

   interface
      integer(c_int) function redi_init_c(self, grid, nz_ml) bind(C, name="redi_init")
         import :: c_int, c_ptr, redi_hgrid_c
         type(c_ptr), intent(out) :: self
         type(redi_hgrid_c), intent(in) :: grid
         integer(c_int), value :: nz_ml
      end function

      integer(c_int) function redi_destroy_c(self) bind(C, name="redi_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: self
      end function

      integer(c_int) function redi_configure_c(self, enable, khtr, continuous, ref_pres) &
         bind(C, name="redi_configure")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: self
         integer(c_int), value :: enable, continuous
         real(c_double), value :: khtr, ref_pres
      end function

      integer(c_int) function redi_calc_coeffs_c(grid, metrics, eos, self, ms, stream) &
         bind(C, name="redi_calc_coeffs")
         import :: c_int, c_ptr, redi_hgrid_c, redi_metrics_c, redi_eos_c, redi_state_c
         type(redi_hgrid_c), intent(in) :: grid
         type(redi_metrics_c), intent(in) :: metrics
         type(redi_eos_c), intent(in) :: eos
         type(c_ptr), value :: self
         type(redi_state_c), intent(in) :: ms
         type(c_ptr), value :: stream
      end function

      integer(c_int) function redi_apply_flux_c(grid, metrics, self, ms, dt, kue, kve, bc, stream) &
         bind(C, name="redi_apply_flux")
         import :: c_int, c_ptr, c_double, redi_hgrid_c, redi_metrics_c, redi_state_c
         type(redi_hgrid_c), intent(in) :: grid
         type(redi_metrics_c), intent(in) :: metrics
         type(c_ptr), value :: self
         type(redi_state_c), intent(in) :: ms
         real(c_double), value :: dt
         type(c_ptr), value :: kue, kve, bc, stream
      end function

      integer(c_int) function redi_apply_flux_hoist_c(grid, metrics, self, ms, dt, kue, kve, bc, &
                                                      stream) bind(C, name="redi_apply_flux_hoist")
         import :: c_int, c_ptr, c_double, redi_hgrid_c, redi_metrics_c, redi_state_c
         type(redi_hgrid_c), intent(in) :: grid
         type(redi_metrics_c), intent(in) :: metrics
         type(c_ptr), value :: self
         type(redi_state_c), intent(in) :: ms
         real(c_double), value :: dt
         type(c_ptr), value :: kue, kve, bc, stream
      end function

      integer(c_int) function redi_export_coeffs_c(self, dst, stream) &
         bind(C, name="redi_export_coeffs")
         import :: c_int, c_ptr, redi_coeffs_c
         type(c_ptr), value :: self
         type(redi_coeffs_c), intent(in) :: dst
         type(c_ptr), value :: stream
      end function

      type(c_ptr) function redi_last_error_c() bind(C, name="redi_last_error")
         import :: c_ptr
      end function

      ! HIP runtime (libamdhip64)
      integer(c_int) function hipMalloc(ptr, bytes) bind(C, name="hipMalloc")
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), intent(out) :: ptr
         integer(c_size_t), value :: bytes
      end function

      integer(c_int) function hipFree(ptr) bind(C, name="hipFree")
         import :: c_int, c_ptr
         type(c_ptr), value :: ptr
      end function

      integer(c_int) function hipMemcpy(dst, src, bytes, kind) bind(C, name="hipMemcpy")
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), value :: dst, src
         integer(c_size_t), value :: bytes
         integer(c_int), value :: kind
      end function

      integer(c_int) function hipDeviceSynchronize() bind(C, name="hipDeviceSynchronize")
         import :: c_int
      end function

      integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
         import :: c_size_t, c_ptr
         type(c_ptr), value :: s
      end function
   end interface

contains

   !> The library's last error message.
   function redi_error_message() result(msg)
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: n, i
      p = redi_last_error_c()
      n = int(c_strlen(p))
      call c_f_pointer(p, chars, [max(n, 1)])
      allocate (character(len=n) :: msg)
      do i = 1, n
         msg(i:i) = chars(i)
      end do
   end function

end module redi_hip_c
