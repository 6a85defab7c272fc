!! Oracle driver for the reference Redi kernel (test infrastructure only).
!!
!! This program is written for this repository; it is NOT a reference file.
!! It `use`s the reference's own modules (constants, grid, ocean_metrics,
!! multilayer_cgrid_state, ocean_boundary_types, ocean_eos, ocean_redi),
!! compiled from their sources under the reference tree by oracle/Makefile,
!! and runs them on a seeded input file instead of the reference drivers'
!! built-in initial condition (redi/drivers/dc_main.F90:172-206).
!!
!! Input file ("REDI" v1, little-endian stream, arrays column-major):
!!   int32  magic(=0x49444552 'REDI'), version(=1)
!!   int32  nx_phys, ny_phys, nz, nghost, nt, idx_T, idx_S, eos_variant,
!!          nsteps, wall_w, wall_e, wall_s, wall_n, use_ext, enable, pad
!!   real64 dt, khtr, rho0, T_ref, S_ref, alpha_T, beta_S, p_ref
!!   real64 h_layer(nx,ny,nz), hTr(nx,ny,nz) x nt,
!!          dy_cu(nx+1,ny), dx_cv(nx,ny+1), idxCu(nx+1,ny), idyCv(nx,ny+1),
!!          areaT(nx,ny), wet_u(nx+1,ny), wet_v(nx,ny+1),
!!          khtr_u_ext(nx+1,ny), khtr_v_ext(nx,ny+1)
!! Output file ("REDO" v1):
!!   int32  magic(=0x4F444552 'REDO'), version, nx, ny, nz, nt, ns, nsteps
!!   after step 1: uPoL,uPoR (real64), uKoL,uKoR (int32), uhEff, then the
!!                 same five v arrays, then hTr x nt
!!   after the last step: hTr x nt
!!
!! Usage: redi_ref <input> <output|-> [mode] [nwarm]
!!   mode 0 = redi_apply_flux (ocean_redi.F90:927), 1 = redi_apply_flux_hoist
!!   (ocean_redi.F90:1201).  nwarm > 0 runs that many extra untimed steps on a
!!   copy of the state first and then prints the timed ms/step of nsteps
!!   (the outputs are always those of a cold start).
program redi_ref_driver
   use, intrinsic :: iso_fortran_env, only: int32, int64, real64, output_unit
   use constants, only: wp, NZ_STACK_MAX
   use grid, only: hgrid_t
   use ocean_metrics, only: ocean_metrics_t
   use multilayer_cgrid_state, only: multilayer_cgrid_state_t
   use ocean_eos, only: ocean_eos_t
   use ocean_boundary_types, only: ocean_bc_state_t, OBC_WALL
   use ocean_redi, only: ocean_redi_t, redi_calc_coeffs, redi_apply_flux, &
                         redi_apply_flux_hoist
   implicit none

   type(hgrid_t) :: grid
   type(ocean_metrics_t) :: metrics
   type(ocean_eos_t) :: eos
   type(ocean_bc_state_t) :: bc
   type(multilayer_cgrid_state_t) :: ms
   type(ocean_redi_t) :: redi
   real(wp), allocatable :: khtr_u_ext(:, :), khtr_v_ext(:, :)
   real(wp), allocatable :: hTr0(:, :, :, :)
   integer(int32) :: hdr(16), magic, version
   real(real64) :: rpar(8)
   integer :: nx, ny, nz, nt, nsteps, mode, nwarm, it, step, iu, ios, ns
   logical :: use_ext
   character(len=1024) :: in_path, out_path, buf
   integer(int64) :: c0, c1, crate
   real(wp) :: ms_step

   if (command_argument_count() < 2) then
      write (output_unit, '(a)') 'usage: redi_ref <input> <output|-> [mode] [nwarm]'
      stop 1
   end if
   call get_command_argument(1, in_path)
   call get_command_argument(2, out_path)
   mode = 0
   nwarm = 0
   if (command_argument_count() >= 3) then
      call get_command_argument(3, buf); read (buf, *) mode
   end if
   if (command_argument_count() >= 4) then
      call get_command_argument(4, buf); read (buf, *) nwarm
   end if

   open (newunit=iu, file=trim(in_path), access='stream', form='unformatted', &
         status='old', iostat=ios)
   if (ios /= 0) then
      write (output_unit, '(2a)') 'ERROR: cannot open ', trim(in_path); stop 1
   end if
   read (iu) magic, version
   if (magic /= int(z'49444552', int32) .or. version /= 1) then
      write (output_unit, '(a)') 'ERROR: bad input magic/version'; stop 1
   end if
   read (iu) hdr
   read (iu) rpar

   grid%nx_phys = hdr(1); grid%ny_phys = hdr(2); nz = hdr(3)
   grid%nghost = hdr(4); nt = hdr(5)
   nx = grid%nx_phys + 2*grid%nghost
   ny = grid%ny_phys + 2*grid%nghost
   grid%nx_total = nx; grid%ny_total = ny
   nsteps = hdr(9)
   if (2*nz + 2 > NZ_STACK_MAX) then
      write (output_unit, '(a)') 'ERROR: 2*nz+2 exceeds NZ_STACK_MAX'; stop 1
   end if

   eos%variant = hdr(8)
   eos%rho0 = rpar(3); eos%T_ref = rpar(4); eos%S_ref = rpar(5)
   eos%alpha_T = rpar(6); eos%beta_S = rpar(7); eos%p_ref = rpar(8)

   bc%west%bc_type = merge(OBC_WALL, 0, hdr(10) /= 0)
   bc%east%bc_type = merge(OBC_WALL, 0, hdr(11) /= 0)
   bc%south%bc_type = merge(OBC_WALL, 0, hdr(12) /= 0)
   bc%north%bc_type = merge(OBC_WALL, 0, hdr(13) /= 0)
   use_ext = hdr(14) /= 0

   ms%nz_ml = nz
   ms%idx_temperature = hdr(6)
   ms%idx_salinity = hdr(7)
   allocate (ms%h_layer(nx, ny, nz))
   read (iu) ms%h_layer
   allocate (ms%tracers(nt))
   do it = 1, nt
      allocate (ms%tracers(it)%hTr(nx, ny, nz))
      read (iu) ms%tracers(it)%hTr
   end do
   allocate (metrics%dy_cu(nx + 1, ny), metrics%dx_cv(nx, ny + 1))
   allocate (metrics%idxCu(nx + 1, ny), metrics%idyCv(nx, ny + 1))
   allocate (metrics%areaT(nx, ny))
   allocate (metrics%wet_u(nx + 1, ny), metrics%wet_v(nx, ny + 1))
   allocate (khtr_u_ext(nx + 1, ny), khtr_v_ext(nx, ny + 1))
   read (iu) metrics%dy_cu, metrics%dx_cv, metrics%idxCu, metrics%idyCv
   read (iu) metrics%areaT, metrics%wet_u, metrics%wet_v
   read (iu) khtr_u_ext, khtr_v_ext
   close (iu)

   call redi%init(grid, nz)
   redi%enable = hdr(15) /= 0
   redi%khtr = rpar(2)
   ns = redi%nsurf

   if (nwarm > 0) then
      ! Timing lane: warm up on the input state, time nsteps, then restore
      ! the cold-start state so the dumped outputs do not depend on nwarm.
      allocate (hTr0(nx, ny, nz, nt))
      do it = 1, nt
         hTr0(:, :, :, it) = ms%tracers(it)%hTr
      end do
      do step = 1, nwarm
         call one_step()
      end do
      call system_clock(c0, crate)
      do step = 1, nsteps
         call one_step()
      end do
      call system_clock(c1)
      ms_step = real(c1 - c0, wp)*1000.0_wp/real(crate, wp)/real(max(nsteps, 1), wp)
      write (output_unit, '(a,f14.4,a,i0,a,i0,a,i0,a,i0)') 'REF_TIMING ms_per_step=', ms_step, &
         ' steps=', nsteps, ' nx=', nx, ' ny=', ny, ' nz=', nz
      do it = 1, nt
         ms%tracers(it)%hTr = hTr0(:, :, :, it)
      end do
      redi%uPoL = 0.0_wp; redi%uPoR = 0.0_wp; redi%uKoL = 1; redi%uKoR = 1
      redi%uhEff = 0.0_wp; redi%vPoL = 0.0_wp; redi%vPoR = 0.0_wp
      redi%vKoL = 1; redi%vKoR = 1; redi%vhEff = 0.0_wp
      if (trim(out_path) == '-') stop 0
   end if

   if (trim(out_path) /= '-') then
      open (newunit=iu, file=trim(out_path), access='stream', form='unformatted', &
            status='replace')
      write (iu) int(z'4F444552', int32), 1_int32, int(nx, int32), int(ny, int32), &
         int(nz, int32), int(nt, int32), int(ns, int32), int(nsteps, int32)
   end if
   do step = 1, nsteps
      call one_step()
      if (step == 1 .and. trim(out_path) /= '-') then
         write (iu) redi%uPoL, redi%uPoR, redi%uKoL, redi%uKoR, redi%uhEff
         write (iu) redi%vPoL, redi%vPoR, redi%vKoL, redi%vKoR, redi%vhEff
         do it = 1, nt
            write (iu) ms%tracers(it)%hTr
         end do
      end if
   end do
   if (trim(out_path) /= '-') then
      do it = 1, nt
         write (iu) ms%tracers(it)%hTr
      end do
      close (iu)
   end if

contains

   subroutine one_step()
      call redi_calc_coeffs(grid, metrics, eos, redi, ms)
      if (mode == 1) then
         if (use_ext) then
            call redi_apply_flux_hoist(grid, metrics, redi, ms, rpar(1), khtr_u_ext, khtr_v_ext, bc)
         else
            call redi_apply_flux_hoist(grid, metrics, redi, ms, rpar(1), bc=bc)
         end if
      else
         if (use_ext) then
            call redi_apply_flux(grid, metrics, redi, ms, rpar(1), khtr_u_ext, khtr_v_ext, bc)
         else
            call redi_apply_flux(grid, metrics, redi, ms, rpar(1), bc=bc)
         end if
      end if
   end subroutine one_step

end program redi_ref_driver
