!-----------------------------------------------------------------------
! sphg_mod.f90 -- Fortran host side of the MI355X per-step particle loop.
!
! Module `sphg` keeps the reference's procedure names and call structure
! (sph.f90 -> time_integration -> single_step -> direct_find / link_list /
! con_density / internal force / art_visc, declarations_sph.f90:525-1435 and
! m_particles.f90:938-3988) and forwards every call through ISO_C_BINDING to
! the C ABI in include/sphg.h, which runs the HIP kernels.  The type
! `particles` holds the arrays of the reference's type(particles) that the
! per-step loop reads or writes, in the same layout (x(2,maxn), ...).
!-----------------------------------------------------------------------
module sphg
use iso_c_binding
implicit none

integer(c_int), parameter :: SPHG_WATER = 1, SPHG_SOIL = 2
integer(c_int64_t), parameter :: SPHG_F_ALL = -1_c_int64_t

! include/sphg.h: struct sphg_params
type, bind(C) :: sphg_params_t
   real(c_double) :: water_rho0, water_b, water_gamma, water_c, water_viscosity
   real(c_double) :: soil_rho0, soil_k, soil_E, soil_niu, soil_c
   real(c_double) :: soil_porosity, soil_permeability, soil_cohesion, soil_phi
   integer(c_int32_t) :: nnps, skf
   real(c_double) :: numeric_gravity, alpha, beta, etq, dd, p1, p2, delta, epsilon, gravity
   integer(c_int32_t) :: single_phase, summation_density, average_velocity, visc_artificial
   integer(c_int32_t) :: artificial_density, soil_artificial_stress, volume_fraction
   integer(c_int32_t) :: water_artificial_volume, volume_fraction_renorm
   integer(c_int32_t) :: stress_integration, yield_criterion, plasticity
   real(c_double) :: x_mingeom, x_maxgeom, y_mingeom, y_maxgeom
   integer(c_int32_t) :: maxngx, maxngy
   real(c_double) :: dt
   integer(c_int32_t) :: precision
   integer(c_int32_t) :: integrate_scheme
   integer(c_int32_t) :: reserved(6)
end type

! include/sphg.h: struct sphg_phase_view
type, bind(C) :: sphg_phase_view_t
   integer(c_int32_t) :: ntotal, nvirt, maxn, max_interaction, material, niac, nfail, reserved0
   real(c_double) :: dspp
   type(c_ptr) :: itype, fail, countiac, pair_i, pair_j
   type(c_ptr) :: x, vx, mass, hsml
   type(c_ptr) :: rho, p, c, vof, vof2, dvof, vof_min
   type(c_ptr) :: rho_min, v_min, drho, dvx, av
   type(c_ptr) :: sxx, syy, sxy, txx, tyy, txy, wxy, vcc, epsilon_p
   type(c_ptr) :: dsxx, dsyy, dsxy, dp
   type(c_ptr) :: w, dwdx
   type(c_ptr) :: sxx_min, sxy_min, syy_min, p_min   ! NULL here: the device keeps them
   type(c_ptr) :: dsxx2, dsxy2, dsyy2, dp2
end type

! The per-step subset of the reference's type(particles) (m_particles.f90:62-384).
type :: particles
   integer :: material = SPHG_WATER
   integer :: ntotal = 0, nvirt = 0, maxn = 0, max_interaction = 0
   integer :: niac = 0, nfail = 0
   real(c_double) :: dspp = 0.d0
   integer(c_int32_t), allocatable :: itype(:), fail(:), countiac(:), pair_i(:), pair_j(:)
   real(c_double), allocatable :: x(:,:), vx(:,:), mass(:), hsml(:)
   real(c_double), allocatable :: rho(:), p(:), c(:), vof(:), vof2(:), dvof(:), vof_min(:)
   real(c_double), allocatable :: rho_min(:), v_min(:,:), drho(:), dvx(:,:), av(:,:)
   real(c_double), allocatable :: sxx(:), syy(:), sxy(:), txx(:), tyy(:), txy(:), wxy(:), vcc(:)
   real(c_double), allocatable :: epsilon_p(:), dsxx(:), dsyy(:), dsxy(:), dp(:)
   real(c_double), allocatable :: w(:), dwdx(:,:)
   type(sphg_phase_view_t) :: view
end type

type(c_ptr), save :: ctx = c_null_ptr
type(sphg_params_t), save :: params
integer, save :: itimestep = 0

interface
   integer(c_int) function sphg_create(p, out) bind(C, name='sphg_create')
      import :: c_int, c_ptr, sphg_params_t
      type(sphg_params_t), intent(in) :: p
      type(c_ptr), intent(out) :: out
   end function
   subroutine sphg_destroy(c) bind(C, name='sphg_destroy')
      import :: c_ptr
      type(c_ptr), value :: c
   end subroutine
   integer(c_int) function sphg_bind_phase(c, phase, v) bind(C, name='sphg_bind_phase')
      import :: c_int, c_ptr, sphg_phase_view_t
      type(c_ptr), value :: c
      integer(c_int), value :: phase
      type(sphg_phase_view_t), intent(inout) :: v
   end function
   integer(c_int) function sphg_sync(c, phase, mask, to_host) bind(C, name='sphg_sync')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: c
      integer(c_int), value :: phase, to_host
      integer(c_int64_t), value :: mask
   end function
   integer(c_int) function sphg_set_itimestep(c, it) bind(C, name='sphg_set_itimestep')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: it
   end function
   integer(c_int) function sphg_get_itimestep(c) bind(C, name='sphg_get_itimestep')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
   end function
   integer(c_int) function sphg_direct_find(c, phase) bind(C, name='sphg_direct_find')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: phase
   end function
   integer(c_int) function sphg_link_list(c, phase) bind(C, name='sphg_link_list')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: phase
   end function
   integer(c_int) function sphg_direct_find_2(c) bind(C, name='sphg_direct_find_2')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
   end function
   integer(c_int) function sphg_con_density(c, phase) bind(C, name='sphg_con_density')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: phase
   end function
   integer(c_int) function sphg_delta_sph_density(c, phase) bind(C, name='sphg_delta_sph_density')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: phase
   end function
   integer(c_int) function sphg_sum_density(c, phase) bind(C, name='sphg_sum_density')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: phase
   end function
   integer(c_int) function sphg_internal_force(c, phase) bind(C, name='sphg_internal_force')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: phase
   end function
   integer(c_int) function sphg_art_visc(c, phase) bind(C, name='sphg_art_visc')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: phase
   end function
   integer(c_int) function sphg_single_step(c, sync) bind(C, name='sphg_single_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: sync
   end function
   integer(c_int) function sphg_time_integration(c, nsteps, sync) bind(C, name='sphg_time_integration')
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: nsteps, sync
   end function
   type(c_ptr) function sphg_strerror(code) bind(C, name='sphg_strerror')
      import :: c_int, c_ptr
      integer(c_int), value :: code
   end function
end interface

contains

! Turn a non-zero code into `error stop` with the reference's message.
subroutine check(rc, where)
   integer(c_int), intent(in) :: rc
   character(len=*), intent(in) :: where
   character(kind=c_char), pointer :: msg(:)
   character(len=256) :: text
   integer :: i
   if (rc == 0) return
   call c_f_pointer(sphg_strerror(rc), msg, [256])
   text = ' '
   do i = 1, 256
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
   enddo
   write(*,*) trim(where), ': ', trim(text)
   error stop 1
end subroutine

subroutine allocate_particles(pl, material, n, max_interaction)
   type(particles), intent(inout) :: pl
   integer, intent(in) :: material, n, max_interaction
   integer :: mi
   pl%material = material
   pl%maxn = n
   pl%max_interaction = max_interaction
   mi = max(max_interaction, 1)
   allocate(pl%itype(n), pl%fail(n), pl%countiac(n), pl%pair_i(mi), pl%pair_j(mi))
   allocate(pl%x(2,n), pl%vx(2,n), pl%mass(n), pl%hsml(n))
   allocate(pl%rho(n), pl%p(n), pl%c(n), pl%vof(n), pl%vof2(n), pl%dvof(n), pl%vof_min(n))
   allocate(pl%rho_min(n), pl%v_min(2,n), pl%drho(n), pl%dvx(2,n), pl%av(2,n))
   allocate(pl%sxx(n), pl%syy(n), pl%sxy(n), pl%txx(n), pl%tyy(n), pl%txy(n), pl%wxy(n), pl%vcc(n))
   allocate(pl%epsilon_p(n), pl%dsxx(n), pl%dsyy(n), pl%dsxy(n), pl%dp(n))
   allocate(pl%w(mi), pl%dwdx(2,mi))
   pl%itype = 0; pl%fail = 0; pl%countiac = 0; pl%pair_i = 0; pl%pair_j = 0
   pl%x = 0.d0; pl%vx = 0.d0; pl%mass = 0.d0; pl%hsml = 0.d0
   pl%rho = 0.d0; pl%p = 0.d0; pl%c = 0.d0; pl%vof = 0.d0; pl%vof2 = 0.d0; pl%dvof = 0.d0
   pl%vof_min = 0.d0; pl%rho_min = 0.d0; pl%v_min = 0.d0; pl%drho = 0.d0; pl%dvx = 0.d0
   pl%av = 0.d0; pl%sxx = 0.d0; pl%syy = 0.d0; pl%sxy = 0.d0; pl%txx = 0.d0; pl%tyy = 0.d0
   pl%txy = 0.d0; pl%wxy = 0.d0; pl%vcc = 0.d0; pl%epsilon_p = 0.d0; pl%dsxx = 0.d0
   pl%dsyy = 0.d0; pl%dsxy = 0.d0; pl%dp = 0.d0; pl%w = 0.d0; pl%dwdx = 0.d0
end subroutine

! Pack the arrays of `pl` into the C view (pointers into pl's storage).
subroutine make_view(pl)
   type(particles), intent(inout), target :: pl
   type(sphg_phase_view_t) :: v
   v%ntotal = pl%ntotal; v%nvirt = pl%nvirt; v%maxn = pl%maxn
   v%max_interaction = pl%max_interaction; v%material = pl%material
   v%niac = pl%niac; v%nfail = pl%nfail; v%reserved0 = 0
   v%dspp = pl%dspp
   v%itype = c_loc(pl%itype); v%fail = c_loc(pl%fail); v%countiac = c_loc(pl%countiac)
   v%pair_i = c_loc(pl%pair_i); v%pair_j = c_loc(pl%pair_j)
   v%x = c_loc(pl%x); v%vx = c_loc(pl%vx); v%mass = c_loc(pl%mass); v%hsml = c_loc(pl%hsml)
   v%rho = c_loc(pl%rho); v%p = c_loc(pl%p); v%c = c_loc(pl%c); v%vof = c_loc(pl%vof)
   v%vof2 = c_loc(pl%vof2); v%dvof = c_loc(pl%dvof); v%vof_min = c_loc(pl%vof_min)
   v%rho_min = c_loc(pl%rho_min); v%v_min = c_loc(pl%v_min); v%drho = c_loc(pl%drho)
   v%dvx = c_loc(pl%dvx); v%av = c_loc(pl%av)
   v%sxx = c_loc(pl%sxx); v%syy = c_loc(pl%syy); v%sxy = c_loc(pl%sxy)
   v%txx = c_loc(pl%txx); v%tyy = c_loc(pl%tyy); v%txy = c_loc(pl%txy)
   v%wxy = c_loc(pl%wxy); v%vcc = c_loc(pl%vcc); v%epsilon_p = c_loc(pl%epsilon_p)
   v%dsxx = c_loc(pl%dsxx); v%dsyy = c_loc(pl%dsyy); v%dsxy = c_loc(pl%dsxy); v%dp = c_loc(pl%dp)
   v%w = c_loc(pl%w); v%dwdx = c_loc(pl%dwdx)
   v%sxx_min = c_null_ptr; v%sxy_min = c_null_ptr; v%syy_min = c_null_ptr; v%p_min = c_null_ptr
   v%dsxx2 = c_null_ptr; v%dsxy2 = c_null_ptr; v%dsyy2 = c_null_ptr; v%dp2 = c_null_ptr
   pl%view = v
end subroutine

! allocate_sph (declarations_sph.f90:338) for the device side.
subroutine sphg_init(p)
   type(sphg_params_t), intent(in) :: p
   params = p
   call check(sphg_create(params, ctx), 'sphg_create')
end subroutine

subroutine sphg_bind(pl)
   type(particles), intent(inout), target :: pl
   call make_view(pl)
   call check(sphg_bind_phase(ctx, int(pl%material, c_int), pl%view), 'sphg_bind_phase')
end subroutine

subroutine sphg_finalize()
   call sphg_destroy(ctx)
   ctx = c_null_ptr
end subroutine

subroutine pull(pl, mask)
   type(particles), intent(inout), target :: pl
   integer(c_int64_t), intent(in) :: mask
   call check(sphg_sync(ctx, int(pl%material, c_int), mask, 1_c_int), 'sphg_sync')
end subroutine

! direct_find(parts), m_particles.f90:938-1001
subroutine direct_find(pl)
   type(particles), intent(inout), target :: pl
   call check(sphg_direct_find(ctx, int(pl%material, c_int)), 'direct_find')
   pl%niac = pl%view%niac
end subroutine

! link_list(parts), m_particles.f90:1146-1265
subroutine link_list(pl)
   type(particles), intent(inout), target :: pl
   call check(sphg_link_list(ctx, int(pl%material, c_int)), 'link_list')
   pl%niac = pl%view%niac
end subroutine

! direct_find_2(parts, soil), m_particles.f90:1004-1062
subroutine direct_find_2(water, soil)
   type(particles), intent(inout), target :: water, soil
   call check(sphg_direct_find_2(ctx), 'direct_find_2')
   water%niac = water%view%niac
end subroutine

! con_density(parts), m_particles.f90:2402-2431
subroutine con_density(pl)
   type(particles), intent(inout), target :: pl
   call check(sphg_con_density(ctx, int(pl%material, c_int)), 'con_density')
end subroutine

! delta_sph(pl, pl%rho, pl%drho), m_particles.f90:2658-2695, as single_step
! calls it after con_density (declarations_sph.f90:1207): adds the density
! diffusion term to drho
subroutine delta_sph_density(pl)
   type(particles), intent(inout), target :: pl
   call check(sphg_delta_sph_density(ctx, int(pl%material, c_int)), 'delta_sph')
end subroutine

! sum_density(parts), m_particles.f90:2243-2298 (single_step_for_water,
! declarations_sph.f90:1465): Shepard-normalised summation density into rho
subroutine sum_density(pl)
   type(particles), intent(inout), target :: pl
   call check(sphg_sum_density(ctx, int(pl%material, c_int)), 'sum_density')
end subroutine

! Internal-force group of single_step, declarations_sph.f90:1222-1308
subroutine internal_force(pl)
   type(particles), intent(inout), target :: pl
   call check(sphg_internal_force(ctx, int(pl%material, c_int)), 'internal_force')
   pl%nfail = pl%view%nfail
end subroutine

! art_visc(parts), m_particles.f90:2504-2564
subroutine art_visc(pl)
   type(particles), intent(inout), target :: pl
   call check(sphg_art_visc(ctx, int(pl%material, c_int)), 'art_visc')
end subroutine

! single_step, declarations_sph.f90:1131-1435
subroutine single_step()
   call check(sphg_set_itimestep(ctx, int(itimestep, c_int)), 'sphg_set_itimestep')
   call check(sphg_single_step(ctx, 1_c_int), 'single_step')
end subroutine

! time_integration, declarations_sph.f90:525-791 (resident state; sync at the end)
subroutine time_integration(maxtimestep)
   integer, intent(in) :: maxtimestep
   call check(sphg_set_itimestep(ctx, int(itimestep, c_int)), 'sphg_set_itimestep')
   call check(sphg_time_integration(ctx, int(maxtimestep, c_int), 1_c_int), 'time_integration')
   itimestep = sphg_get_itimestep(ctx)
end subroutine

! Binary state snapshot (sources_amd/snapshot.py) -- read.
subroutine read_state(fname, water, soil, nph)
   character(len=*), intent(in) :: fname
   type(particles), intent(inout) :: water, soil
   integer, intent(out) :: nph
   integer :: u, hdr(4), psize, k, mat, nt, nv, n
   character(kind=c_char), allocatable :: buf(:)
   real(c_double) :: dspp
   open(newunit=u, file=fname, access='stream', form='unformatted', status='old')
   read(u) hdr
   if (hdr(1) /= int(z'53504847') .or. hdr(2) /= 1) error stop 'read_state: not a version-1 snapshot'
   nph = hdr(3); itimestep = hdr(4)
   read(u) psize
   if (psize /= int(c_sizeof(params))) error stop 'read_state: params size mismatch'
   allocate(buf(psize))
   read(u) buf
   params = transfer(buf, params)
   do k = 1, nph
      read(u) mat, nt, nv
      read(u) dspp
      n = nt + nv
      if (k == 1) then
         call read_phase(water)
      else
         call read_phase(soil)
      endif
   enddo
   close(u)
contains
   subroutine read_phase(pl)
      type(particles), intent(inout) :: pl
      call allocate_particles(pl, mat, n, 1)
      pl%ntotal = nt; pl%nvirt = nv; pl%dspp = dspp
      read(u) pl%itype, pl%x, pl%vx
      read(u) pl%mass, pl%hsml, pl%rho, pl%p, pl%c, pl%vof, pl%sxx, pl%sxy, pl%syy, pl%epsilon_p, pl%drho
      read(u) pl%dvx
      read(u) pl%dvof, pl%dsxx, pl%dsxy, pl%dsyy, pl%dp
   end subroutine
end subroutine

! Binary state snapshot -- write.
subroutine write_state(fname, water, soil, nph)
   character(len=*), intent(in) :: fname
   type(particles), intent(in) :: water, soil
   integer, intent(in) :: nph
   integer :: u, k
   character(kind=c_char), allocatable :: buf(:)
   allocate(buf(c_sizeof(params)))
   buf = transfer(params, buf)
   open(newunit=u, file=fname, access='stream', form='unformatted', status='replace')
   write(u) int(z'53504847'), 1, nph, itimestep
   write(u) int(c_sizeof(params))
   write(u) buf
   do k = 1, nph
      if (k == 1) then
         call write_phase(water)
      else
         call write_phase(soil)
      endif
   enddo
   close(u)
contains
   subroutine write_phase(pl)
      type(particles), intent(in) :: pl
      write(u) int(pl%material), pl%ntotal, pl%nvirt
      write(u) pl%dspp
      write(u) pl%itype, pl%x, pl%vx
      write(u) pl%mass, pl%hsml, pl%rho, pl%p, pl%c, pl%vof, pl%sxx, pl%sxy, pl%syy, pl%epsilon_p, pl%drho
      write(u) pl%dvx
      write(u) pl%dvof, pl%dsxx, pl%dsxy, pl%dsyy, pl%dp
   end subroutine
end subroutine

end module sphg
