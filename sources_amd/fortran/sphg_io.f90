!-----------------------------------------------------------------------
! sphg_io.f90 -- host-side front and back end of the Fortran driver:
!   read_cas / treat_cmd   the steering file (declarations_sph.f90:61-335):
!                          records of `NAME = value[, value ...]`, names
!                          case-insensitive, '#' comments, a trailing ','
!                          continues a command on the next record, a record
!                          starting with '&' ends the file;
!   output                 Tecplot BLOCK zones of both phases
!                          (declarations_sph.f90:1703-1810);
!   interaction_statistics / minimum_time_step
!                          the reference's diagnostics (m_particles.f90:555-630).
! The numerical keywords update a set of overrides that apply_cas copies
! into the sphg_params of the run (after the initial state is read).
!-----------------------------------------------------------------------
module sphg_io
use iso_c_binding
use sphg
implicit none

integer, save :: print_step = 100, save_step = 500, save_step_from = 0
integer, save :: integrate_scheme = 1
character(len=256), save :: results = 'f_xv.dat', res_soil = 'f_state.dat', res_other = 'f_other.dat'
character(len=32), save :: first_phase = 'water', second_phase = 'soil'
character(len=32), save :: particle_size = '', operating_mode = ''
logical, save :: dry_soil = .false., with_nozzle = .false., single_phase_cas = .false.
real(c_double), save :: inlet_velocity = 0.d0, time = 0.d0
integer, save :: f_xv = -1, f_state = -1, f_other = -1

! overrides of sphg_params fields read from the steering file
logical, save :: set_nnps = .false., set_water = .false., set_soil = .false., set_darcy = .false.
logical, save :: set_failure = .false., set_dt = .false., set_single = .false.
integer, save :: cas_nnps = 1
real(c_double), save :: cas_water(5) = 0.d0, cas_soil(5) = 0.d0, cas_darcy(2) = 0.d0
real(c_double), save :: cas_failure(2) = 0.d0, cas_dt = 0.d0

contains

subroutine read_cas(cas_file)
   character(len=*), intent(in) :: cas_file
   character(len=4096) :: cmd
   character(len=1024) :: record
   integer :: u, stat, lenrc
   logical :: alive
   inquire(file=cas_file, exist=alive)
   if (.not. alive) then
      write(*,*) "Steering file ", trim(cas_file), " doesn't exist."
      stop
   endif
   open(newunit=u, file=cas_file, status='old', action='read')
   cmd = ''
   stat = 0
   do while (stat == 0)
      read(u, '(A)', iostat=stat) record
      record = adjustl(record)
      if (stat /= 0 .or. record(1:1) == '&') then
         call treat_cmd(trim(cmd))
         exit
      endif
      lenrc = len_trim(record)
      if (lenrc == 0) cycle
      if (record(1:1) == '#') cycle
      cmd = trim(cmd) // trim(record)
      if (record(lenrc:lenrc) == ',') cycle
      call treat_cmd(trim(cmd))
      cmd = ''
   enddo
   close(u)
end subroutine

subroutine make_upper(s)
   character(len=*), intent(inout) :: s
   integer :: i, code
   do i = 1, len_trim(s)
      code = ichar(s(i:i))
      if (code >= ichar('a') .and. code <= ichar('z')) s(i:i) = char(code - 32)
   enddo
end subroutine

! The first value of a comma-separated list.
function first_value(v) result(f)
   character(len=*), intent(in) :: v
   character(len=len(v)) :: f
   integer :: k
   k = index(v, ',')
   if (k > 0) then
      f = v(1:k-1)
   else
      f = v
   endif
end function

subroutine treat_cmd(cmd_in)
   character(len=*), intent(in) :: cmd_in
   character(len=len(cmd_in)) :: cmd, pname, pvalu
   integer :: lenc, idx
   cmd = adjustl(cmd_in)
   lenc = len_trim(cmd)
   if (lenc == 0) return
   if (cmd(lenc:lenc) == ',') then
      cmd(lenc:lenc) = ' '
      lenc = lenc - 1
   endif
   idx = index(cmd, '=')
   if (idx == 0) return
   pname = cmd(1:idx-1)
   pvalu = cmd(idx+1:lenc)
   if (len_trim(adjustl(pname)) == 0) return
   if (len_trim(adjustl(pvalu)) == 0) return
   pvalu = adjustl(pvalu)
   call make_upper(pname)
   select case (trim(adjustl(pname)))
   case ('PRINT STEP')
      read(pvalu, *) print_step
      write(*,*) 'Print step = ', print_step
   case ('SAVE STEP')
      read(pvalu, *) save_step
      write(*,*) 'Save step = ', save_step
   case ('PARTICLE PAIRS FIND METHOD')
      read(pvalu, *) cas_nnps
      set_nnps = .true.
      write(*,*) 'Particle pairs find method = ', cas_nnps
   case ('RESULTS FILE')
      read(pvalu, *) results, res_soil, res_other
      write(*,*) 'Results file = ', trim(results), ', ', trim(res_soil), ', ', trim(res_other)
   case ('INLET VELOCITY')
      read(pvalu, *) inlet_velocity
      write(*,*) 'Inlet velocity = ', inlet_velocity
   case ('SINGLE PHASE')
      read(first_value(pvalu), *) single_phase_cas
      set_single = .true.
      write(*,*) 'Single phase = ', single_phase_cas
   case ('FIRST PHASE')
      read(pvalu, *) first_phase
      write(*,*) 'First phase = ', first_phase
   case ('SECOND PHASE')
      read(pvalu, *) second_phase
      write(*,*) 'Second phase = ', second_phase
   case ('DRY SOIL')
      read(pvalu, *) dry_soil
      write(*,*) 'Dry soil = ', dry_soil
   case ('PARTICLE SIZE')
      read(first_value(pvalu), *) particle_size
      write(*,*) 'Particle size = ', particle_size
   case ('OPERATING MODE')
      read(pvalu, *) operating_mode
      write(*,*) 'Operating mode = ', operating_mode
   case ('IMMERSE DEPTH')
      ! read and ignored, as in the reference
   case ('WITH NOZZLE')
      read(pvalu, *) with_nozzle
      write(*,*) 'With nozzle = ', with_nozzle
   case ('WATER PROPERTY')
      read(pvalu, *) cas_water
      set_water = .true.
      write(*,*) 'Water property: rho0, b, gamma, c, viscosity'
      write(*,*) cas_water
   case ('SOIL PROPERTY')
      read(pvalu, *) cas_soil
      set_soil = .true.
      write(*,*) 'Soil property: rho0, k, Youngs, Poisson, c'
      write(*,*) cas_soil
   case ('DARCY PARAMETER')
      read(pvalu, *) cas_darcy
      set_darcy = .true.
      write(*,*) 'Darcy parameter: porosity, permeability '
      write(*,*) cas_darcy
   case ('INTEGRATION SCHEME')
      read(pvalu, *) integrate_scheme
      write(*,*) 'Integration scheme: ', integrate_scheme
   case ('SOIL FAILURE CRITERIA')
      read(pvalu, *) cas_failure
      set_failure = .true.
      write(*,*) 'Soil failure criteria: cohesion, phi'
      write(*,*) cas_failure
   case ('TIME STEP SIZE')
      read(first_value(pvalu), *) cas_dt
      set_dt = .true.
      write(*,*) 'Time step size = ', cas_dt
   case ('MAXIMUM TIME STEP')
      ! read and ignored, as in the reference
   case default
      write(*,*) "Error: No such keyword -- ", trim(pname)
      stop
   end select
end subroutine

! Copy the steering-file overrides into the run's sphg_params.
subroutine apply_cas(p)
   type(sphg_params_t), intent(inout) :: p
   if (set_nnps) p%nnps = cas_nnps
   if (set_single) p%single_phase = merge(1, 0, single_phase_cas)
   p%integrate_scheme = integrate_scheme
   if (set_water) then
      p%water_rho0 = cas_water(1); p%water_b = cas_water(2); p%water_gamma = cas_water(3)
      p%water_c = cas_water(4); p%water_viscosity = cas_water(5)
   endif
   if (set_soil) then
      p%soil_rho0 = cas_soil(1); p%soil_k = cas_soil(2); p%soil_E = cas_soil(3)
      p%soil_niu = cas_soil(4); p%soil_c = cas_soil(5)
   endif
   if (set_darcy) then
      p%soil_porosity = cas_darcy(1); p%soil_permeability = cas_darcy(2)
   endif
   if (set_failure) then
      p%soil_cohesion = cas_failure(1); p%soil_phi = cas_failure(2)
   endif
   if (set_dt) p%dt = cas_dt
end subroutine

! The parsed steering file as `name = value` lines (used by the tests).
subroutine dump_cas()
   write(*,'(A,I0)') 'print_step = ', print_step
   write(*,'(A,I0)') 'save_step = ', save_step
   write(*,'(A,I0)') 'nnps = ', merge(cas_nnps, -1, set_nnps)
   write(*,'(A,A)') 'results = ', trim(results)
   write(*,'(A,A)') 'res_soil = ', trim(res_soil)
   write(*,'(A,A)') 'res_other = ', trim(res_other)
   write(*,'(A,L1)') 'single_phase = ', single_phase_cas
   write(*,'(A,A)') 'first_phase = ', trim(first_phase)
   write(*,'(A,A)') 'second_phase = ', trim(second_phase)
   write(*,'(A,L1)') 'dry_soil = ', dry_soil
   write(*,'(A,I0)') 'integrate_scheme = ', integrate_scheme
   write(*,'(A,5ES25.16E3)') 'water = ', cas_water
   write(*,'(A,5ES25.16E3)') 'soil = ', cas_soil
   write(*,'(A,2ES25.16E3)') 'darcy = ', cas_darcy
   write(*,'(A,2ES25.16E3)') 'failure = ', cas_failure
   write(*,'(A,ES25.16E3)') 'dt = ', cas_dt
end subroutine

subroutine open_files()
   open(newunit=f_xv, file=trim(results))
   open(newunit=f_state, file=trim(res_soil))
   open(newunit=f_other, file=trim(res_other))
end subroutine

subroutine close_files()
   if (f_xv >= 0) close(f_xv)
   if (f_state >= 0) close(f_state)
   if (f_other >= 0) close(f_other)
end subroutine

! Tecplot BLOCK zones: water (or the single phase) into f_xv, soil into
! f_state, with the reference's variable lists.  The reference's particle
! zone numbers are not part of the state: the "zone" column is 0.
subroutine output(parts, soil, nph)
   type(particles), intent(in) :: parts, soil
   integer, intent(in) :: nph
   integer :: n, n2
   real(c_double), allocatable :: zone(:)
   n = parts%ntotal + parts%nvirt
   allocate(zone(max(n, soil%ntotal + soil%nvirt)))
   zone = 0.d0
   if (parts%material == SPHG_WATER) then
      write(f_xv,*) 'VARIABLES="X","Y","Pressure","VoF",'
      write(f_xv,*) '"vx","vy","rho","zone","VoF2","mass" '
      write(f_xv,*) 'ZONE I=', n, ' F=BLOCK'
      write(f_xv,*) parts%x(1,1:n)
      write(f_xv,*) parts%x(2,1:n)
      write(f_xv,*) parts%p(1:n)
      write(f_xv,*) parts%vof(1:n)
      write(f_xv,*) parts%vx(1,1:n)
      write(f_xv,*) parts%vx(2,1:n)
      write(f_xv,*) parts%rho(1:n)
      write(f_xv,*) zone(1:n)
      write(f_xv,*) parts%vof2(1:n)
      write(f_xv,*) parts%mass(1:n)
   else
      call soil_zone(f_xv, parts, n)
   endif
   if (nph < 2) return
   n2 = soil%ntotal + soil%nvirt
   call soil_zone(f_state, soil, n2)
contains
   subroutine soil_zone(f, pl, m)
      integer, intent(in) :: f, m
      type(particles), intent(in) :: pl
      write(f,*) 'VARIABLES="X","Y","Pressure", "VoF", "ep", '
      write(f,*) '"sxy", "sxx","syy","vx","vy", "rho","mass" '
      write(f,*) '"sigma_yy", "zone" '
      write(f,*) 'ZONE I=', m, ' F=BLOCK'
      write(f,*) pl%x(1,1:m)
      write(f,*) pl%x(2,1:m)
      write(f,*) pl%p(1:m)
      write(f,*) pl%vof(1:m)
      write(f,*) pl%epsilon_p(1:m)
      write(f,*) pl%sxy(1:m)
      write(f,*) pl%sxx(1:m)
      write(f,*) pl%syy(1:m)
      write(f,*) pl%vx(1,1:m)
      write(f,*) pl%vx(2,1:m)
      write(f,*) pl%rho(1:m)
      write(f,*) pl%mass(1:m)
      write(f,*) -pl%p(1:m) + pl%syy(1:m)
      write(f,*) zone(1:m)
   end subroutine
end subroutine

subroutine interaction_statistics(pl)
   type(particles), intent(in) :: pl
   integer :: n, i, sumiac, maxiac, miniac, noiac, maxp, minp
   n = pl%ntotal + pl%nvirt
   sumiac = 0; maxiac = 0; miniac = 1000; noiac = 0; maxp = 0; minp = 0
   do i = 1, n
      sumiac = sumiac + pl%countiac(i)
      if (pl%countiac(i) > maxiac) then
         maxiac = pl%countiac(i)
         maxp = i
      endif
      if (pl%countiac(i) < miniac) then
         miniac = pl%countiac(i)
         minp = i
      endif
      if (pl%countiac(i) == 0) noiac = noiac + 1
   enddo
   print *, ' >> Statistics: interactions per particle:'
   print *, '**** Particle:', maxp, ' maximal interactions:', maxiac
   print *, '**** Particle:', minp, ' minimal interactions:', miniac
   print *, '**** Average :', real(sumiac) / real(n)
   print *, '**** Total pairs : ', pl%niac
   print *, '**** Particles with no interactions:', noiac
end subroutine

subroutine minimum_time_step(pl, c, viscosity, rho0)
   type(particles), intent(in) :: pl
   real(c_double), intent(in) :: c, viscosity, rho0
   integer :: n
   real(c_double) :: miu, mint1, mint21, mint22, mint3
   n = pl%ntotal
   miu = viscosity / rho0
   mint1 = 0.25 * minval(pl%hsml(1:n) / c)
   mint21 = 0.25 * minval(sqrt(pl%hsml(1:n) / pl%dvx(1,1:n)))
   mint22 = 0.25 * minval(sqrt(pl%hsml(1:n) / pl%dvx(2,1:n)))
   mint3 = 0.125 * minval(pl%hsml(1:n)**2 / miu)
   write(*,*) 'Minimum time step is: '
   write(*,*) ' mint1 = ', mint1
   write(*,*) ' mint21 = ', mint21, ' mint22 = ', mint22
   write(*,*) ' mint3 = ', mint3
end subroutine

end module
