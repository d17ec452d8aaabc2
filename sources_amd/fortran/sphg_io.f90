!-----------------------------------------------------------------------
! sphg_io.f90 -- host-side front and back end of the Fortran driver:
!   read_cas / dispatch    the steering file (its syntax and keywords are
!                          those of declarations_sph.f90:61-335): lines of
!                          `NAME = value[, value ...]`, names case-insensitive,
!                          '#' comments, a trailing ',' continues a command on
!                          the next line, a line starting with '&' ends the
!                          file.  A keyword table drives the parse;
!   output                 Tecplot BLOCK zones of both phases
!                          (declarations_sph.f90:1703-1810);
!   interaction_statistics / minimum_time_step
!                          the reference's diagnostics (m_particles.f90:555-630).
! Parsed values land in one slot per keyword; settle() and apply_cas copy
! them into the run-control variables and the run's sphg_params (after the
! initial state is read).
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

! ---- the keyword table -------------------------------------------------
! Each steering keyword (declarations_sph.f90:125-325 names the set) is a
! row: how many values of which kind it takes and how its value is echoed.
integer, parameter :: KW_INT = 1, KW_REAL = 2, KW_LOGICAL = 3, KW_WORD = 4, KW_IGNORED = 5
integer, parameter :: ECHO_NONE = 0, ECHO_INLINE = 1, ECHO_BELOW = 2, ECHO_JOINED = 3
integer, parameter :: MAXVAL_KW = 5

type keyword
   character(len=28) :: name          ! upper case, as matched
   integer :: kind, nval, echo
   character(len=48) :: label
end type

integer, parameter :: K_PRINT = 1, K_SAVE = 2, K_NNPS = 3, K_RESULTS = 4, K_INLET = 5, K_SINGLE = 6, &
                      K_FIRST = 7, K_SECOND = 8, K_DRY = 9, K_PSIZE = 10, K_MODE = 11, K_IMMERSE = 12, &
                      K_NOZZLE = 13, K_WATER = 14, K_SOIL = 15, K_DARCY = 16, K_SCHEME = 17, &
                      K_FAILURE = 18, K_DT = 19, K_MAXDT = 20, NKW = 20

type(keyword), parameter :: table(NKW) = [ &
   keyword('PRINT STEP',                 KW_INT,     1, ECHO_INLINE, 'Print step ='), &
   keyword('SAVE STEP',                  KW_INT,     1, ECHO_INLINE, 'Save step ='), &
   keyword('PARTICLE PAIRS FIND METHOD', KW_INT,     1, ECHO_INLINE, 'Particle pairs find method ='), &
   keyword('RESULTS FILE',               KW_WORD,    3, ECHO_JOINED, 'Results file ='), &
   keyword('INLET VELOCITY',             KW_REAL,    1, ECHO_INLINE, 'Inlet velocity ='), &
   keyword('SINGLE PHASE',               KW_LOGICAL, 1, ECHO_INLINE, 'Single phase ='), &
   keyword('FIRST PHASE',                KW_WORD,    1, ECHO_INLINE, 'First phase ='), &
   keyword('SECOND PHASE',               KW_WORD,    1, ECHO_INLINE, 'Second phase ='), &
   keyword('DRY SOIL',                   KW_LOGICAL, 1, ECHO_INLINE, 'Dry soil ='), &
   keyword('PARTICLE SIZE',              KW_WORD,    1, ECHO_INLINE, 'Particle size ='), &
   keyword('OPERATING MODE',             KW_WORD,    1, ECHO_INLINE, 'Operating mode ='), &
   keyword('IMMERSE DEPTH',              KW_IGNORED, 0, ECHO_NONE,   ''), &
   keyword('WITH NOZZLE',                KW_LOGICAL, 1, ECHO_INLINE, 'With nozzle ='), &
   keyword('WATER PROPERTY',             KW_REAL,    5, ECHO_BELOW,  'Water property: rho0, b, gamma, c, viscosity'), &
   keyword('SOIL PROPERTY',              KW_REAL,    5, ECHO_BELOW,  'Soil property: rho0, k, Youngs, Poisson, c'), &
   keyword('DARCY PARAMETER',            KW_REAL,    2, ECHO_BELOW,  'Darcy parameter: porosity, permeability'), &
   keyword('INTEGRATION SCHEME',         KW_INT,     1, ECHO_INLINE, 'Integration scheme:'), &
   keyword('SOIL FAILURE CRITERIA',      KW_REAL,    2, ECHO_BELOW,  'Soil failure criteria: cohesion, phi'), &
   keyword('TIME STEP SIZE',             KW_REAL,    1, ECHO_INLINE, 'Time step size ='), &
   keyword('MAXIMUM TIME STEP',          KW_IGNORED, 0, ECHO_NONE,   '') ]

! ---- the parsed values, one slot per keyword row ---------------------
logical, save :: given(NKW) = .false.
integer, save :: int_of(NKW) = 0
real(c_double), save :: real_of(MAXVAL_KW, NKW) = 0.d0
logical, save :: flag_of(NKW) = .false.
character(len=256), save :: word_of(3, NKW) = ''

contains

! The steering file is read whole and cut into lines.  A line is blank,
! a comment ('#' first), the end mark ('&' first: the rest of the file is
! not read) or part of a command; a command continues on the next line
! while its line ends with ','.  Leading and trailing blanks do not count.
subroutine read_cas(cas_file)
   character(len=*), intent(in) :: cas_file
   character(len=:), allocatable :: text, command
   integer :: u, nbytes, first, last
   logical :: present
   inquire(file=cas_file, exist=present)
   if (.not. present) then
      write(*,*) "Steering file ", trim(cas_file), " doesn't exist."
      stop
   endif
   open(newunit=u, file=cas_file, access='stream', form='unformatted', status='old', action='read')
   inquire(unit=u, size=nbytes)
   allocate(character(len=max(nbytes, 0)) :: text)
   if (nbytes > 0) read(u) text
   close(u)
   command = ''
   first = 1
   do while (first <= nbytes)
      last = index(text(first:), achar(10))
      if (last == 0) then
         last = nbytes
      else
         last = first + last - 2
      endif
      block
         character(len=:), allocatable :: ln
         ln = trimmed_line(text(first:last))
         first = last + 2
         if (len(ln) == 0) cycle
         if (ln(1:1) == '&') exit
         if (ln(1:1) == '#') cycle
         command = command // ln
         if (ln(len(ln):) /= ',') then
            call dispatch(command)
            command = ''
         endif
      end block
   enddo
   call dispatch(command)
   call settle()
end subroutine

! A line without its carriage return and its leading / trailing blanks.
function trimmed_line(raw) result(ln)
   character(len=*), intent(in) :: raw
   character(len=:), allocatable :: ln
   integer :: a, b
   b = len(raw)
   if (b > 0) then
      if (raw(b:b) == achar(13)) b = b - 1
   endif
   a = verify(raw(1:b), ' ' // achar(9))
   if (a == 0) then
      ln = ''
   else
      ln = trim(raw(a:b))
   endif
end function

! Upper-case copy of an ASCII string.
function upper(s) result(u)
   character(len=*), intent(in) :: s
   character(len=len(s)) :: u
   integer :: k
   u = s
   do k = 1, len(s)
      if (lge(s(k:k), 'a') .and. lle(s(k:k), 'z')) u(k:k) = achar(iachar(s(k:k)) - 32)
   enddo
end function

! Row of the keyword table with this (upper-case) name; 0 if none.
integer function row_of(name)
   character(len=*), intent(in) :: name
   integer :: k
   row_of = 0
   do k = 1, NKW
      if (table(k)%name == name) then
         row_of = k
         return
      endif
   enddo
end function

! One command `NAME = v1[, v2 ...]`: looked up in the table, its values
! read list-directed into the keyword's slot and echoed.
subroutine dispatch(command)
   character(len=*), intent(in) :: command
   character(len=:), allocatable :: name, vals
   integer :: eq, k
   eq = index(command, '=')
   if (eq == 0) return
   name = upper(trim(adjustl(command(1:eq-1))))
   vals = trim(adjustl(command(eq+1:)))
   if (len(vals) > 0) then
      if (vals(len(vals):) == ',') vals = vals(1:len(vals)-1)
   endif
   if (len(name) == 0 .or. len_trim(vals) == 0) return
   k = row_of(name)
   if (k == 0) then
      write(*,*) "Error: No such keyword -- ", name
      stop
   endif
   select case (table(k)%kind)
   case (KW_INT)
      read(vals, *) int_of(k)
   case (KW_REAL)
      read(vals, *) real_of(1:table(k)%nval, k)
   case (KW_LOGICAL)
      read(vals, *) flag_of(k)
   case (KW_WORD)
      read(vals, *) word_of(1:table(k)%nval, k)
   case default
      return
   end select
   given(k) = .true.
   call echo(k)
end subroutine

subroutine echo(k)
   integer, intent(in) :: k
   character(len=:), allocatable :: lab
   lab = trim(table(k)%label) // ' '
   select case (table(k)%echo)
   case (ECHO_BELOW)
      write(*,*) trim(table(k)%label)
      write(*,*) real_of(1:table(k)%nval, k)
   case (ECHO_JOINED)
      write(*,*) lab, trim(word_of(1, k)), ', ', trim(word_of(2, k)), ', ', trim(word_of(3, k))
   case (ECHO_INLINE)
      select case (table(k)%kind)
      case (KW_INT)
         write(*,*) lab, int_of(k)
      case (KW_REAL)
         write(*,*) lab, real_of(1, k)
      case (KW_LOGICAL)
         write(*,*) lab, flag_of(k)
      case (KW_WORD)
         write(*,*) lab, trim(word_of(1, k))
      end select
   end select
end subroutine

! The run-control variables the driver uses, from the parsed slots.
subroutine settle()
   if (given(K_PRINT)) print_step = int_of(K_PRINT)
   if (given(K_SAVE)) save_step = int_of(K_SAVE)
   if (given(K_SCHEME)) integrate_scheme = int_of(K_SCHEME)
   if (given(K_RESULTS)) then
      results = word_of(1, K_RESULTS); res_soil = word_of(2, K_RESULTS); res_other = word_of(3, K_RESULTS)
   endif
   if (given(K_INLET)) inlet_velocity = real_of(1, K_INLET)
   if (given(K_SINGLE)) single_phase_cas = flag_of(K_SINGLE)
   if (given(K_FIRST)) first_phase = word_of(1, K_FIRST)
   if (given(K_SECOND)) second_phase = word_of(1, K_SECOND)
   if (given(K_DRY)) dry_soil = flag_of(K_DRY)
   if (given(K_PSIZE)) particle_size = word_of(1, K_PSIZE)
   if (given(K_MODE)) operating_mode = word_of(1, K_MODE)
   if (given(K_NOZZLE)) with_nozzle = flag_of(K_NOZZLE)
end subroutine

! Copy the steering-file values into the run's sphg_params.
subroutine apply_cas(p)
   type(sphg_params_t), intent(inout) :: p
   if (given(K_NNPS)) p%nnps = int_of(K_NNPS)
   if (given(K_SINGLE)) p%single_phase = merge(1, 0, flag_of(K_SINGLE))
   p%integrate_scheme = integrate_scheme
   if (given(K_WATER)) then
      p%water_rho0 = real_of(1, K_WATER); p%water_b = real_of(2, K_WATER); p%water_gamma = real_of(3, K_WATER)
      p%water_c = real_of(4, K_WATER); p%water_viscosity = real_of(5, K_WATER)
   endif
   if (given(K_SOIL)) then
      p%soil_rho0 = real_of(1, K_SOIL); p%soil_k = real_of(2, K_SOIL); p%soil_E = real_of(3, K_SOIL)
      p%soil_niu = real_of(4, K_SOIL); p%soil_c = real_of(5, K_SOIL)
   endif
   if (given(K_DARCY)) then
      p%soil_porosity = real_of(1, K_DARCY); p%soil_permeability = real_of(2, K_DARCY)
   endif
   if (given(K_FAILURE)) then
      p%soil_cohesion = real_of(1, K_FAILURE); p%soil_phi = real_of(2, K_FAILURE)
   endif
   if (given(K_DT)) p%dt = real_of(1, K_DT)
end subroutine

! The parsed steering file as `name = value` lines (used by the tests).
subroutine dump_cas()
   write(*,'(A,I0)') 'print_step = ', print_step
   write(*,'(A,I0)') 'save_step = ', save_step
   write(*,'(A,I0)') 'nnps = ', merge(int_of(K_NNPS), -1, given(K_NNPS))
   write(*,'(A,A)') 'results = ', trim(results)
   write(*,'(A,A)') 'res_soil = ', trim(res_soil)
   write(*,'(A,A)') 'res_other = ', trim(res_other)
   write(*,'(A,L1)') 'single_phase = ', single_phase_cas
   write(*,'(A,A)') 'first_phase = ', trim(first_phase)
   write(*,'(A,A)') 'second_phase = ', trim(second_phase)
   write(*,'(A,L1)') 'dry_soil = ', dry_soil
   write(*,'(A,I0)') 'integrate_scheme = ', integrate_scheme
   write(*,'(A,5ES25.16E3)') 'water = ', real_of(1:5, K_WATER)
   write(*,'(A,5ES25.16E3)') 'soil = ', real_of(1:5, K_SOIL)
   write(*,'(A,2ES25.16E3)') 'darcy = ', real_of(1:2, K_DARCY)
   write(*,'(A,2ES25.16E3)') 'failure = ', real_of(1:2, K_FAILURE)
   write(*,'(A,ES25.16E3)') 'dt = ', real_of(1, K_DT)
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

! Interactions per particle (countiac of the last pair search): the first
! particle with the most and with the fewest partners (the fewest counted
! below 1000, the reference's starting bound), the mean, the pair total and
! the particles without a partner (m_particles.f90:555-597 prints these).
subroutine interaction_statistics(pl)
   type(particles), intent(in) :: pl
   integer :: n, most, fewest, at_most, at_fewest
   n = pl%ntotal + pl%nvirt
   associate (cnt => pl%countiac(1:n))
      most = max(maxval(cnt), 0)
      at_most = 0
      if (most > 0) at_most = maxloc(cnt, dim=1)
      fewest = min(minval(cnt), 1000)
      at_fewest = 0
      if (fewest < 1000) at_fewest = minloc(cnt, dim=1)
      print *, ' >> Statistics: interactions per particle:'
      print *, '**** Particle:', at_most, ' maximal interactions:', most
      print *, '**** Particle:', at_fewest, ' minimal interactions:', fewest
      print *, '**** Average :', real(sum(cnt)) / real(n)
      print *, '**** Total pairs : ', pl%niac
      print *, '**** Particles with no interactions:', count(cnt == 0)
   end associate
end subroutine

! Time-step limits of the real particles (m_particles.f90:600-630 prints
! these): CFL h/c, the acceleration limits sqrt(h/a) per direction and the
! viscous diffusion limit h^2/nu, with the reference's safety factors.
subroutine minimum_time_step(pl, c, viscosity, rho0)
   type(particles), intent(in) :: pl
   real(c_double), intent(in) :: c, viscosity, rho0
   real(c_double) :: lim(4)
   integer :: n
   n = pl%ntotal
   associate (h => pl%hsml(1:n), ax => pl%dvx(1,1:n), ay => pl%dvx(2,1:n))
      lim = [0.25d0 * minval(h / c), 0.25d0 * minval(sqrt(h / ax)), 0.25d0 * minval(sqrt(h / ay)), &
             0.125d0 * minval(h**2 / (viscosity / rho0))]
   end associate
   write(*,*) 'Minimum time step is: '
   write(*,*) ' mint1 = ', lim(1)
   write(*,*) ' mint21 = ', lim(2), ' mint22 = ', lim(3)
   write(*,*) ' mint3 = ', lim(4)
end subroutine

end module
