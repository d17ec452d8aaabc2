!-----------------------------------------------------------------------
! sph_cas.f90 -- Fortran driver with the run-time structure of the
! reference's program SPH (sph.f90:1-78): steering file, initial state,
! Tecplot output of the initial state, minimum_time_step, then the
! interactive loop "maximal time steps?" / "more time steps? (0=No, 1=yes)"
! on stdin.  Steps run on the GPU through module sphg; every print_step a
! progress header is printed and every save_step the state is written with
! `output` (declarations_sph.f90:860-886).
!   usage: sph_cas <steering-file> <state-in>     (steps read from stdin)
!          sph_cas <steering-file> --parse-only   (print the parsed keywords)
!-----------------------------------------------------------------------
program sph_cas
use sphg
use sphg_io
implicit none
type(particles), target :: parts, soil
character(len=1024) :: cas_file, state_file
integer :: nph, maxtimestep, yesorno, ios

call get_command_argument(1, cas_file)
call get_command_argument(2, state_file)
write(*,*) 'cas_file:', trim(cas_file)
call read_cas(trim(cas_file))
if (trim(state_file) == '--parse-only') then
   call dump_cas()
   stop
endif

call open_files()
call read_state(trim(state_file), parts, soil, nph)
call apply_cas(params)
write(*,*) '  **************************************************'
write(*,*) '  Initial particle configuration generated   '
write(*,*) '  Total number of real particles   ', parts%ntotal
write(*,*) '  Total number of virtual particles  ', parts%nvirt
write(*,*) '  **************************************************'
if (nph > 1) then
   write(*,*) 'Total number of real soil particles:', soil%ntotal
   write(*,*) 'Total number of virtual particles:  ', soil%nvirt
endif

call sphg_init(params)
call sphg_bind(parts)
if (nph > 1) call sphg_bind(soil)

call output(parts, soil, nph)
call minimum_time_step(parts, params%water_c, params%water_viscosity, params%water_rho0)

do
   write(*,*) '  ***************************************************'
   write(*,*) '          Please input the maximal time steps '
   write(*,*) '  ***************************************************'
   read(*, *, iostat=ios) maxtimestep
   if (ios /= 0) exit
   call run_steps(maxtimestep)
   write(*,*) '  ***************************************************'
   write(*,*) 'Are you going to run more time steps ? (0=No, 1=yes)'
   write(*,*) '  ***************************************************'
   read(*, *, iostat=ios) yesorno
   if (ios /= 0) exit
   if (yesorno == 0) exit
enddo

call close_files()
call sphg_finalize()

contains

! maxtimestep steps in chunks that end on print / save boundaries.
subroutine run_steps(nsteps)
   integer, intent(in) :: nsteps
   integer :: left, chunk, it
   left = nsteps
   do while (left > 0)
      it = itimestep
      chunk = min(left, print_step - mod(it, print_step), save_step - mod(it, save_step))
      call time_integration(chunk)
      time = time + chunk * params%dt
      left = left - chunk
      if (mod(itimestep, print_step) == 0) then
         write(*,*) '______________________________________________'
         write(*,*) '  current number of time step =', itimestep, '     current time=', real(time + params%dt)
         write(*,*) '______________________________________________'
      endif
      if (itimestep >= save_step_from .and. mod(itimestep, save_step) == 0) call output(parts, soil, nph)
   enddo
end subroutine

end program
