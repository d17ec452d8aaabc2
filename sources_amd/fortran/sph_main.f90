!-----------------------------------------------------------------------
! sph_main.f90 -- Fortran driver with the structure of the reference's
! program SPH (sph.f90:1-78): read the initial state, allocate, run
! time_integration for the requested number of steps, write the result.
!   usage: sph_gpu <state-in> <state-out> <maxtimestep>
!-----------------------------------------------------------------------
program sph_gpu
use sphg
implicit none
type(particles), target :: parts, soil
character(len=1024) :: infile, outfile, arg
integer :: maxtimestep, nph

call get_command_argument(1, infile)
call get_command_argument(2, outfile)
call get_command_argument(3, arg)
read(arg, *) maxtimestep

call read_state(trim(infile), parts, soil, nph)
write(*,*) ' Total number of real particles   ', parts%ntotal
write(*,*) ' Total number of virtual particles ', parts%nvirt
if (nph > 1) then
   write(*,*) 'Total number of real soil particles:', soil%ntotal
   write(*,*) 'Total number of virtual particles:  ', soil%nvirt
endif

call sphg_init(params)
call sphg_bind(parts)
if (nph > 1) call sphg_bind(soil)

call time_integration(maxtimestep)
write(*,*) ' time steps done: ', itimestep

call write_state(trim(outfile), parts, soil, nph)
call sphg_finalize()
end program
