!-----------------------------------------------------------------------
! sph_main.f90 -- Fortran driver with the structure of the reference's
! program SPH (sph.f90:1-78): read the initial state, allocate, run
! time_integration for the requested number of steps, write the result.
!   usage: sph_gpu <state-in> <state-out> <maxtimestep>
!          sph_gpu <state-in> <state-out> density
! The second form calls the granular density procedures at the reference's
! call sites instead: con_density and delta_sph(rho, drho) on every phase
! (single_step, declarations_sph.f90:1195-1207), then sum_density on the
! water (single_step_for_water, :1465), and writes the state.
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
maxtimestep = -1
if (trim(arg) /= 'density') read(arg, *) maxtimestep

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

if (maxtimestep >= 0) then
   call time_integration(maxtimestep)
   write(*,*) ' time steps done: ', itimestep
else
   call con_density(parts)
   call delta_sph_density(parts)
   if (nph > 1) then
      call con_density(soil)
      call delta_sph_density(soil)
   endif
   call sum_density(parts)
   write(*,*) ' density procedures done'
endif

call write_state(trim(outfile), parts, soil, nph)
call sphg_finalize()
end program
