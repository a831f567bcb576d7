! ELMM's resident pressure solve on this library: the reference's own adapter state
! (benchmarks/elmm/sources/resident_pressure_state.f90, with gpu_pressure.f90) drives
! elmm_gpu_pressure_*_v1 as ELMM's pressure-solver template does (benchmarks/elmm/templates/
! pressure-solver.f90): start with the configuration, bind Phi(-1:) and RHS(0:) persistent, solve,
! observe, finish. RHS comes in and Phi goes out as field dumps (in/fields-RHS.bin, out/fields-Phi.*).
!   resident_pressure nx ny nz dx flags      flags 0: FP64 solve, 2: split-bf16 with FP64 refinement
program resident_pressure
  use, intrinsic :: iso_c_binding
  use poisson_solver_device, only: POISSON_SOLVER_PERIODIC, POISSON_SOLVER_NEUMANN_STAGGERED
  use elmm_gpu_pressure, only: elmm_gpu_pressure_config_v1
  use elmm_resident_pressure_state, only: resident_flow
  implicit none
  real(c_double), allocatable, target :: Phi(:,:,:), RHS(:,:,:)
  type(elmm_gpu_pressure_config_v1) :: config
  character(len=32) :: word
  integer :: nx, ny, nz, unit
  integer(c_int32_t) :: flags
  real(c_double) :: dx

  call get_command_argument(1, word); read(word, *) nx
  call get_command_argument(2, word); read(word, *) ny
  call get_command_argument(3, word); read(word, *) nz
  call get_command_argument(4, word); read(word, *) dx
  call get_command_argument(5, word); read(word, *) flags
  allocate(Phi(-1:nx+2, -1:ny+2, -1:nz+2), RHS(0:nx+1, 0:ny+1, 0:nz+1))
  open(newunit=unit, file='in/fields-RHS.bin', access='stream', form='unformatted', status='old')
  read(unit) RHS
  close(unit)
  Phi = 0
  Phi(-1, -1, -1) = 123  ! a ghost the solve leaves alone

  config%solver%sizes = [nx, ny, nz]
  config%solver%lengths = [nx * dx, ny * dx, nz * dx]
  config%solver%boundary_conditions = [POISSON_SOLVER_PERIODIC, POISSON_SOLVER_PERIODIC, POISSON_SOLVER_PERIODIC, &
                                       POISSON_SOLVER_PERIODIC, POISSON_SOLVER_NEUMANN_STAGGERED, &
                                       POISSON_SOLVER_NEUMANN_STAGGERED]
  config%solver%approximation = 2
  config%solver%flags = flags
  call resident_flow%start(config)
  resident_flow%phi = resident_flow%bind3(Phi, lbound(Phi), .true., .true., origin='pressure::phi')
  resident_flow%rhs = resident_flow%bind3(RHS, lbound(RHS), .true., .true., origin='pressure::rhs')
  call resident_flow%solve()
  call resident_flow%observe_scratch()
  call resident_flow%finish()

  open(newunit=unit, file='out/fields-Phi.bin', access='stream', form='unformatted', status='replace')
  write(unit) Phi
  close(unit)
  open(newunit=unit, file='out/fields-Phi.shape', status='replace')
  write(unit, *) storage_size(Phi), shape(Phi), lbound(Phi), ubound(Phi)
  close(unit)
  print '(a)', 'done'
end program resident_pressure
