! The Fortran face of the pressure provider this library puts in place of the poisson-solver submodule
! (jiriklepl/poisson-solver, absent from the reference tree): the module the reference's ELMM adapter
! uses for its configuration type (benchmarks/elmm/sources/gpu_pressure.f90:3), with the provider calls of
! include/ftc_poisson.h. The type has the layout of poisson_solver_config_v1 in that header, whose fields
! are the ones gpu_pressure.cpp:51-69 reads and benchmarks/elmm/templates/pressure-solver.f90 sets.
module poisson_solver_device
  use iso_c_binding
  implicit none
  private
  public :: poisson_solver_config_v1, poisson_solver_workspace_v1
  public :: poisson_solver_create_v1, poisson_solver_workspace_query_v1, poisson_solver_destroy_v1
  public :: POISSON_SOLVER_REAL32, POISSON_SOLVER_MIXED, POISSON_SOLVER_PERIODIC, POISSON_SOLVER_NEUMANN_STAGGERED

  ! flags: FP32 sweeps (REAL32), or split-bf16 transforms with FP64 refinement (MIXED, this library's own)
  integer(c_int32_t), parameter :: POISSON_SOLVER_REAL32 = 1, POISSON_SOLVER_MIXED = 2
  ! boundary_conditions, per face (x-, x+, y-, y+, z-, z+)
  integer(c_int32_t), parameter :: POISSON_SOLVER_PERIODIC = 0, POISSON_SOLVER_NEUMANN_STAGGERED = 4

  type, bind(C) :: poisson_solver_config_v1
    integer(c_int32_t) :: version = 1, flags = 0
    integer(c_int32_t) :: dimensions = 3, approximation = 2
    integer(c_int32_t) :: sizes(3) = 0
    real(c_double) :: lengths(3) = 0
    integer(c_int32_t) :: boundary_conditions(6) = 0
    integer(c_int64_t) :: device_budget_bytes = 0
  end type

  type, bind(C) :: poisson_solver_workspace_v1
    integer(c_int32_t) :: version = 1, flags = 0
    integer(c_int64_t) :: persistent_device_bytes = 0, peak_device_bytes = 0, execution_extra_bytes = 0
  end type

  interface
    function poisson_solver_create_v1(config, solver) bind(C, name='poisson_solver_create_v1') result(status)
      import :: c_int, c_ptr, poisson_solver_config_v1
      integer(c_int) :: status
      type(poisson_solver_config_v1), intent(in) :: config
      type(c_ptr), intent(out) :: solver
    end function
    function poisson_solver_workspace_query_v1(solver, bytes) bind(C, name='poisson_solver_workspace_query_v1') &
        result(status)
      import :: c_int, c_ptr, poisson_solver_workspace_v1
      integer(c_int) :: status
      type(c_ptr), value :: solver
      type(poisson_solver_workspace_v1), intent(inout) :: bytes
    end function
    function poisson_solver_destroy_v1(solver) bind(C, name='poisson_solver_destroy_v1') result(status)
      import :: c_int, c_ptr
      integer(c_int) :: status
      type(c_ptr), intent(inout) :: solver
    end function
  end interface
end module poisson_solver_device
