! Fortran side of the pressure provider (include/ftc_poisson.h): the bind(C) mirrors of the configuration
! structs and the ELMM adapter calls, with the fields benchmarks/elmm/templates/pressure-solver.f90:16-24
! sets (version, flags, dimensions, approximation, sizes, lengths, boundary_conditions, budget).
module ftc_pressure
  use iso_c_binding
  implicit none
  private
  public :: poisson_solver_config_v1, elmm_gpu_pressure_config_v1, fort_scope_external_report_v1
  public :: elmm_gpu_pressure_create_v1, elmm_gpu_pressure_solve_v1, elmm_gpu_pressure_destroy_v1
  public :: POISSON_SOLVER_REAL32, POISSON_SOLVER_MIXED, BC_PERIODIC, BC_NEUMANN_STAGGERED

  integer(c_int32_t), parameter :: POISSON_SOLVER_REAL32 = 1, POISSON_SOLVER_MIXED = 2
  integer(c_int32_t), parameter :: BC_PERIODIC = 0, BC_NEUMANN_STAGGERED = 4

  type, bind(C) :: poisson_solver_config_v1
    integer(c_int32_t) :: version = 1, flags = 0
    integer(c_int32_t) :: dimensions = 3, approximation = 2
    integer(c_int32_t) :: sizes(3) = 0
    real(c_double) :: lengths(3) = 0
    integer(c_int32_t) :: boundary_conditions(6) = 0
    integer(c_int64_t) :: device_budget_bytes = 0
  end type

  type, bind(C) :: elmm_gpu_pressure_config_v1
    integer(c_int32_t) :: version = 1, reserved = 0
    type(poisson_solver_config_v1) :: solver
    integer(c_int64_t) :: phi_lower_bounds(3) = -1, rhs_lower_bounds(3) = 0
  end type

  type, bind(C) :: fort_scope_external_report_v1
    integer(c_int32_t) :: version = 1, status = 0
    integer(c_int32_t) :: started = 0, committed = 0
  end type

  interface
    ! domain: a fort_scope_domain_t; handle: out, the adapter (a C pointer).
    function elmm_gpu_pressure_create_v1(domain, config, handle) bind(C, name='elmm_gpu_pressure_create_v1') &
        result(status)
      import :: c_int, c_int64_t, c_ptr, elmm_gpu_pressure_config_v1
      integer(c_int) :: status
      integer(c_int64_t), value :: domain
      type(elmm_gpu_pressure_config_v1), intent(in) :: config
      type(c_ptr), intent(out) :: handle
    end function
    ! Solves lap(phi) = rhs for the registered arrays phi (n+4 extents) and rhs (n+2 extents).
    function elmm_gpu_pressure_solve_v1(context, handle, phi, rhs, report) bind(C, name='elmm_gpu_pressure_solve_v1') &
        result(status)
      import :: c_int, c_int64_t, c_ptr, fort_scope_external_report_v1
      integer(c_int) :: status
      integer(c_int64_t), value :: context
      type(c_ptr), value :: handle
      integer(c_int64_t), value :: phi, rhs
      type(fort_scope_external_report_v1), intent(inout) :: report
    end function
    function elmm_gpu_pressure_destroy_v1(handle) bind(C, name='elmm_gpu_pressure_destroy_v1') result(status)
      import :: c_int, c_ptr
      integer(c_int) :: status
      type(c_ptr), intent(inout) :: handle
    end function
  end interface
end module ftc_pressure
