! Fortran interfaces of the C-ABI entries in include/ftc_les.h, ftc_dense.h, ftc_reduce.h, ftc_poisson.h and ftc_box.h.
!
! One module per entry, each exporting `run`, in the shape the reference compiler gives its generated
! interfaces (compiler/emission/cuda/scoped.py:514-522): the context and the mode by value, array handles
! as integer(c_int64_t) by value, scalars by reference with intent(in). A spliced block calls
!   fort_status = run(ctx, FORT_SCOPE_GPU, u2_handle, u_handle, v_handle, w_handle, dxmin, dymin, dzmin, nx, ny, nz)
! and runs its original statements when fort_status is FORT_SCOPE_DECLINED (9).
module fort_shared_momentumadvection_cdu
  use iso_c_binding
  implicit none
  private
  public :: run
  interface
    function run(fort_context, fort_mode, u2_handle, u_handle, v_handle, w_handle, dxmin, dymin, dzmin, &
                 unx, uny, unz) bind(C, name='cpp_fort_shared_momentumadvection_cdu') result(fort_status)
      import :: c_int, c_int64_t, c_double
      integer(c_int) :: fort_status
      integer(c_int64_t), value :: fort_context
      integer(c_int), value :: fort_mode
      integer(c_int64_t), value :: u2_handle, u_handle, v_handle, w_handle
      real(c_double), intent(in) :: dxmin, dymin, dzmin
      integer(c_int), intent(in) :: unx, uny, unz
    end function
  end interface
end module fort_shared_momentumadvection_cdu

module fort_shared_momentumadvection_cdv
  use iso_c_binding
  implicit none
  private
  public :: run
  interface
    function run(fort_context, fort_mode, v2_handle, u_handle, v_handle, w_handle, dxmin, dymin, dzmin, &
                 vnx, vny, vnz) bind(C, name='cpp_fort_shared_momentumadvection_cdv') result(fort_status)
      import :: c_int, c_int64_t, c_double
      integer(c_int) :: fort_status
      integer(c_int64_t), value :: fort_context
      integer(c_int), value :: fort_mode
      integer(c_int64_t), value :: v2_handle, u_handle, v_handle, w_handle
      real(c_double), intent(in) :: dxmin, dymin, dzmin
      integer(c_int), intent(in) :: vnx, vny, vnz
    end function
  end interface
end module fort_shared_momentumadvection_cdv

module fort_shared_momentumadvection_cdw
  use iso_c_binding
  implicit none
  private
  public :: run
  interface
    function run(fort_context, fort_mode, w2_handle, u_handle, v_handle, w_handle, dxmin, dymin, dzmin, &
                 wnx, wny, wnz) bind(C, name='cpp_fort_shared_momentumadvection_cdw') result(fort_status)
      import :: c_int, c_int64_t, c_double
      integer(c_int) :: fort_status
      integer(c_int64_t), value :: fort_context
      integer(c_int), value :: fort_mode
      integer(c_int64_t), value :: w2_handle, u_handle, v_handle, w_handle
      real(c_double), intent(in) :: dxmin, dymin, dzmin
      integer(c_int), intent(in) :: wnx, wny, wnz
    end function
  end interface
end module fort_shared_momentumadvection_cdw

module fort_shared_les_tendency
  use iso_c_binding
  implicit none
  private
  public :: run
  interface
    function run(fort_context, fort_mode, tu_handle, tv_handle, tw_handle, ts_handle, u_handle, v_handle, &
                 w_handle, s_handle, dx, dy, dz, cs, nx, ny, nz, sgs) &
        bind(C, name='cpp_fort_shared_les_tendency') result(fort_status)
      import :: c_int, c_int64_t, c_double
      integer(c_int) :: fort_status
      integer(c_int64_t), value :: fort_context
      integer(c_int), value :: fort_mode
      integer(c_int64_t), value :: tu_handle, tv_handle, tw_handle, ts_handle
      integer(c_int64_t), value :: u_handle, v_handle, w_handle, s_handle
      real(c_double), intent(in) :: dx, dy, dz, cs
      integer(c_int), intent(in) :: nx, ny, nz, sgs
    end function
  end interface
end module fort_shared_les_tendency

module fort_shared_les_step
  use iso_c_binding
  implicit none
  private
  public :: run
  interface
    function run(fort_context, fort_mode, un_handle, vn_handle, wn_handle, sn_handle, u_handle, v_handle, &
                 w_handle, s_handle, dx, dy, dz, dt, cs, nx, ny, nz, sgs) &
        bind(C, name='cpp_fort_shared_les_step') result(fort_status)
      import :: c_int, c_int64_t, c_double
      integer(c_int) :: fort_status
      integer(c_int64_t), value :: fort_context
      integer(c_int), value :: fort_mode
      integer(c_int64_t), value :: un_handle, vn_handle, wn_handle, sn_handle
      integer(c_int64_t), value :: u_handle, v_handle, w_handle, s_handle
      real(c_double), intent(in) :: dx, dy, dz, dt, cs
      integer(c_int), intent(in) :: nx, ny, nz, sgs
    end function
  end interface
end module fort_shared_les_step

module fort_shared_chain_contraction
  use iso_c_binding
  implicit none
  private
  public :: run
  interface
    function run(fort_context, fort_mode, d_handle, a_handle, b_handle, c_handle, n, precision) &
        bind(C, name='cpp_fort_shared_chain_contraction') result(fort_status)
      import :: c_int, c_int64_t
      integer(c_int) :: fort_status
      integer(c_int64_t), value :: fort_context
      integer(c_int), value :: fort_mode
      integer(c_int64_t), value :: d_handle, a_handle, b_handle, c_handle
      integer(c_int), intent(in) :: n, precision
    end function
  end interface
end module fort_shared_chain_contraction

module fort_shared_fixed_order_reduction
  use iso_c_binding
  implicit none
  private
  public :: run
  interface
    ! value: in, the variable before the nest; out, after (the result is on the host on return).
    function run(fort_context, fort_mode, x_handle, op, lo3, hi3, value) &
        bind(C, name='cpp_fort_shared_fixed_order_reduction') result(fort_status)
      import :: c_int, c_int64_t, c_long_long, c_double
      integer(c_int) :: fort_status
      integer(c_int64_t), value :: fort_context
      integer(c_int), value :: fort_mode
      integer(c_int64_t), value :: x_handle
      integer(c_int), intent(in) :: op
      integer(c_long_long), intent(in) :: lo3(3), hi3(3)
      real(c_double), intent(inout) :: value
    end function
  end interface
end module fort_shared_fixed_order_reduction

module fort_shared_pressure_divergence
  use iso_c_binding
  implicit none
  private
  public :: run
  interface
    function run(fort_context, fort_mode, rhs_handle, u_handle, v_handle, w_handle, dx, dy, dz, dt, nx, ny, nz) &
        bind(C, name='cpp_fort_shared_pressure_divergence') result(fort_status)
      import :: c_int, c_int64_t, c_double
      integer(c_int) :: fort_status
      integer(c_int64_t), value :: fort_context
      integer(c_int), value :: fort_mode
      integer(c_int64_t), value :: rhs_handle, u_handle, v_handle, w_handle
      real(c_double), intent(in) :: dx, dy, dz, dt
      integer(c_int), intent(in) :: nx, ny, nz
    end function
  end interface
end module fort_shared_pressure_divergence

module fort_shared_pressure_projection
  use iso_c_binding
  implicit none
  private
  public :: run
  interface
    function run(fort_context, fort_mode, u_handle, v_handle, w_handle, phi_handle, dx, dy, dz, dt, nx, ny, nz) &
        bind(C, name='cpp_fort_shared_pressure_projection') result(fort_status)
      import :: c_int, c_int64_t, c_double
      integer(c_int) :: fort_status
      integer(c_int64_t), value :: fort_context
      integer(c_int), value :: fort_mode
      integer(c_int64_t), value :: u_handle, v_handle, w_handle, phi_handle
      real(c_double), intent(in) :: dx, dy, dz, dt
      integer(c_int), intent(in) :: nx, ny, nz
    end function
  end interface
end module fort_shared_pressure_projection

module fort_shared_pressure_fix_reference
  use iso_c_binding
  implicit none
  private
  public :: run
  interface
    function run(fort_context, fort_mode, phi_handle, phi_ref, nx, ny, nz) &
        bind(C, name='cpp_fort_shared_pressure_fix_reference') result(fort_status)
      import :: c_int, c_int64_t, c_double
      integer(c_int) :: fort_status
      integer(c_int64_t), value :: fort_context
      integer(c_int), value :: fort_mode
      integer(c_int64_t), value :: phi_handle
      real(c_double), intent(in) :: phi_ref
      integer(c_int), intent(in) :: nx, ny, nz
    end function
  end interface
end module fort_shared_pressure_fix_reference

! The box application's device procedures (include/ftc_box.h).
module fort_shared_boundaries_periodic
  use iso_c_binding
  implicit none
  private
  public :: run
  interface
    function run(fort_context, fort_mode, a_handle) bind(C, name='cpp_fort_shared_boundaries_periodic') &
        result(fort_status)
      import :: c_int, c_int64_t
      integer(c_int) :: fort_status
      integer(c_int64_t), value :: fort_context
      integer(c_int), value :: fort_mode
      integer(c_int64_t), value :: a_handle
    end function
  end interface
end module fort_shared_boundaries_periodic

module fort_shared_stepper_diffuse
  use iso_c_binding
  implicit none
  private
  public :: run
  interface
    function run(fort_context, fort_mode, a_handle, b_handle, dx, dt, nu) &
        bind(C, name='cpp_fort_shared_stepper_diffuse') result(fort_status)
      import :: c_int, c_int64_t, c_double
      integer(c_int) :: fort_status
      integer(c_int64_t), value :: fort_context
      integer(c_int), value :: fort_mode
      integer(c_int64_t), value :: a_handle, b_handle
      real(c_double), intent(in) :: dx, dt, nu
    end function
  end interface
end module fort_shared_stepper_diffuse

module fort_shared_stepper_store
  use iso_c_binding
  implicit none
  private
  public :: run
  interface
    function run(fort_context, fort_mode, a_handle, b_handle) bind(C, name='cpp_fort_shared_stepper_store') &
        result(fort_status)
      import :: c_int, c_int64_t
      integer(c_int) :: fort_status
      integer(c_int64_t), value :: fort_context
      integer(c_int), value :: fort_mode
      integer(c_int64_t), value :: a_handle, b_handle
    end function
  end interface
end module fort_shared_stepper_store

module fort_shared_relax_sweep
  use iso_c_binding
  implicit none
  private
  public :: run
  interface
    function run(fort_context, fort_mode, v_handle, u_handle, n) &
        bind(C, name='cpp_fort_shared_relax_sweep') result(fort_status)
      import :: c_int, c_int64_t
      integer(c_int) :: fort_status
      integer(c_int64_t), value :: fort_context
      integer(c_int), value :: fort_mode
      integer(c_int64_t), value :: v_handle, u_handle
      integer(c_int), intent(in) :: n
    end function
  end interface
end module fort_shared_relax_sweep

module fort_shared_relax_change
  use iso_c_binding
  implicit none
  private
  public :: run
  interface
    function run(fort_context, fort_mode, v_handle, u_handle, n, change) &
        bind(C, name='cpp_fort_shared_relax_change') result(fort_status)
      import :: c_int, c_int64_t, c_double
      integer(c_int) :: fort_status
      integer(c_int64_t), value :: fort_context
      integer(c_int), value :: fort_mode
      integer(c_int64_t), value :: v_handle, u_handle
      integer(c_int), intent(in) :: n
      real(c_double), intent(inout) :: change
    end function
  end interface
end module fort_shared_relax_change

module fort_shared_relax_copy
  use iso_c_binding
  implicit none
  private
  public :: run
  interface
    function run(fort_context, fort_mode, u_handle, v_handle, n) &
        bind(C, name='cpp_fort_shared_relax_copy') result(fort_status)
      import :: c_int, c_int64_t
      integer(c_int) :: fort_status
      integer(c_int64_t), value :: fort_context
      integer(c_int), value :: fort_mode
      integer(c_int64_t), value :: u_handle, v_handle
      integer(c_int), intent(in) :: n
    end function
  end interface
end module fort_shared_relax_copy
