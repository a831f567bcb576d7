! ORACLE — test infrastructure only. A C-callable wrapper around the reference's own CDW subroutine
! (the module MomentumAdvection of the reference stencil source named by oracle/Makefile), so that tests can
! call the reference nest on the same arrays as the oracle and the HIP path.
module ref_cdw_wrap
  use iso_c_binding
  use MomentumAdvection
  implicit none
contains
  subroutine ref_cdw(out, u, v, w, dx, dy, dz, nx, ny, nz, e1, e2, e3) bind(C, name="ref_cdw")
    integer(c_int), value :: nx, ny, nz, e1, e2, e3
    real(c_double), value :: dx, dy, dz
    real(c_double), intent(inout) :: out(e1, e2, e3)
    real(c_double), intent(in) :: u(e1, e2, e3), v(e1, e2, e3), w(e1, e2, e3)
    call CDW(out, u, v, w, dx, dy, dz, nx, ny, nz)
  end subroutine ref_cdw
end module ref_cdw_wrap
