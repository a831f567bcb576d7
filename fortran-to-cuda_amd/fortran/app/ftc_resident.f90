! The residency domain of an application built by resident_app.py: one resident_domain object (the
! reference's compiler/runtime/resident_domain.f90) that every spliced block and expanded directive of
! the application uses, as the domain module of a build by the reference's resident flow does
! (compiler/resident/directives.py:85-90 name it through `use <module>, only: <object>`).
module ftc_resident
  use, intrinsic :: iso_c_binding
  use fort_resident_domain, only: resident_domain, ftc_require => resident_require
  implicit none
  private
  public :: ftc_domain, ftc_require, ftc_address
  type(resident_domain), save :: ftc_domain
contains
  ! The address of a variable that need not have the TARGET attribute (a field of an element of an array
  ! of derived type): the domain compares it with the addresses of the fields it mirrors and never
  ! dereferences it.
  function ftc_address(x) result(address)
    type(*), target, intent(in) :: x(..)
    type(c_ptr) :: address
    address = c_loc(x)
  end function ftc_address
end module ftc_resident
