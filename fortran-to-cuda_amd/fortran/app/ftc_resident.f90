! The residency domain of an application built by resident_app.py: one resident_domain object (the
! reference's compiler/runtime/resident_domain.f90) that every spliced block and expanded directive of
! the application uses, as the domain module of a build by the reference's resident flow does
! (compiler/resident/directives.py:85-90 name it through `use <module>, only: <object>`).
module ftc_resident
  use, intrinsic :: iso_c_binding
  use fort_resident_domain, only: resident_domain, ftc_require => resident_require
  implicit none
  private
  public :: ftc_domain, ftc_require
  type(resident_domain), save :: ftc_domain
end module ftc_resident
