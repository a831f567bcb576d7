! A program that calls this library's device entries the way the reference's spliced application code
! does (compiler/scopes/resident_splice.py:113-196): inside an epoch of the reference's own
! fort_resident_domain (compiler/runtime/resident_domain.f90), every array argument is bound at each
! call, tolerant, with its own lower bounds, and the entry's module procedure `run` is called with the
! domain's context, FORT_SCOPE_GPU, the handles and the scalars by reference. The entries' interface
! modules are the ones this library ships (fortran-to-cuda_amd/fortran/ftc_entries.f90), in the form the
! reference emits for an entry (compiler/emission/cuda/scoped.py:514-521).
!
! Inputs and outputs are field dumps in the reference's storage-shape-bounds-v1 format
! (benchmarks/elmm/templates/field-dumps.f90): fields-<name>.bin in storage order and fields-<name>.shape.
!   les   nx ny nz steps dx dt cs   reads in/fields-{U,V,W,S}.bin, takes `steps` explicit LES steps
!                                    alternating between two sets of arrays, writes out/fields-{U,V,W,S}
!   chain n precision               reads in/fields-{A,B,C}.bin, writes out/fields-D (D = A*B*C)
module application
  use, intrinsic :: iso_c_binding
  use, intrinsic :: iso_fortran_env, only: error_unit
  use fort_scoped_memory, only: FORT_SCOPE_GPU, FORT_SCOPE_OK, fort_scope_error
  use fort_resident_domain
  implicit none
  type(resident_domain) :: domain
contains
  subroutine check(status)
    integer(c_int), intent(in) :: status
    character(kind=c_char), pointer :: text(:)
    integer :: n
    if (status == FORT_SCOPE_OK) return
    write(error_unit, '(a,i0)') 'entry failed with status ', status
    if (c_associated(fort_scope_error())) then
      call c_f_pointer(fort_scope_error(), text, [512])
      n = 0
      do while (n < 511)
        if (text(n + 1) == c_null_char) exit
        n = n + 1
      end do
      write(error_unit, '(512a1)') text(:n)
    end if
    error stop 'resident entry failure'
  end subroutine

  subroutine les_step(un, vn, wn, sn, u, v, w, s, dx, dt, cs)
    use fort_shared_les_step, only: les_run => run
    real(c_double), dimension(0:,0:,0:), intent(inout), target :: un, vn, wn, sn, u, v, w, s
    real(c_double), intent(in) :: dx, dt, cs
    integer(c_int64_t) :: h(8)
    integer(c_int) :: nx, ny, nz
    nx = size(u, 1) - 2
    ny = size(u, 2) - 2
    nz = size(u, 3) - 2
    h(1) = domain%bind_real64(un, lbound(un), .true., origin='les_step::un', tolerant=.true.)
    h(2) = domain%bind_real64(vn, lbound(vn), .true., origin='les_step::vn', tolerant=.true.)
    h(3) = domain%bind_real64(wn, lbound(wn), .true., origin='les_step::wn', tolerant=.true.)
    h(4) = domain%bind_real64(sn, lbound(sn), .true., origin='les_step::sn', tolerant=.true.)
    h(5) = domain%bind_real64(u, lbound(u), .true., origin='les_step::u', tolerant=.true.)
    h(6) = domain%bind_real64(v, lbound(v), .true., origin='les_step::v', tolerant=.true.)
    h(7) = domain%bind_real64(w, lbound(w), .true., origin='les_step::w', tolerant=.true.)
    h(8) = domain%bind_real64(s, lbound(s), .true., origin='les_step::s', tolerant=.true.)
    if (any(h == 0)) error stop 'an array could not be kept on the device'
    call check(les_run(domain%context, FORT_SCOPE_GPU, h(1), h(2), h(3), h(4), h(5), h(6), h(7), h(8), &
                       dx, dx, dx, dt, cs, nx, ny, nz, 1_c_int))
  end subroutine

  subroutine chain(d, a, b, c, precision)
    use fort_shared_chain_contraction, only: chain_run => run
    real(c_double), dimension(:,:), intent(inout), target :: d, a, b, c
    integer(c_int), intent(in) :: precision
    integer(c_int64_t) :: h(4)
    h(1) = domain%bind_real64(d, lbound(d), .true., origin='chain::d', tolerant=.true.)
    h(2) = domain%bind_real64(a, lbound(a), .true., origin='chain::a', tolerant=.true.)
    h(3) = domain%bind_real64(b, lbound(b), .true., origin='chain::b', tolerant=.true.)
    h(4) = domain%bind_real64(c, lbound(c), .true., origin='chain::c', tolerant=.true.)
    if (any(h == 0)) error stop 'an array could not be kept on the device'
    call check(chain_run(domain%context, FORT_SCOPE_GPU, h(1), h(2), h(3), h(4), int(size(a, 1), c_int), precision))
  end subroutine

  subroutine load(name, values)
    character(*), intent(in) :: name
    real(c_double), intent(out) :: values(..)
    integer :: unit
    open(newunit=unit, file='in/fields-'//name//'.bin', access='stream', form='unformatted', status='old')
    select rank (values)
    rank (2)
      read(unit) values
    rank (3)
      read(unit) values
    end select
    close(unit)
  end subroutine

  ! As the dump template writes a field: the values, then storage_size, shape, lbound and ubound. The
  ! dummy is allocatable so that lbound and ubound are the array's own, as where the template runs.
  subroutine dump(name, values)
    character(*), intent(in) :: name
    real(c_double), allocatable, intent(in) :: values(..)
    integer :: unit
    open(newunit=unit, file='out/fields-'//name//'.bin', access='stream', form='unformatted', status='replace')
    select rank (values)
    rank (2)
      write(unit) values
    rank (3)
      write(unit) values
    end select
    close(unit)
    open(newunit=unit, file='out/fields-'//name//'.shape', status='replace')
    write(unit, *) storage_size(values), shape(values), lbound(values), ubound(values)
    close(unit)
  end subroutine
end module

program resident_entries
  use, intrinsic :: iso_c_binding
  use application
  implicit none
  real(c_double), allocatable, target :: u1(:,:,:), v1(:,:,:), w1(:,:,:), s1(:,:,:)
  real(c_double), allocatable, target :: u2(:,:,:), v2(:,:,:), w2(:,:,:), s2(:,:,:)
  real(c_double), allocatable, target :: a(:,:), b(:,:), c(:,:), d(:,:)
  character(len=32) :: mode, word
  integer :: nx, ny, nz, steps, step, n
  integer(c_int) :: precision
  real(c_double) :: dx, dt, cs

  call get_command_argument(1, mode)
  select case (trim(mode))
  case ('les')
    call get_command_argument(2, word); read(word, *) nx
    call get_command_argument(3, word); read(word, *) ny
    call get_command_argument(4, word); read(word, *) nz
    call get_command_argument(5, word); read(word, *) steps
    call get_command_argument(6, word); read(word, *) dx
    call get_command_argument(7, word); read(word, *) dt
    call get_command_argument(8, word); read(word, *) cs
    allocate(u1(0:nx+1,0:ny+1,0:nz+1), v1(0:nx+1,0:ny+1,0:nz+1), w1(0:nx+1,0:ny+1,0:nz+1), s1(0:nx+1,0:ny+1,0:nz+1))
    allocate(u2(0:nx+1,0:ny+1,0:nz+1), v2(0:nx+1,0:ny+1,0:nz+1), w2(0:nx+1,0:ny+1,0:nz+1), s2(0:nx+1,0:ny+1,0:nz+1))
    call load('U', u1); call load('V', v1); call load('W', w1); call load('S', s1)
    u2 = 0; v2 = 0; w2 = 0; s2 = 0
    ! The epoch keeps every array on the device from the first call to the end.
    call domain%begin_epoch(0_c_int)
    do step = 1, steps
      if (mod(step, 2) == 1) then
        call les_step(u2, v2, w2, s2, u1, v1, w1, s1, dx, dt, cs)
      else
        call les_step(u1, v1, w1, s1, u2, v2, w2, s2, dx, dt, cs)
      end if
    end do
    call domain%end_epoch()
    if (mod(steps, 2) == 1) then
      call dump('U', u2); call dump('V', v2); call dump('W', w2); call dump('S', s2)
    else
      call dump('U', u1); call dump('V', v1); call dump('W', w1); call dump('S', s1)
    end if
  case ('chain')
    call get_command_argument(2, word); read(word, *) n
    call get_command_argument(3, word); read(word, *) precision
    allocate(a(n,n), b(n,n), c(n,n), d(n,n))
    call load('A', a); call load('B', b); call load('C', c)
    d = 0
    call domain%begin_epoch(0_c_int)
    call chain(d, a, b, c, precision)
    call domain%end_epoch()
    call dump('D', d)
  case default
    error stop 'unknown case'
  end select
  call domain%close()
  print '(a)', 'done'
end program
