! A program written against the reference's Fortran host layer (module fort_resident_domain,
! compiler/runtime/resident_domain.f90, and the bind(C) interfaces of fort_scoped_memory,
! compiler/runtime/scoped_memory.f90), linked against this library's runtime on its CPU backend, where
! "device" memory is host memory that only the runtime's transfers keep in step with the arrays.
!
! The first argument selects a case; each prints "key value" lines.
!   epochs     binding, host and device accesses, transfers, field mirrors, workspace slots, two epochs
!   undefined  a device read of values that were never defined
!   foreign    a host access through a handle the runtime never issued (the layer stops the program)
program resident_host
  use, intrinsic :: iso_c_binding
  use fort_scoped_memory
  use fort_resident_domain
  implicit none
  type :: particle
    real(c_double) :: x
    integer(c_int) :: id
  end type
  type(resident_domain) :: owner
  real(c_double), allocatable, target :: u(:,:,:), other(:)
  integer(c_int), allocatable, target :: kinds(:,:)
  type(particle), allocatable, target :: parts(:)
  real(c_double), pointer, contiguous :: work(:,:,:)
  real(c_double), pointer :: dev(:)
  type(fort_scope_access) :: access
  type(fort_scope_stats) :: stats
  type(c_ptr) :: device, first_work
  integer(c_int64_t) :: hu, hk, hp, ho
  integer(c_int) :: status
  character(len=16) :: mode
  integer :: i

  call get_command_argument(1, mode)
  allocate(u(-1:2,0:2,1:2), kinds(3,2), other(5), parts(3))
  u = reshape([(real(i, c_double), i = 1, 24)], shape(u))
  kinds = 7
  other = 0
  do i = 1, 3
    parts(i) = particle(10.0d0*i, i)
  end do
  call owner%begin_epoch(0_c_int, 1048576_c_size_t)
  print '(a,1x,i0)', 'generation', owner%generation

  select case (trim(mode))
  case ('epochs')
    hu = owner%bind_real64(u, lbound(u), .true., origin='resident_host::u')
    hk = owner%bind_integer32(kinds, lbound(kinds), .true., origin='resident_host::kinds')
    ! A section names the storage of its whole array; an unbound array names none.
    print '(a,1x,l1,1x,l1,1x,l1)', 'held', owner%held(u) == hu, owner%held(u(:,:,2)) == hu, owner%held(other) == 0
    print '(a,1x,l1)', 'distinct', hu /= hk .and. hu /= 0 .and. hk /= 0
    print '(a,1x,l1)', 'rebind', owner%bind_real64(u, lbound(u), .true.) == hu

    ! The device overwrites u: nothing is uploaded, and the host keeps its values until observed.
    access%flags = ior(FORT_SCOPE_WRITE_ALL, FORT_SCOPE_OVERWRITE_ALL)
    call require(fort_scope_device_begin(owner%context, hu, access, device))
    call c_f_pointer(device, dev, [24])
    dev = [(100.0d0 + i, i = 1, 24)]
    call require(fort_scope_device_end(owner%context, hu))
    print '(a,1x,f0.1)', 'host_before_observe', sum(u)
    call owner%observe([hu, hu, 0_c_int64_t])
    print '(a,1x,f0.1)', 'host_after_observe', sum(u)

    ! Host statements change u; the next device read sees the change.
    call owner%host_modify([hu])
    u = u + 1
    call owner%host_done([hu])
    access%flags = FORT_SCOPE_READ_ALL
    call require(fort_scope_device_begin(owner%context, hu, access, device))
    call c_f_pointer(device, dev, [24])
    print '(a,1x,f0.1)', 'device_after_host_change', sum(dev)
    call require(fort_scope_device_end(owner%context, hu))
    call require(fort_scope_stats_get(owner%context, stats))
    print '(a,4(1x,i0))', 'transfers', stats%uploads, stats%downloads, stats%upload_bytes, stats%download_bytes

    ! One field of an array of derived type, kept through a copy of the field alone.
    hp = owner%bind_field(c_loc(parts(1)%x), int(storage_size(parts(1))/8, c_size_t), size(parts, kind=c_size_t), &
                          FORT_SCOPE_REAL64, .true., origin='resident_host::parts%x')
    access%flags = ior(FORT_SCOPE_READ_ALL, FORT_SCOPE_WRITE_ALL)
    call require(fort_scope_device_begin(owner%context, hp, access, device))
    call c_f_pointer(device, dev, [3])
    print '(a,3(1x,f0.1))', 'field_on_device', dev
    dev = dev + 0.5d0
    call require(fort_scope_device_end(owner%context, hp))
    print '(a,3(1x,f0.1))', 'field_before_observe', parts%x
    call owner%observe_field(c_loc(parts(1)%x))
    print '(a,3(1x,f0.1),3(1x,i0))', 'field_after_observe', parts%x, parts%id

    ! A workspace slot: bound uninitialized and persistent, recycled when the site asks again.
    call owner%acquire_work_real64_3(7_c_int64_t, [1, 1, 1], [4, 3, 2], work)
    first_work = c_loc(work)
    print '(a,3(1x,i0))', 'work_bounds', ubound(work)
    call owner%release_work_real64_3(7_c_int64_t, work)
    call owner%end_epoch()
    print '(a,1x,f0.1,1x,l1)', 'published', sum(u), owner%held(u) == 0

    call owner%begin_epoch(0_c_int, 1048576_c_size_t)
    print '(a,1x,i0)', 'generation', owner%generation
    call owner%acquire_work_real64_3(7_c_int64_t, [1, 1, 1], [4, 3, 2], work)
    print '(a,1x,l1)', 'work_reused', c_associated(c_loc(work), first_work)
    call owner%release_work_real64_3(7_c_int64_t, work)
  case ('undefined')
    ho = owner%bind_real64(other, lbound(other), .false., origin='resident_host::other')
    access%flags = FORT_SCOPE_READ_ALL
    status = fort_scope_device_begin(owner%context, ho, access, device)
    print '(a,1x,i0,1x,l1)', 'status', status, status == FORT_SCOPE_UNINITIALIZED
    print '(a,1x,a)', 'message', trim(message())
  case ('foreign')
    access%flags = FORT_SCOPE_READ_ALL
    call owner%host_begin(987654321_c_int64_t, access)
  end select
  call owner%close()
  print '(a)', 'closed'

contains
  subroutine require(code)
    integer(c_int), intent(in) :: code
    if (code /= FORT_SCOPE_OK) then
      print '(a,1x,i0,1x,a)', 'failed', code, trim(message())
      error stop 1
    end if
  end subroutine

  function message() result(text)
    character(len=512) :: text
    character(kind=c_char), pointer :: chars(:)
    type(c_ptr) :: p
    integer :: n
    text = ''
    p = fort_scope_error()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, chars, [512])
    n = 0
    do while (n < 511)
      if (chars(n + 1) == c_null_char) exit
      n = n + 1
      text(n:n) = chars(n)
    end do
  end function
end program
