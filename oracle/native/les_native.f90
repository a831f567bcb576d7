! ORACLE — test infrastructure only: the native CPU build of the synthetic LES step, used as the timed CPU
! baseline (bench.py's cpu_baseline leg) and bit-compared with the C restatement (oracle/step.c) by
! tests/test_native_baseline.py. Nothing in the product path runs it.
!
! One explicit step F_new = F + dt * (advection + Smagorinsky diffusion) of U, V, W (staggered faces) and
! the scalar S (cell centres) on (nx+2, ny+2, nz+2) arrays with one ghost layer, then the halo fill of the
! new fields: periodic in x and y; at the z walls U and V mirrored with a sign change, W zero on and beyond
! the walls (its loop domain is the interior faces k = 2..nz), S zero-gradient.
!
! The arithmetic of every point is the one the C restatement states, operation for operation, so the
! result is bit-identical when both are built without contraction of a*b+c (-ffp-contract=off):
!   advection of u, v, w: the flux-form divergence and advective-form terms of
!     benchmarks/cdu/src/cdu.f90:48-107 and compiler/tests/fixtures/stencils/{cvd,cdw}.f90:44-98,
!     halved; the scalar in flux form with two-point face means (oracle/advection.c:orc_scalar_adv);
!   nu_t = (Cs*Delta)^2 * sqrt(2 S_ij S_ij) at cell centres, ghost values from their images, and the
!     eddy-viscosity diffusion of u, v, w and s with face and edge averages (oracle/sgs.c).
! Every sum is parenthesised in the order the C statement evaluates it. The nests are written here for
! a CPU: each new value is formed in one pass over the grid (the C restatement forms the advection and
! diffusion terms in separate passes over a tendency array); OpenMP splits the k loop.
module les_native
  implicit none
  private
  integer, parameter :: rk = kind(1.0d0)
  public :: rk, sgs_coefficients, les_step, halo_fill

contains

  ! (1/dx, 1/dy, 1/dz, 0.25/dx, 0.25/dy, 0.25/dz, 1/dx^2, 1/dy^2, 1/dz^2, (Cs*Delta)^2), Delta = (dx dy dz)^(1/3)
  subroutine sgs_coefficients(dx, dy, dz, cs, q)
    real(rk), intent(in) :: dx, dy, dz, cs
    real(rk), intent(out) :: q(10)
    real(rk) :: delta
    interface
      real(c_double) function cbrt(x) bind(c, name='cbrt')
        use, intrinsic :: iso_c_binding, only: c_double
        real(c_double), value :: x
      end function cbrt
    end interface
    delta = cbrt(dx * dy * dz)
    q(1) = 1.0_rk / dx
    q(2) = 1.0_rk / dy
    q(3) = 1.0_rk / dz
    q(4) = 0.25_rk / dx
    q(5) = 0.25_rk / dy
    q(6) = 0.25_rk / dz
    q(7) = 1.0_rk / (dx * dx)
    q(8) = 1.0_rk / (dy * dy)
    q(9) = 1.0_rk / (dz * dz)
    q(10) = (cs * delta) * (cs * delta)
  end subroutine sgs_coefficients

  ! nu_t on the padded box: interior cells from the strain rate, every ghost from its image.
  subroutine viscosity(nu, u, v, w, q, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(rk), intent(out) :: nu(nx+2, ny+2, nz+2)
    real(rk), intent(in) :: u(nx+2, ny+2, nz+2), v(nx+2, ny+2, nz+2), w(nx+2, ny+2, nz+2), q(10)
    real(rk) :: s11, s22, s33, dudy, dvdx, dudz, dwdx, dvdz, dwdy, s12, s13, s23, ss
    integer :: i, j, k, ii, jj, kk
    !$omp parallel do private(i, j, s11, s22, s33, dudy, dvdx, dudz, dwdx, dvdz, dwdy, s12, s13, s23, ss)
    do k = 2, nz + 1
      do j = 2, ny + 1
        do i = 2, nx + 1
          s11 = (u(i,j,k) - u(i-1,j,k)) * q(1)
          s22 = (v(i,j,k) - v(i,j-1,k)) * q(2)
          s33 = (w(i,j,k) - w(i,j,k-1)) * q(3)
          dudy = ((u(i,j+1,k) + u(i-1,j+1,k)) - (u(i,j-1,k) + u(i-1,j-1,k))) * q(5)
          dvdx = ((v(i+1,j,k) + v(i+1,j-1,k)) - (v(i-1,j,k) + v(i-1,j-1,k))) * q(4)
          dudz = ((u(i,j,k+1) + u(i-1,j,k+1)) - (u(i,j,k-1) + u(i-1,j,k-1))) * q(6)
          dwdx = ((w(i+1,j,k) + w(i+1,j,k-1)) - (w(i-1,j,k) + w(i-1,j,k-1))) * q(4)
          dvdz = ((v(i,j,k+1) + v(i,j-1,k+1)) - (v(i,j,k-1) + v(i,j-1,k-1))) * q(6)
          dwdy = ((w(i,j+1,k) + w(i,j+1,k-1)) - (w(i,j-1,k) + w(i,j-1,k-1))) * q(5)
          s12 = 0.5_rk * (dudy + dvdx)
          s13 = 0.5_rk * (dudz + dwdx)
          s23 = 0.5_rk * (dvdz + dwdy)
          ss = 2.0_rk * ((s11 * s11 + s22 * s22) + s33 * s33) + 4.0_rk * ((s12 * s12 + s13 * s13) + s23 * s23)
          nu(i,j,k) = q(10) * sqrt(ss)
        end do
      end do
    end do
    !$omp end parallel do
    !$omp parallel do private(i, j, ii, jj, kk)
    do k = 1, nz + 2
      do j = 1, ny + 2
        do i = 1, nx + 2
          ii = merge(nx + 1, merge(2, i, i == nx + 2), i == 1)
          jj = merge(ny + 1, merge(2, j, j == ny + 2), j == 1)
          kk = min(max(k, 2), nz + 1)
          if (ii /= i .or. jj /= j .or. kk /= k) nu(i,j,k) = nu(ii,jj,kk)
        end do
      end do
    end do
    !$omp end parallel do
  end subroutine viscosity

  ! One step (u, v, w, s) -> (un, vn, wn, sn); nu is work space of the padded size. The walls of wn
  ! (faces k = 1 and nz + 1) are written zero by the halo fill.
  subroutine les_step(un, vn, wn, sn, nu, u, v, w, s, dx, dy, dz, dt, q, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(rk), intent(out) :: un(nx+2, ny+2, nz+2), vn(nx+2, ny+2, nz+2), wn(nx+2, ny+2, nz+2), sn(nx+2, ny+2, nz+2)
    real(rk), intent(inout) :: nu(nx+2, ny+2, nz+2)
    real(rk), intent(in) :: u(nx+2, ny+2, nz+2), v(nx+2, ny+2, nz+2), w(nx+2, ny+2, nz+2), s(nx+2, ny+2, nz+2)
    real(rk), intent(in) :: dx, dy, dz, dt, q(10)
    real(rk) :: ax, ay, az, bx, by, bz, hx, hy, hz, t, f1, f2, f3, ua, va, wa
    real(rk) :: fxp, fxm, fyp, fym, fzp, fzm
    integer :: i, j, k

    call viscosity(nu, u, v, w, q, nx, ny, nz)
    ax = 0.25_rk / dx
    ay = 0.25_rk / dy
    az = 0.25_rk / dz
    bx = 0.5_rk / dx
    by = 0.5_rk / dy
    bz = 0.5_rk / dz
    hx = 0.125_rk / dx
    hy = 0.125_rk / dy
    hz = 0.125_rk / dz
    !$omp parallel do private(i, j, t, f1, f2, f3, ua, va, wa, fxp, fxm, fyp, fym, fzp, fzm)
    do k = 2, nz + 1
      do j = 2, ny + 1
        do i = 2, nx + 1
          ! u on the x face: flux divergence, advective form, half, then the diffusion
          f1 = ax * (u(i+1,j,k) + u(i,j,k)) * (u(i+1,j,k) + u(i,j,k)) - ax * (u(i,j,k) + u(i-1,j,k)) * (u(i,j,k) + u(i-1,j,k))
          f2 = ay * (u(i,j+1,k) + u(i,j,k)) * (v(i+1,j,k) + v(i,j,k)) - ay * (u(i,j,k) + u(i,j-1,k)) * (v(i+1,j-1,k) + v(i,j-1,k))
          f3 = az * (u(i,j,k+1) + u(i,j,k)) * (w(i+1,j,k) + w(i,j,k)) - az * (u(i,j,k) + u(i,j,k-1)) * (w(i+1,j,k-1) + w(i,j,k-1))
          t = -((f1 + f2) + f3)
          va = ((v(i,j,k) + v(i+1,j,k)) + v(i,j-1,k)) + v(i+1,j-1,k)
          wa = ((w(i,j,k) + w(i+1,j,k)) + w(i,j,k-1)) + w(i+1,j,k-1)
          t = t - ((bx * (u(i+1,j,k) - u(i-1,j,k)) * u(i,j,k) + hy * (u(i,j+1,k) - u(i,j-1,k)) * va) &
                   + hz * (u(i,j,k+1) - u(i,j,k-1)) * wa)
          t = t * 0.5_rk
          fxp = nu(i+1,j,k) * (u(i+1,j,k) - u(i,j,k))
          fxm = nu(i,j,k) * (u(i,j,k) - u(i-1,j,k))
          fyp = (0.25_rk * (((nu(i,j,k) + nu(i+1,j,k)) + nu(i,j+1,k)) + nu(i+1,j+1,k))) * (u(i,j+1,k) - u(i,j,k))
          fym = (0.25_rk * (((nu(i,j-1,k) + nu(i+1,j-1,k)) + nu(i,j,k)) + nu(i+1,j,k))) * (u(i,j,k) - u(i,j-1,k))
          fzp = (0.25_rk * (((nu(i,j,k) + nu(i+1,j,k)) + nu(i,j,k+1)) + nu(i+1,j,k+1))) * (u(i,j,k+1) - u(i,j,k))
          fzm = (0.25_rk * (((nu(i,j,k-1) + nu(i+1,j,k-1)) + nu(i,j,k)) + nu(i+1,j,k))) * (u(i,j,k) - u(i,j,k-1))
          t = t + (((fxp - fxm) * q(7) + (fyp - fym) * q(8)) + (fzp - fzm) * q(9))
          un(i,j,k) = u(i,j,k) + dt * t

          ! v on the y face
          f2 = ay * (v(i,j+1,k) + v(i,j,k)) * (v(i,j+1,k) + v(i,j,k)) - ay * (v(i,j,k) + v(i,j-1,k)) * (v(i,j,k) + v(i,j-1,k))
          f1 = ax * (v(i+1,j,k) + v(i,j,k)) * (u(i,j+1,k) + u(i,j,k)) - ax * (v(i,j,k) + v(i-1,j,k)) * (u(i-1,j+1,k) + u(i-1,j,k))
          f3 = az * (v(i,j,k+1) + v(i,j,k)) * (w(i,j+1,k) + w(i,j,k)) - az * (v(i,j,k) + v(i,j,k-1)) * (w(i,j+1,k-1) + w(i,j,k-1))
          t = -((f2 + f1) + f3)
          ua = ((u(i,j,k) + u(i,j+1,k)) + u(i-1,j,k)) + u(i-1,j+1,k)
          wa = ((w(i,j,k) + w(i,j+1,k)) + w(i,j,k-1)) + w(i,j+1,k-1)
          t = t - ((hx * (v(i+1,j,k) - v(i-1,j,k)) * ua + by * (v(i,j+1,k) - v(i,j-1,k)) * v(i,j,k)) &
                   + hz * (v(i,j,k+1) - v(i,j,k-1)) * wa)
          t = t * 0.5_rk
          fxp = (0.25_rk * (((nu(i,j,k) + nu(i+1,j,k)) + nu(i,j+1,k)) + nu(i+1,j+1,k))) * (v(i+1,j,k) - v(i,j,k))
          fxm = (0.25_rk * (((nu(i-1,j,k) + nu(i,j,k)) + nu(i-1,j+1,k)) + nu(i,j+1,k))) * (v(i,j,k) - v(i-1,j,k))
          fyp = nu(i,j+1,k) * (v(i,j+1,k) - v(i,j,k))
          fym = nu(i,j,k) * (v(i,j,k) - v(i,j-1,k))
          fzp = (0.25_rk * (((nu(i,j,k) + nu(i,j+1,k)) + nu(i,j,k+1)) + nu(i,j+1,k+1))) * (v(i,j,k+1) - v(i,j,k))
          fzm = (0.25_rk * (((nu(i,j,k-1) + nu(i,j+1,k-1)) + nu(i,j,k)) + nu(i,j+1,k))) * (v(i,j,k) - v(i,j,k-1))
          t = t + (((fxp - fxm) * q(7) + (fyp - fym) * q(8)) + (fzp - fzm) * q(9))
          vn(i,j,k) = v(i,j,k) + dt * t

          ! w on the interior z faces
          if (k <= nz) then
            f3 = az * (w(i,j,k+1) + w(i,j,k)) * (w(i,j,k+1) + w(i,j,k)) - az * (w(i,j,k) + w(i,j,k-1)) * (w(i,j,k) + w(i,j,k-1))
            f2 = ay * (w(i,j+1,k) + w(i,j,k)) * (v(i,j,k+1) + v(i,j,k)) - ay * (w(i,j,k) + w(i,j-1,k)) * (v(i,j-1,k) + v(i,j-1,k+1))
            f1 = ax * (w(i+1,j,k) + w(i,j,k)) * (u(i,j,k+1) + u(i,j,k)) - ax * (w(i,j,k) + w(i-1,j,k)) * (u(i-1,j,k+1) + u(i-1,j,k))
            t = -((f3 + f2) + f1)
            ua = ((u(i,j,k) + u(i,j,k+1)) + u(i-1,j,k)) + u(i-1,j,k+1)
            va = ((v(i,j,k) + v(i,j,k+1)) + v(i,j-1,k)) + v(i,j-1,k+1)
            t = t - ((hx * (w(i+1,j,k) - w(i-1,j,k)) * ua + hy * (w(i,j+1,k) - w(i,j-1,k)) * va) &
                     + bz * (w(i,j,k+1) - w(i,j,k-1)) * w(i,j,k))
            t = t * 0.5_rk
            fxp = (0.25_rk * (((nu(i,j,k) + nu(i+1,j,k)) + nu(i,j,k+1)) + nu(i+1,j,k+1))) * (w(i+1,j,k) - w(i,j,k))
            fxm = (0.25_rk * (((nu(i-1,j,k) + nu(i,j,k)) + nu(i-1,j,k+1)) + nu(i,j,k+1))) * (w(i,j,k) - w(i-1,j,k))
            fyp = (0.25_rk * (((nu(i,j,k) + nu(i,j+1,k)) + nu(i,j,k+1)) + nu(i,j+1,k+1))) * (w(i,j+1,k) - w(i,j,k))
            fym = (0.25_rk * (((nu(i,j-1,k) + nu(i,j,k)) + nu(i,j-1,k+1)) + nu(i,j,k+1))) * (w(i,j,k) - w(i,j-1,k))
            fzp = nu(i,j,k+1) * (w(i,j,k+1) - w(i,j,k))
            fzm = nu(i,j,k) * (w(i,j,k) - w(i,j,k-1))
            t = t + (((fxp - fxm) * q(7) + (fyp - fym) * q(8)) + (fzp - fzm) * q(9))
            wn(i,j,k) = w(i,j,k) + dt * t
          end if

          ! the scalar
          f1 = bx * u(i,j,k) * (s(i+1,j,k) + s(i,j,k)) - bx * u(i-1,j,k) * (s(i,j,k) + s(i-1,j,k))
          f2 = by * v(i,j,k) * (s(i,j+1,k) + s(i,j,k)) - by * v(i,j-1,k) * (s(i,j,k) + s(i,j-1,k))
          f3 = bz * w(i,j,k) * (s(i,j,k+1) + s(i,j,k)) - bz * w(i,j,k-1) * (s(i,j,k) + s(i,j,k-1))
          t = -((f1 + f2) + f3)
          fxp = (0.5_rk * (nu(i,j,k) + nu(i+1,j,k))) * (s(i+1,j,k) - s(i,j,k))
          fxm = (0.5_rk * (nu(i-1,j,k) + nu(i,j,k))) * (s(i,j,k) - s(i-1,j,k))
          fyp = (0.5_rk * (nu(i,j,k) + nu(i,j+1,k))) * (s(i,j+1,k) - s(i,j,k))
          fym = (0.5_rk * (nu(i,j-1,k) + nu(i,j,k))) * (s(i,j,k) - s(i,j-1,k))
          fzp = (0.5_rk * (nu(i,j,k) + nu(i,j,k+1))) * (s(i,j,k+1) - s(i,j,k))
          fzm = (0.5_rk * (nu(i,j,k-1) + nu(i,j,k))) * (s(i,j,k) - s(i,j,k-1))
          t = t + (((fxp - fxm) * q(7) + (fyp - fym) * q(8)) + (fzp - fzm) * q(9))
          sn(i,j,k) = s(i,j,k) + dt * t
        end do
      end do
    end do
    !$omp end parallel do
    call halo_fill(un, 0, nx, ny, nz)
    call halo_fill(vn, 0, nx, ny, nz)
    call halo_fill(wn, 1, nx, ny, nz)
    call halo_fill(sn, 2, nx, ny, nz)
  end subroutine les_step

  ! Ghost layer of one field: x faces, then y faces over the whole x range, then z faces over the whole
  ! x-y range, so edges and corners follow. kind 0: tangential velocity, 1: wall-normal velocity, 2: scalar.
  subroutine halo_fill(f, kind, nx, ny, nz)
    integer, intent(in) :: kind, nx, ny, nz
    real(rk), intent(inout) :: f(nx+2, ny+2, nz+2)
    integer :: j, k
    !$omp parallel do private(j)
    do k = 2, nz + 1
      do j = 2, ny + 1
        f(1,j,k) = f(nx+1,j,k)
        f(nx+2,j,k) = f(2,j,k)
      end do
      f(:,1,k) = f(:,ny+1,k)
      f(:,ny+2,k) = f(:,2,k)
    end do
    !$omp end parallel do
    select case (kind)
    case (0)
      f(:,:,1) = -f(:,:,2)
      f(:,:,nz+2) = -f(:,:,nz+1)
    case (1)
      f(:,:,1) = 0.0_rk
      f(:,:,nz+1) = 0.0_rk
      f(:,:,nz+2) = 0.0_rk
    case default
      f(:,:,1) = f(:,:,2)
      f(:,:,nz+2) = f(:,:,nz+1)
    end select
  end subroutine halo_fill
end module les_native

! les_native NX NY NZ STEPS REPEATS INPUT [OUTPUT [FINAL]]
! Reads U, V, W, S (each (NX+2)(NY+2)(NZ+2) REAL64 values, Fortran order, one after another) from INPUT,
! dx = dy = dz = 10 m, dt = 0.2 s, Cs = 0.17. Runs one warm-up step (written to OUTPUT, if given, for the
! bit comparison), then REPEATS repetitions of STEPS steps, each timed with system_clock around its step
! loop, and prints the threads and every repetition's ms per step, then the median. FINAL, if given,
! receives the fields after all 1 + REPEATS*STEPS steps.
program les_native_bench
  use les_native
  use omp_lib, only: omp_get_max_threads
  implicit none
  integer :: nx, ny, nz, steps, repeats, r, n, unit, idx
  character(len=512) :: arg, input, output, final
  real(rk), allocatable, target :: f(:,:,:,:,:)
  real(rk), allocatable :: nu(:,:,:), times(:)
  real(rk) :: q(10), swap
  integer(kind=8) :: t0, t1, rate
  integer :: cur

  call get_command_argument(1, arg); read(arg, *) nx
  call get_command_argument(2, arg); read(arg, *) ny
  call get_command_argument(3, arg); read(arg, *) nz
  call get_command_argument(4, arg); read(arg, *) steps
  call get_command_argument(5, arg); read(arg, *) repeats
  call get_command_argument(6, input)
  output = ''
  final = ''
  if (command_argument_count() >= 7) call get_command_argument(7, output)
  if (command_argument_count() >= 8) call get_command_argument(8, final)

  ! f(:,:,:,field,copy): fields U, V, W, S of the current and the next state
  allocate(f(nx+2, ny+2, nz+2, 4, 2), nu(nx+2, ny+2, nz+2), times(max(repeats, 1)))
  f = 0.0_rk
  open(newunit=unit, file=trim(input), access='stream', form='unformatted', status='old')
  read(unit) f(:,:,:,:,1)
  close(unit)
  call sgs_coefficients(10.0_rk, 10.0_rk, 10.0_rk, 0.17_rk, q)

  cur = 1
  call les_step(f(:,:,:,1,3-cur), f(:,:,:,2,3-cur), f(:,:,:,3,3-cur), f(:,:,:,4,3-cur), nu, &
                f(:,:,:,1,cur), f(:,:,:,2,cur), f(:,:,:,3,cur), f(:,:,:,4,cur), &
                10.0_rk, 10.0_rk, 10.0_rk, 0.2_rk, q, nx, ny, nz)
  cur = 3 - cur
  if (len_trim(output) > 0) then
    open(newunit=unit, file=trim(output), access='stream', form='unformatted', status='replace')
    write(unit) f(:,:,:,:,cur)
    close(unit)
  end if

  do r = 1, repeats
    call system_clock(t0, rate)
    do n = 1, steps
      call les_step(f(:,:,:,1,3-cur), f(:,:,:,2,3-cur), f(:,:,:,3,3-cur), f(:,:,:,4,3-cur), nu, &
                    f(:,:,:,1,cur), f(:,:,:,2,cur), f(:,:,:,3,cur), f(:,:,:,4,cur), &
                    10.0_rk, 10.0_rk, 10.0_rk, 0.2_rk, q, nx, ny, nz)
      cur = 3 - cur
    end do
    call system_clock(t1)
    times(r) = 1000.0_rk * real(t1 - t0, rk) / real(rate, rk) / real(max(steps, 1), rk)
  end do

  write(*, '(a,i0)') 'threads ', omp_get_max_threads()
  do r = 1, repeats
    write(*, '(a,i0,a,f0.4)') 'repeat ', r, ' ms_per_step ', times(r)
  end do
  ! median (insertion sort of the repeats)
  do r = 2, repeats
    swap = times(r)
    idx = r - 1
    do while (idx >= 1)
      if (times(idx) <= swap) exit
      times(idx + 1) = times(idx)
      idx = idx - 1
    end do
    times(idx + 1) = swap
  end do
  if (repeats > 0) write(*, '(a,f0.4)') 'median_ms_per_step ', times((repeats + 1) / 2)
  if (len_trim(final) > 0) then
    open(newunit=unit, file=trim(final), access='stream', form='unformatted', status='replace')
    write(unit) f(:,:,:,:,cur)
    close(unit)
  end if
end program les_native_bench
