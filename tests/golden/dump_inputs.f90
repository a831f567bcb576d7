! Writes the inputs the reference stencil drivers (compiler/tests/fixtures/stencils/cdu_main.f90:38-49,
! identical in cvd_main.f90 and cdw_main.f90) build, as raw float64 in Fortran order: U, then V, then W,
! each (NX+2, NY+2, NZ+2). The expressions are those of the drivers, so the same compiler gives the same
! bytes the drivers fed to the kernels.
program dump_inputs
    implicit none
    integer, parameter :: knd = kind(1.0d0)
    integer, parameter :: NX = 16, NY = 16, NZ = 16
    real(knd) :: U(NX+2, NY+2, NZ+2), V(NX+2, NY+2, NZ+2), W(NX+2, NY+2, NZ+2)
    integer :: i, j, k, unit
    do k = 1, NZ+2
        do j = 1, NY+2
            do i = 1, NX+2
                U(i,j,k) = sin(real(i,knd)*0.3_knd) * cos(real(j,knd)*0.5_knd) &
                          * (1.0_knd + 0.1_knd*real(k,knd))
                V(i,j,k) = cos(real(i,knd)*0.7_knd) * sin(real(k,knd)*0.4_knd) &
                          * (1.0_knd + 0.1_knd*real(j,knd))
                W(i,j,k) = sin(real(j,knd)*0.6_knd + real(k,knd)*0.2_knd)
            end do
        end do
    end do
    open(newunit=unit, file='stencil_inputs.bin', access='stream', form='unformatted', status='replace')
    write(unit) U, V, W
    close(unit)
end program dump_inputs
