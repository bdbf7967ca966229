! A Fortran caller of the library's Fortran-callable transposes, written as
! the reference's Fortran driver calls its own (OpenMP/mhd.f90:410-474;
! the GPU form HIP/cu/cuda_transpose12.cu:69-127 takes host arrays, every
! argument by reference): `call mhd_transpose12(...)` links to the C-ABI
! symbol mhd_transpose12_.  The state u(5,nx,ny,nz), b(3,nx,ny,nz) is the
! library's AoS layout as it is.  Checks the component maps of
! Heterogeneous_MHD/C/transpose12.c:22-29 and transpose13.c (u 1,3,2,4,5 /
! 1,4,3,2,5; b 2,1,3 / 3,2,1) and that each transpose is an involution.
program transpose_f
  implicit none
  integer, parameter :: nx = 6, ny = 4, nz = 5
  real :: u(5, nx, ny, nz), b(3, nx, ny, nz), dt
  real :: ut(5, ny, nx, nz), bt(3, ny, nx, nz), u2(5, nx, ny, nz), b2(3, nx, ny, nz)
  real :: uz(5, nz, ny, nx), bz(3, nz, ny, nx)
  integer :: i, j, k, c, nbad, mx, my, mz
  integer :: p12u(5) = (/1, 3, 2, 4, 5/), p12b(3) = (/2, 1, 3/)
  integer :: p13u(5) = (/1, 4, 3, 2, 5/), p13b(3) = (/3, 2, 1/)

  do k = 1, nz
    do j = 1, ny
      do i = 1, nx
        do c = 1, 5
          u(c, i, j, k) = real(c) + 10.0 * i + 100.0 * j + 1000.0 * k
        end do
        do c = 1, 3
          b(c, i, j, k) = -(real(c) + 10.0 * i + 100.0 * j + 1000.0 * k)
        end do
      end do
    end do
  end do
  dt = 0.0
  nbad = 0

  mx = nx; my = ny; mz = nz
  call mhd_transpose12(ut, bt, u, b, mx, my, mz, dt)
  do k = 1, nz
    do j = 1, ny
      do i = 1, nx
        do c = 1, 5
          if (ut(c, j, i, k) /= u(p12u(c), i, j, k)) nbad = nbad + 1
        end do
        do c = 1, 3
          if (bt(c, j, i, k) /= b(p12b(c), i, j, k)) nbad = nbad + 1
        end do
      end do
    end do
  end do
  ! back: the (ny,nx,nz) frame transposed again is the original state
  mx = ny; my = nx; mz = nz
  call mhd_transpose12(u2, b2, ut, bt, mx, my, mz, dt)
  if (any(u2 /= u) .or. any(b2 /= b)) nbad = nbad + 1

  ! transpose13: (i,j,k) -> (k,j,i), any shape here
  mx = nx; my = ny; mz = nz
  call mhd_transpose13(uz, bz, u, b, mx, my, mz, dt)
  do k = 1, nz
    do j = 1, ny
      do i = 1, nx
        do c = 1, 5
          if (uz(c, k, j, i) /= u(p13u(c), i, j, k)) nbad = nbad + 1
        end do
        do c = 1, 3
          if (bz(c, k, j, i) /= b(p13b(c), i, j, k)) nbad = nbad + 1
        end do
      end do
    end do
  end do

  print '(a,i0)', 'transpose_f: mismatches ', nbad
  if (nbad /= 0) stop 1
end program transpose_f
