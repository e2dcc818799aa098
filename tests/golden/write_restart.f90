! Writes a small restart directory with the Fortran I/O statements SALMON uses for its
! restart data (src/io/checkpoint_restart.f90: info.bin and occupation.bin sequential
! unformatted, wfn.bin stream access in (ix,iy,iz,is,io,ik) order, Vh_stock.bin two
! sequential records, rtdata.bin the three records of write_rtdata), with values given by
! closed formulas the test recomputes.
! Usage: write_restart <directory> [real|complex [nspin]]
program write_restart
  implicit none
  integer, parameter :: nx = 5, ny = 4, nz = 3, no = 3, nk = 2
  integer :: nspin
  complex(8), allocatable :: zwf(:, :, :, :, :, :)
  real(8), allocatable :: rwf(:, :, :, :, :, :), rocc(:, :, :)
  real(8) :: vh1(nx, ny, nz), vh2(nx, ny, nz)
  real(8) :: ac_tot(3, 0:1236), ac_ind(3, 0:1236)
  integer :: ix, iy, iz, is, io, ik, itt
  character(256) :: dir, mode, arg3
  logical :: if_real

  call get_command_argument(1, dir)
  call get_command_argument(2, mode)
  call get_command_argument(3, arg3)
  if_real = trim(mode) == 'real'
  nspin = 1
  if (len_trim(arg3) > 0) read(arg3, *) nspin
  allocate(zwf(nx, ny, nz, nspin, no, nk), rwf(nx, ny, nz, nspin, no, nk), rocc(no, nk, nspin))
  do ik = 1, nk
  do io = 1, no
  do is = 1, nspin
  do iz = 1, nz
  do iy = 1, ny
  do ix = 1, nx
    zwf(ix, iy, iz, is, io, ik) = cmplx(ix + 10*iy + 100*iz + 0.5d0*is, io + 10*ik + 0.25d0*(ix - iy), kind=8)
    rwf(ix, iy, iz, is, io, ik) = ix - 2*iy + 3*iz + 0.125d0*io - 7*ik
  end do
  end do
  end do
  rocc(io, ik, is) = 2d0 - 0.1d0*io - 0.01d0*ik - 0.001d0*(is - 1)
  end do
  end do
  end do
  do iz = 1, nz
  do iy = 1, ny
  do ix = 1, nx
    vh1(ix, iy, iz) = 0.5d0*ix - iy + 0.25d0*iz
    vh2(ix, iy, iz) = -vh1(ix, iy, iz) + 1d0/3d0
  end do
  end do
  end do

  open(97, file=trim(dir)//'/info.bin', form='unformatted')
  write(97) nk
  write(97) no
  write(97) 1234
  write(97) 4
  write(97) if_real
  close(97)
  open(97, file=trim(dir)//'/occupation.bin', form='unformatted')
  write(97) rocc(1:no, 1:nk, 1:nspin)
  close(97)
  open(87, file=trim(dir)//'/wfn.bin', form='unformatted', access='stream')
  do ik = 1, nk
  do io = 1, no
  do is = 1, nspin
    if (if_real) then
      write(87) rwf(1:nx, 1:ny, 1:nz, is, io, ik)
    else
      write(87) zwf(1:nx, 1:ny, 1:nz, is, io, ik)
    end if
  end do
  end do
  end do
  close(87)
  open(97, file=trim(dir)//'/Vh_stock.bin', form='unformatted')
  write(97) vh1
  write(97) vh2
  close(97)
  ! write_rtdata (checkpoint_restart.f90:1925-1957) at itt = 1234
  do itt = 0, 1236
    do ix = 1, 3
      ac_tot(ix, itt) = 1d-3 * itt + 0.1d0 * ix
      ac_ind(ix, itt) = -2d-4 * itt + 0.01d0 * ix * ix
    end do
  end do
  itt = 1234
  open(97, file=trim(dir)//'/rtdata.bin', form='unformatted')
  write(97) ac_tot(:, itt), ac_tot(:, itt + 1)
  write(97) ac_ind(:, itt), ac_ind(:, itt + 1)
  write(97) ac_ind(:, 0)
  close(97)
end program write_restart
