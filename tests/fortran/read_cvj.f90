! Reads a per-step record file with the statement form of the reference's
! writes of Cvj.LLLL / Cvjprod.LLLL ("write(15) it,Cvj",
! SRC/main_madwave3.f:353-365): "read(15) it, cvj" until the end of the
! file, cvj dimensioned (nv, nj, nch) as Cvj(nvini:nvmax, jini:jmax, ncan).
! dims.txt holds nv, nj, nch and the file name; prints per record it, the
! sum of the values and the elements (1,1,1), (2,1,1), (1,2,1), (nv,nj,nch).
program read_cvj
  implicit none
  integer :: nv, nj, nch, it, ios
  character(len=64) :: name
  real*8, allocatable :: cvj(:, :, :)
  open(10, file='dims.txt', status='old')
  read(10, *) nv, nj, nch, name
  close(10)
  allocate(cvj(nv, nj, nch))
  open(15, file=trim(name), status='old', form='unformatted')
  do
    read(15, iostat=ios) it, cvj
    if (ios /= 0) exit
    write(6, '(i8, 5es25.16)') it, sum(cvj), cvj(1, 1, 1), cvj(2, 1, 1), cvj(1, 2, 1), &
        cvj(nv, nj, nch)
  end do
  close(15)
end program read_cvj
