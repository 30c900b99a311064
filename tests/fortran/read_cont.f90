! Reads continuation files the way the reference's start-up does
! (SRC/main_madwave3.f:223-261, SRC/mod_flux_01y2.f:298-322,
! SRC/mod_coortrans01_02.f:180-196): list-directed cont.data, unformatted
! sequential cont00.paq / cont00.S2 / cont00.S2prod with whole-array reads.
! Sizes come from dims.txt (ntot, netot, ns2, nflux, nprod); prints the
! counters and the sums of every array for the test to compare.
program read_cont
  implicit none
  integer :: ntot, netot, ns2, nflux, nprod, it0, indt, iloop, nfl, mcan
  real*8 :: abstot
  real*8, allocatable :: rflanz0(:), rpaq(:)
  complex*16, allocatable :: zs2(:, :), zcr(:, :), zcdr(:, :), zs2prod(:, :)
  open(10, file='dims.txt', status='old')
  read(10, *) ntot, netot, ns2, nflux, nprod
  close(10)
  allocate(rflanz0(ntot), rpaq(ntot), zs2(netot, ns2), zcr(netot, nflux), zcdr(netot, nflux))
  open(3, file='cont.data', status='old')
  read(3, *) it0, indt, iloop
  read(3, *) abstot, nfl, mcan
  close(3)
  open(4, file='cont00.paq', status='old', form='unformatted')
  read(4) rflanz0
  read(4) rpaq
  close(4)
  open(4, file='cont00.S2', status='old', form='unformatted')
  read(4) zs2
  read(4) zcr
  read(4) zcdr
  close(4)
  write(6, '(5i12)') it0, indt, iloop, nfl, mcan
  write(6, '(2es25.16)') sum(rflanz0), sum(rpaq)
  write(6, '(2es25.16)') sum(dble(zs2)), sum(dimag(zs2))
  write(6, '(2es25.16)') sum(dble(zcr)), sum(dimag(zcdr))
  write(6, '(2es25.16)') dble(zs2(netot, ns2)), dimag(zcr(1, nflux))
  if (nprod > 0) then
    allocate(zs2prod(netot, nprod))
    open(4, file='cont00.S2prod', status='old', form='unformatted')
    read(4) zs2prod
    close(4)
    write(6, '(2es25.16)') sum(dble(zs2prod)), sum(dimag(zs2prod))
  end if
end program read_cont
