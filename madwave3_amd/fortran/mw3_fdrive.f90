! mw3_fdrive.f90 -- a Fortran host of libmw3 through module mw3_capi.
!
! Usage:  mw3_fdrive <case file> <result file>
!
! The case file (stream, native byte order) holds, in this order:
!   4 bytes 'MW3D', int32 format version (1);
!   every member of struct mw3_setup in declaration order: int32 and double
!     members as their value, pointer members as an int64 element count
!     followed by the elements (int32 or double);
!   int64 n, double psi0(n)   the initial packet in A0 layout;
!   int32 nsteps, int32 precision (MW3_FP64 or MW3_MIXED);
!   int32 hasprod, and when 1 every member of struct mw3_prod_setup in the
!     same encoding (the product-projection tables, iprod = 2).
! The program runs mw3_cheby_begin (after mw3_prod_attach when hasprod) and
! nsteps steps of mw3_cheby_steps and writes to the result file (stream):
! int32 nsteps, double norms(nsteps), int32 netot, double P(netot)
! (mw3_flux_result), int64 m, double zS2(m) (mw3_zs2_get, interleaved
! complex), int64 mp, double zS2prod(mp) (mw3_zs2prod_get; mp = 0 without
! product tables).  It is the Fortran side of the C ABI exercised end to end;
! errors stop with the library's message.
program mw3_fdrive
  use, intrinsic :: iso_c_binding
  use mw3_capi
  implicit none
  type(mw3_setup), target :: s
  type(mw3_prod_setup), target :: ps
  integer(c_int32_t), allocatable, target :: nobcprod(:)
  real(c_double), allocatable, target :: transrgjac(:), prodwf(:), dtrans(:), zs2prod(:)
  integer(c_int32_t) :: hasprod
  integer(c_int64_t) :: mp
  integer(c_int32_t), allocatable, target :: iomdiat(:), iomatom(:), indangreal(:), &
      npunreal(:), indr1(:), indr2(:), iombas(:), nelebas(:), nojbas(:), j00(:), &
      jbas(:), noBCstates(:)
  real(c_double), allocatable, target :: sigdiat(:), sigatom(:), VVV(:), Djmmp(:), &
      absfr1(:), absfr2(:), etots2(:), reacfct(:), fd(:), efe(:), psi0(:), norms(:), &
      ptot(:), zs2(:)
  character(len=4096) :: fin, fout
  character(len=4) :: magic
  integer(c_int32_t) :: ver, nsteps, prec, rc, nj, nv
  integer(c_int64_t) :: n, m
  type(c_ptr) :: ctx
  integer :: u

  if (command_argument_count() /= 2) then
    write(6, *) 'usage: mw3_fdrive <case file> <result file>'
    stop 2
  end if
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) magic, ver
  if (magic /= 'MW3D' .or. ver /= 2) then
    write(6, *) 'mw3_fdrive: not a version-2 case file: ', trim(fin)
    stop 2
  end if
  read(u) s%abi_version
  read(u) s%npun1, s%npun1min, s%npun2, s%nangu, s%inc
  read(u) s%Jtot, s%iparity, s%iommin, s%iommax, s%j0
  read(u) s%nelecmax
  read(u) s%rmis1, s%ah1, s%rmis2, s%ah2, s%hbr
  read(u) s%xm1red, s%xm2red, s%radcutmax, s%rotcutmax, s%emindlt, s%delta2
  read(u) s%proj_j2_up
  call rdi(iomdiat, s%iomdiat)
  call rdi(iomatom, s%iomatom)
  call rdd(sigdiat, s%sigdiat)
  call rdd(sigatom, s%sigatom)
  read(u) s%npuntot
  call rdi(indangreal, s%indangreal)
  call rdi(npunreal, s%npunreal)
  call rdi(indr1, s%indr1)
  call rdi(indr2, s%indr2)
  call rdd(VVV, s%VVV)
  read(u) s%ncan
  call rdi(iombas, s%iombas)
  call rdi(nelebas, s%nelebas)
  call rdi(nojbas, s%nojbas)
  call rdi(j00, s%j00)
  call rdi(jbas, s%jbas)
  read(u) s%ndj
  call rdd(Djmmp, s%Djmmp)
  call rdd(absfr1, s%absfr1)
  call rdd(absfr2, s%absfr2)
  read(u) s%ir2abs
  read(u) s%netot
  call rdd(etots2, s%etots2)
  call rdd(reacfct, s%reacfct)
  read(u) s%r1flux
  read(u) s%nvini, s%nvmax, s%jini, s%jmax
  call rdi(noBCstates, s%noBCstates)
  call rdd(fd, s%fd)
  read(u) s%iphoto, s%iprod, s%kminelastic
  read(u) s%icanref
  read(u) s%nvref, s%jref, s%iomref
  read(u) s%nofe
  call rdd(efe, s%efe)
  read(u) n
  allocate(psi0(n))
  read(u) psi0
  read(u) nsteps, prec
  read(u) hasprod
  if (hasprod == 1) then
    read(u) ps%n2prod0, ps%n2prod1, ps%nangproj0, ps%nangproj1
    read(u) ps%iomminprod, ps%iommaxprod
    read(u) ps%nviniprod, ps%nvmaxprod, ps%jiniprod, ps%jmaxprod
    call rdi(nobcprod, ps%noBCprod)
    call rdd(transrgjac, ps%transRgjac)
    call rdd(prodwf, ps%prodwf)
    call rdd(dtrans, ps%dtrans)
  end if
  close(u)

  if (mw3_abi_version() /= MW3_ABI_EXPECTED) then
    write(6, *) 'mw3_fdrive: libmw3 ABI version differs from mw3_capi'
    stop 1
  end if
  rc = mw3_create(s, prec, ctx)
  call mw3_check(rc, c_null_ptr, 'mw3_create')
  if (mw3_ntot(ctx) /= n) then
    write(6, *) 'mw3_fdrive: psi0 length differs from mw3_ntot'
    stop 1
  end if
  if (hasprod == 1) call mw3_check(mw3_prod_attach(ctx, ps), ctx, 'mw3_prod_attach')
  call mw3_check(mw3_cheby_begin(ctx, psi0, c_null_ptr), ctx, 'mw3_cheby_begin')
  allocate(norms(max(nsteps, 1)), ptot(max(s%netot, 1)))
  call mw3_check(mw3_cheby_steps(ctx, nsteps, c_loc(norms), c_null_ptr), ctx, &
                 'mw3_cheby_steps')
  call mw3_check(mw3_flux_result(ctx, c_loc(ptot), c_null_ptr), ctx, 'mw3_flux_result')
  nj = s%jmax - s%jini + 1
  nv = s%nvmax - s%nvini + 1
  m = 2_c_int64_t * s%ncan * nj * nv * s%netot
  allocate(zs2(max(m, 1_c_int64_t)))
  call mw3_check(mw3_zs2_get(ctx, zs2), ctx, 'mw3_zs2_get')
  mp = 0
  if (hasprod == 1) then
    mp = 2_c_int64_t * s%netot * (ps%nvmaxprod - ps%nviniprod + 1) * &
         (ps%jmaxprod - ps%jiniprod + 1) * (ps%iommaxprod - ps%iomminprod + 1)
    allocate(zs2prod(mp))
    call mw3_check(mw3_zs2prod_get(ctx, zs2prod), ctx, 'mw3_zs2prod_get')
  end if
  call mw3_destroy(ctx)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) nsteps, norms(1:nsteps)
  write(u) s%netot, ptot(1:s%netot)
  write(u) m, zs2(1:m)
  write(u) mp
  if (mp > 0) write(u) zs2prod(1:mp)
  close(u)

contains

  subroutine rdi(a, p)
    integer(c_int32_t), allocatable, target, intent(inout) :: a(:)
    type(c_ptr), intent(out) :: p
    integer(c_int64_t) :: k
    read(u) k
    allocate(a(max(k, 1_c_int64_t)))
    if (k > 0) read(u) a(1:k)
    p = c_null_ptr
    if (k > 0) p = c_loc(a)
  end subroutine rdi

  subroutine rdd(a, p)
    real(c_double), allocatable, target, intent(inout) :: a(:)
    type(c_ptr), intent(out) :: p
    integer(c_int64_t) :: k
    read(u) k
    allocate(a(max(k, 1_c_int64_t)))
    if (k > 0) read(u) a(1:k)
    p = c_null_ptr
    if (k > 0) p = c_loc(a)
  end subroutine rdd

end program mw3_fdrive
