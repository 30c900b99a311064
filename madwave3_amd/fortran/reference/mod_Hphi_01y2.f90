! mod_Hphi_01y2.f90 -- drop-in replacement of MADWAVE3's module mod_Hphi_01y2
! (SRC/mod_Hphi_01y2.f) that forwards H psi to libmw3 on the GPU.
!
! It keeps the names the rest of MADWAVE3 uses (SRC/main_madwave3.f,
! mod_colini_01y2, mod_photoini_01y2, mod_flux_01y2, mod_PsiE_01y2,
! mod_coortrans01_02):
!   rpaqproc, rflanz0proc, rHpaqrec, rpaq0   state vectors (nbastotproc)
!   xpr1, xp2r1, xpr2, xp2r2                 momentum work arrays (mod_flux_01y2.f:134)
!   set_vectors_Hphi                         allocation + libmw3 context
!   difs                                     rHpaqrec = H rpaqproc (zots)
!   Hdiatom                                  not provided: stops
! set_vectors_Hphi must run where the reference calls it
! (SRC/main_madwave3.f:159): after pot2, when the grid, potential, basis,
! angular functions and absorption are set.  The context it creates holds the
! operator only (no energy projection); module mod_mw3_cheby creates the
! propagation context once ini_flux has run.  Single process (nproc = 1).
module mod_Hphi_01y2
  use, intrinsic :: iso_c_binding
  use mod_gridYpara_01y2
  use mod_pot_01y2
  use mod_baseYfunciones_01y2
  use mod_absorcion_01y2
  use mw3_capi
  implicit none
  save
  private
  public :: rpaqproc, rflanz0proc, rHpaqrec, rpaq0, xpr1, xp2r1, xpr2, xp2r2
  public :: set_vectors_Hphi, difs, Hdiatom
  public :: mw3_hctx, mw3_precision, mw3_fill_hphi, mw3_setup_keep_free

  real*8, allocatable :: rpaqproc(:), rflanz0proc(:), rHpaqrec(:), rpaq0(:)
  real*8, allocatable :: xpr1(:), xp2r1(:), xpr2(:), xp2r2(:)

  ! libmw3 operator context; precision MW3_FP64 or MW3_MIXED
  type(c_ptr) :: mw3_hctx = c_null_ptr
  integer(c_int32_t) :: mw3_precision = MW3_FP64

  ! copies of the module arrays whose addresses a mw3_setup holds; they live
  ! until mw3_setup_keep_free (mw3_create reads them during the call only)
  integer(c_int32_t), allocatable, target :: k_iomdiat(:), k_iomatom(:), k_indangreal(:), &
      k_npunreal(:), k_indr1(:), k_indr2(:), k_iombas(:), k_nelebas(:), k_nojbas(:), &
      k_j00(:), k_jbas(:), k_nobc(:)
  real(c_double), allocatable, target :: k_sigdiat(:), k_sigatom(:), k_vvv(:), k_djmmp(:), &
      k_absfr1(:), k_absfr2(:), k_fd(:)

contains

  subroutine set_vectors_Hphi
    type(mw3_setup) :: s
    integer :: ierror
    integer(c_int32_t) :: rc
    if (nproc /= 1) then
      write(6, *) ' set_vectors_Hphi: the libmw3 module runs one process (nproc = 1)'
      call flush(6)
      stop
    end if
    allocate(rpaqproc(nbastotproc), rflanz0proc(nbastotproc), rpaq0(nbastotproc), &
             rHpaqrec(nbastotproc), xpr1(npun1), xp2r1(npun1), xpr2(npun2), xp2r2(npun2), &
             stat=ierror)
    if (ierror /= 0) then
      write(*, *) ' error in initmem for paqs and Hpaqs '
      stop
    end if
    rpaq0(:) = 0.d0
    call mw3_fill_hphi(s)
    rc = mw3_create(s, mw3_precision, mw3_hctx)
    call mw3_setup_keep_free()
    call mw3_check(rc, c_null_ptr, 'set_vectors_Hphi: mw3_create')
    if (mw3_ntot(mw3_hctx) /= int(nbastotproc, c_int64_t)) then
      write(6, *) ' set_vectors_Hphi: libmw3 state length differs from nbastotproc'
      call flush(6)
      stop
    end if
    write(6, *) ' set_vectors_Hphi: libmw3 context created, state length ', nbastotproc
    call flush(6)
  end subroutine set_vectors_Hphi

  ! rHpaqrec = H rpaqproc (SRC/mod_Hphi_01y2.f:623-943), zots, unnormalised
  subroutine difs
    call mw3_check(mw3_hphi(mw3_hctx, rpaqproc, rHpaqrec), mw3_hctx, 'difs: mw3_hphi')
  end subroutine difs

  subroutine Hdiatom
    write(6, *) ' Hdiatom: the diatomic-only path (xm2 < 1e-3) is not in libmw3'
    call flush(6)
    stop
  end subroutine Hdiatom

  ! the operator members of a mw3_setup from the reference modules; the energy
  ! projection members are left empty (netot = 0)
  subroutine mw3_fill_hphi(s)
    type(mw3_setup), intent(out) :: s
    s%abi_version = MW3_ABI_EXPECTED
    s%npun1 = npun1
    s%npun1min = npun1min
    s%npun2 = npun2
    s%nangu = nangu
    s%inc = inc
    s%Jtot = Jtot
    s%iparity = iparity
    s%iommin = iommin
    s%iommax = iommax
    s%j0 = j0
    s%nelecmax = nelecmax
    s%rmis1 = rmis1
    s%ah1 = ah1
    s%rmis2 = rmis2
    s%ah2 = ah2
    s%hbr = hbr
    s%xm1red = xm1red
    s%xm2red = xm2red
    s%radcutmax = radcutmax
    s%rotcutmax = rotcutmax
    s%emindlt = emindlt
    s%delta2 = delta2
    s%proj_j2_up = Proj_j2_Up
    k_iomdiat = iomdiat(1:nelecmax)
    k_iomatom = iomatom(1:nelecmax)
    k_sigdiat = sigdiat(1:nelecmax)
    k_sigatom = sigatom(1:nelecmax)
    s%iomdiat = c_loc(k_iomdiat)
    s%iomatom = c_loc(k_iomatom)
    s%sigdiat = c_loc(k_sigdiat)
    s%sigatom = c_loc(k_sigatom)
    s%npuntot = npuntot
    k_indangreal = indangreal(1:nangu, 0)
    k_npunreal = npunreal(1:nangu)
    k_indr1 = reshape(indr1(1:npuntot, 1:nangu), [npuntot * nangu])
    k_indr2 = reshape(indr2(1:npuntot, 1:nangu), [npuntot * nangu])
    k_vvv = reshape(VVV(1:npuntot, 1:nangu, 1:nelecmax, 1:nelecmax), &
                    [npuntot * nangu * nelecmax * nelecmax])
    s%indangreal = c_loc(k_indangreal)
    s%npunreal = c_loc(k_npunreal)
    s%indr1 = c_loc(k_indr1)
    s%indr2 = c_loc(k_indr2)
    s%VVV = c_loc(k_vvv)
    s%ncan = ncan
    k_iombas = iombas(1:ncan)
    k_nelebas = nelebas(1:ncan)
    k_nojbas = nojbas(1:ncan)
    k_j00 = j00(1:ncan)
    k_jbas = reshape(jbas(1:nangu, 1:ncan), [nangu * ncan])
    s%iombas = c_loc(k_iombas)
    s%nelebas = c_loc(k_nelebas)
    s%nojbas = c_loc(k_nojbas)
    s%j00 = c_loc(k_j00)
    s%jbas = c_loc(k_jbas)
    s%ndj = nangu2 + 1
    k_djmmp = reshape(Djmmp(1:nangu, 0:nangu2, 1:nelecmax, 0:Jtot), &
                      [nangu * (nangu2 + 1) * nelecmax * (Jtot + 1)])
    s%Djmmp = c_loc(k_djmmp)
    k_absfr1 = absfr1(1:npun1)
    k_absfr2 = absfr2(1:npun2)
    s%absfr1 = c_loc(k_absfr1)
    s%absfr2 = c_loc(k_absfr2)
    s%ir2abs = ir2abs
    s%netot = 0
    s%etots2 = c_null_ptr
    s%reacfct = c_null_ptr
    s%r1flux = 0.d0
    s%nvini = nvini
    s%nvmax = nvmax
    s%jini = jini
    s%jmax = jmax
    k_nobc = reshape(noBCstates(jini:jmax, 1:nelecmax), [(jmax - jini + 1) * nelecmax])
    k_fd = reshape(fd(1:npun1, nvini:nvmax, jini:jmax, 1:nelecmax), &
                   [npun1 * (nvmax - nvini + 1) * (jmax - jini + 1) * nelecmax])
    s%noBCstates = c_loc(k_nobc)
    s%fd = c_loc(k_fd)
    s%iphoto = iphoto
    s%iprod = iprod
    s%kminelastic = 0
    s%icanref = 0
    s%nvref = nvref
    s%jref = jref
    s%iomref = iomref
    s%nofe = 0
    s%efe = c_null_ptr
  end subroutine mw3_fill_hphi

  subroutine mw3_setup_keep_free()
    if (allocated(k_iomdiat)) deallocate(k_iomdiat, k_iomatom, k_sigdiat, k_sigatom)
    if (allocated(k_indangreal)) deallocate(k_indangreal, k_npunreal, k_indr1, k_indr2, k_vvv)
    if (allocated(k_iombas)) deallocate(k_iombas, k_nelebas, k_nojbas, k_j00, k_jbas)
    if (allocated(k_djmmp)) deallocate(k_djmmp, k_absfr1, k_absfr2, k_nobc, k_fd)
  end subroutine mw3_setup_keep_free

end module mod_Hphi_01y2
