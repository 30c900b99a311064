! mod_mw3_cheby.f90 -- the Chebyshev propagation of MADWAVE3's main program on
! libmw3: the recursion and the per-step projections stay on the GPU, and the
! accumulators are copied into the reference's module arrays only when the
! analysis line needs them.
!
! Calls, in the order of SRC/main_madwave3.f:
!   after ini_flux, input_PsiE and ini_transcoor (188-197):
!                                             call mw3_cheby_init(kminelastic)
!   in place of 231-250 (first step or restart):
!                                             call mw3_cheby_start(iit0)
!   in place of the "do it" loop (324-368):    call mw3_cheby_block(ntimes, xnorm)
!   before totalflux / wvpcont / printprob (375-380):
!                                             call mw3_cheby_sync
! With iprod = 2 the product projection of prodpaq / prodcvj runs on the GPU
! too, on the tables ini_transcoor built (module mod_coortrans01_02).  The
! operator context of mod_Hphi_01y2 (set_vectors_Hphi) stays usable for difs.
! Single process (nproc = 1).
module mod_mw3_cheby
  use, intrinsic :: iso_c_binding
  use mw3_capi
  use mod_gridYpara_01y2
  use mod_pot_01y2
  use mod_baseYfunciones_01y2
  use mod_Hphi_01y2
  use mod_flux_01y2
  use mod_PsiE_01y2
  use mod_coortrans01_02
  implicit none
  save
  private
  public :: mw3_pctx, mw3_cheby_init, mw3_cheby_start, mw3_cheby_block, mw3_cheby_sync

  type(c_ptr) :: mw3_pctx = c_null_ptr
  real(c_double), allocatable, target :: k_etot(:), k_reacfct(:), k_efe(:)

contains

  ! propagation context: the operator of mw3_fill_hphi plus the energy
  ! projection of ini_flux (SRC/mod_flux_01y2.f:54-334), Cvjflux_k and funE
  subroutine mw3_cheby_init(kminelastic)
    integer, intent(in) :: kminelastic
    type(mw3_setup) :: s
    integer(c_int32_t) :: rc
    call mw3_fill_hphi(s)
    k_etot = etotS2(1:netot)
    k_reacfct = reacfct(1:netot)
    s%netot = netot
    s%etots2 = c_loc(k_etot)
    s%reacfct = c_loc(k_reacfct)
    s%r1flux = r1flux
    s%kminelastic = kminelastic
    s%icanref = icanref
    s%nofe = nofe
    if (nofe > 0) then
      k_efe = efe(1:nofe)
      s%efe = c_loc(k_efe)
    end if
    rc = mw3_create(s, mw3_precision, mw3_pctx)
    call mw3_setup_keep_free()
    deallocate(k_etot, k_reacfct)
    if (allocated(k_efe)) deallocate(k_efe)
    call mw3_check(rc, c_null_ptr, 'mw3_cheby_init: mw3_create')
    if (iprod == 2 .and. npun1 > 1 .and. npun2 > 1 .and. allocated(transRgjac)) &
      call mw3_prod_tables()
  end subroutine mw3_cheby_init

  ! the tables of ini_transcoor (SRC/mod_coortrans01_02.f:31-209) to the
  ! context: prodpaq and prodcvj then run in every step
  subroutine mw3_prod_tables()
    type(mw3_prod_setup) :: ps
    integer(c_int32_t), allocatable, target :: k_nobc(:)
    real(c_double), allocatable, target :: k_tr(:), k_wf(:), k_dt(:)
    ps%n2prod0 = n2prod0
    ps%n2prod1 = n2prod1
    ps%nangproj0 = nangproj0
    ps%nangproj1 = nangproj1
    ps%iomminprod = iomminprod
    ps%iommaxprod = iommaxprod
    ps%nviniprod = nviniprod
    ps%nvmaxprod = nvmaxprod
    ps%jiniprod = jiniprod
    ps%jmaxprod = jmaxprod
    k_nobc = noBCprod(jiniprod:jmaxprod)
    k_tr = transRgjac(1:nbastotproc)
    k_wf = reshape(prodwf, [size(prodwf)])
    k_dt = reshape(dtrans(:, :, :, 1:ncan), [size(dtrans(:, :, :, 1:ncan))])
    ps%noBCprod = c_loc(k_nobc)
    ps%transRgjac = c_loc(k_tr)
    ps%prodwf = c_loc(k_wf)
    ps%dtrans = c_loc(k_dt)
    call mw3_check(mw3_prod_attach(mw3_pctx, ps), mw3_pctx, 'mw3_cheby_init: mw3_prod_attach')
    deallocate(k_nobc, k_tr, k_wf, k_dt)
  end subroutine mw3_prod_tables

  ! iit0 = 0: psi_0 = rpaq0 and the undamped first step (main_madwave3.f:
  ! 231-250); iit0 > 0: continue from rflanz0proc, rpaqproc read from the
  ! contNN.paq file (250-259)
  subroutine mw3_cheby_start(iit0)
    integer, intent(inout) :: iit0
    if (iit0 == 0) then
      call mw3_check(mw3_cheby_begin(mw3_pctx, rpaq0, c_null_ptr), mw3_pctx, &
                     'mw3_cheby_start: mw3_cheby_begin')
      iit0 = 1
    else
      call mw3_check(mw3_state_put(mw3_pctx, rflanz0proc, rpaqproc, int(iit0, c_int32_t), &
                                   c_null_ptr), mw3_pctx, 'mw3_cheby_start: mw3_state_put')
      ! accumulators read from contNN.S2 / contNN.S2prod by ini_flux and
      ! ini_transcoor (SRC/mod_flux_01y2.f:311-322, mod_coortrans01_02.f:186-193)
      if (ncontfile == 1) then
        call mw3_accum_put()
        if (allocated(zS2prod) .and. iprod == 2) call mw3_check(mw3_zs2prod_put(mw3_pctx, &
            zs2prod_flat()), mw3_pctx, 'mw3_cheby_start: mw3_zs2prod_put')
      end if
    end if
  end subroutine mw3_cheby_start

  subroutine mw3_accum_put()
    complex(c_double_complex), allocatable, target :: b1(:, :), b2(:, :), b3(:)
    integer(c_int64_t) :: nf
    type(c_ptr) :: p3
    nf = mw3_nflux(mw3_pctx)
    allocate(b1(netot, max(nf, 1_c_int64_t)), b2(netot, max(nf, 1_c_int64_t)))
    if (nf > 0) then
      b1 = zCR(1:netot, 1:nf)
      b2 = zCdR(1:netot, 1:nf)
    end if
    p3 = c_null_ptr
    if (allocated(zS2)) then
      b3 = reshape(zS2(:, :, :, 1:ncan), [size(zS2(:, :, :, 1:ncan))])
      p3 = c_loc(b3)
    end if
    call mw3_check(mw3_flux_accum_put(mw3_pctx, c_loc(b1), c_loc(b2), p3), mw3_pctx, &
                   'mw3_cheby_start: mw3_flux_accum_put')
  end subroutine mw3_accum_put

  function zs2prod_flat() result(z)
    real(c_double), allocatable :: z(:)
    complex(c_double_complex), allocatable :: t(:)
    t = reshape(zS2prod, [size(zS2prod)])
    allocate(z(2 * size(t)))
    z(1::2) = real(t)
    z(2::2) = aimag(t)
  end function zs2prod_flat

  ! ntimes steps it = iit0+1 .. iit0+ntimes of main_madwave3.f:324-368 with
  ! difs, the damped update, CoefEcheby, totalflux_k, Cvjflux_k, check and funE;
  ! xnorm(i) = |psi|^2 after step i; autocor(i) = <psi_0|psi> (the auto file
  ! of check, iphoto >= 1), when present
  subroutine mw3_cheby_block(ntimes, xnorm, autocor)
    integer, intent(in) :: ntimes
    real(c_double), intent(out), target :: xnorm(ntimes)
    real(c_double), intent(out), target, optional :: autocor(ntimes)
    type(c_ptr) :: pa
    pa = c_null_ptr
    if (present(autocor)) pa = c_loc(autocor)
    call mw3_check(mw3_cheby_steps(mw3_pctx, int(ntimes, c_int32_t), c_loc(xnorm), pa), &
                   mw3_pctx, 'mw3_cheby_block: mw3_cheby_steps')
  end subroutine mw3_cheby_block

  ! device accumulators and state into the reference's module arrays (zCR,
  ! zCdR, zS2 of mod_flux_01y2, zfe of mod_PsiE_01y2, rflanz0proc and
  ! rpaqproc) before totalflux, wvpcont and printprob
  subroutine mw3_cheby_sync
    complex(c_double_complex), allocatable, target :: b1(:, :), b2(:, :)
    real(c_double), allocatable :: z(:)
    integer(c_int64_t) :: nf
    integer(c_int32_t) :: k
    integer :: nj, nv, ie, iv, j, ic, i
    nf = mw3_nflux(mw3_pctx)
    if (nf > 0) then
      allocate(b1(netot, nf), b2(netot, nf))
      call mw3_check(mw3_flux_accum_get(mw3_pctx, c_loc(b1), c_loc(b2)), mw3_pctx, &
                     'mw3_cheby_sync: mw3_flux_accum_get')
      zCR(1:netot, 1:nf) = b1
      zCdR(1:netot, 1:nf) = b2
      deallocate(b1, b2)
    end if
    nj = jmax - jini + 1
    nv = nvmax - nvini + 1
    if (allocated(zS2) .and. nj > 0 .and. nv > 0) then
      allocate(z(2 * netot * nv * nj * ncan))
      call mw3_check(mw3_zs2_get(mw3_pctx, z), mw3_pctx, 'mw3_cheby_sync: mw3_zs2_get')
      i = 0
      do ic = 1, ncan
        do j = jini, jmax
          do iv = nvini, nvmax
            do ie = 1, netot
              zS2(ie, iv, j, ic) = dcmplx(z(i + 1), z(i + 2))
              i = i + 2
            end do
          end do
        end do
      end do
      deallocate(z)
    end if
    if (nofe > 0) then
      allocate(z(2 * nbastotproc * nofe))
      call mw3_check(mw3_psie_get(mw3_pctx, z), mw3_pctx, 'mw3_cheby_sync: mw3_psie_get')
      zfe(1:nbastotproc, 1:nofe) = reshape(cmplx(z(1::2), z(2::2), kind=c_double_complex), &
                                           [nbastotproc, nofe])
      deallocate(z)
    end if
    if (iprod == 2 .and. allocated(zS2prod) .and. allocated(transRgjac)) then
      ! zS2prod and the last step's Cvjprod of prodpaq / prodcvj
      allocate(z(2 * size(zS2prod)))
      call mw3_check(mw3_zs2prod_get(mw3_pctx, z), mw3_pctx, 'mw3_cheby_sync: mw3_zs2prod_get')
      zS2prod = reshape(cmplx(z(1::2), z(2::2), kind=c_double_complex), shape(zS2prod))
      deallocate(z)
      allocate(z(size(Cvjprod)))
      call mw3_check(mw3_cvjprod_get(mw3_pctx, z), mw3_pctx, 'mw3_cheby_sync: mw3_cvjprod_get')
      Cvjprod = reshape(z, shape(Cvjprod))
      deallocate(z)
    end if
    call mw3_check(mw3_state_get(mw3_pctx, rflanz0proc, rpaqproc, k), mw3_pctx, &
                   'mw3_cheby_sync: mw3_state_get')
  end subroutine mw3_cheby_sync

end module mod_mw3_cheby
