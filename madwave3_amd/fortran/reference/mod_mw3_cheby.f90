! mod_mw3_cheby.f90 -- the Chebyshev propagation of MADWAVE3's main program on
! libmw3: the recursion and the per-step projections stay on the GPU, and the
! accumulators are copied into the reference's module arrays only when the
! analysis line needs them.
!
! Calls, in the order of SRC/main_madwave3.f:
!   after ini_flux and input_PsiE (188-192):  call mw3_cheby_init(kminelastic)
!   in place of 231-250 (first step or restart):
!                                             call mw3_cheby_start(iit0)
!   in place of the "do it" loop (324-368):    call mw3_cheby_block(ntimes, xnorm)
!   before totalflux / wvpcont / printprob (375-380):
!                                             call mw3_cheby_sync
! The operator context of mod_Hphi_01y2 (set_vectors_Hphi) stays usable for
! difs.  Single process (nproc = 1).
module mod_mw3_cheby
  use, intrinsic :: iso_c_binding
  use mw3_capi
  use mod_gridYpara_01y2
  use mod_pot_01y2
  use mod_baseYfunciones_01y2
  use mod_Hphi_01y2
  use mod_flux_01y2
  use mod_PsiE_01y2
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
  end subroutine mw3_cheby_init

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
    end if
  end subroutine mw3_cheby_start

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
    call mw3_check(mw3_state_get(mw3_pctx, rflanz0proc, rpaqproc, k), mw3_pctx, &
                   'mw3_cheby_sync: mw3_state_get')
  end subroutine mw3_cheby_sync

end module mod_mw3_cheby
