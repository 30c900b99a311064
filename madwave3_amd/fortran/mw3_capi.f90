! mw3_capi.f90 -- Fortran 2003 ISO_C_BINDING interface to libmw3 (include/mw3.h).
!
! Module mw3_capi gives a Fortran host (MADWAVE3's main program or the
! mod_Hphi_01y2 replacement in fortran/reference) the C ABI of libmw3:
!   type(mw3_setup)        the C struct mw3_setup, field for field (bind(C));
!   type(mw3_prod_setup)   the C struct mw3_prod_setup (product projection);
!   mw3_create ... mw3_kernel_name   one interface per C entry point;
!   mw3_error_message(ctx) the message of mw3_last_error as a Fortran string;
!   mw3_check(rc, ctx, where)  the reference's error convention: write the
!                          message to unit 6 and stop (e.g.
!                          SRC/mod_Hphi_01y2.f:70-73).
! Array arguments are assumed-size real(c_double) / integer(c_int32_t) arrays
! passed in Fortran memory order; arguments the C side accepts as NULL are
! type(c_ptr), value (pass c_loc(x) or c_null_ptr).
module mw3_capi
  use, intrinsic :: iso_c_binding
  implicit none
  private :: c_strlen

  integer(c_int32_t), parameter :: MW3_ABI_EXPECTED = 1  ! MW3_ABI_VERSION of include/mw3.h
  integer(c_int32_t), parameter :: MW3_OK = 0, MW3_EINVAL = -1, MW3_ENOMEM = -2, &
                                   MW3_EDEVICE = -3, MW3_ESTATE = -4, MW3_EUNSUP = -5
  integer(c_int32_t), parameter :: MW3_FP64 = 0, MW3_MIXED = 1
  integer(c_int32_t), parameter :: MW3_NKERNELS = 11

  ! struct mw3_setup: same members in the same order (C layout by bind(C))
  type, bind(C) :: mw3_setup
    integer(c_int32_t) :: abi_version
    integer(c_int32_t) :: npun1, npun1min, npun2, nangu, inc
    integer(c_int32_t) :: Jtot, iparity, iommin, iommax, j0
    integer(c_int32_t) :: nelecmax
    real(c_double) :: rmis1, ah1
    real(c_double) :: rmis2, ah2
    real(c_double) :: hbr
    real(c_double) :: xm1red, xm2red
    real(c_double) :: radcutmax, rotcutmax
    real(c_double) :: emindlt, delta2
    integer(c_int32_t) :: proj_j2_up
    type(c_ptr) :: iomdiat, iomatom
    type(c_ptr) :: sigdiat, sigatom
    integer(c_int32_t) :: npuntot
    type(c_ptr) :: indangreal, npunreal, indr1, indr2
    type(c_ptr) :: VVV
    integer(c_int32_t) :: ncan
    type(c_ptr) :: iombas, nelebas, nojbas, j00
    type(c_ptr) :: jbas
    integer(c_int32_t) :: ndj
    type(c_ptr) :: Djmmp
    type(c_ptr) :: absfr1, absfr2
    integer(c_int32_t) :: ir2abs
    integer(c_int32_t) :: netot
    type(c_ptr) :: etots2, reacfct
    real(c_double) :: r1flux
    integer(c_int32_t) :: nvini, nvmax, jini, jmax
    type(c_ptr) :: noBCstates
    type(c_ptr) :: fd
    integer(c_int32_t) :: iphoto, iprod, kminelastic
    integer(c_int32_t) :: icanref
    integer(c_int32_t) :: nvref, jref, iomref
    integer(c_int32_t) :: nofe
    type(c_ptr) :: efe
  end type mw3_setup

  ! struct mw3_prod_setup: the product-projection tables (iprod = 2)
  type, bind(C) :: mw3_prod_setup
    integer(c_int32_t) :: n2prod0, n2prod1, nangproj0, nangproj1
    integer(c_int32_t) :: iomminprod, iommaxprod
    integer(c_int32_t) :: nviniprod, nvmaxprod, jiniprod, jmaxprod
    type(c_ptr) :: noBCprod
    type(c_ptr) :: transRgjac, prodwf, dtrans
  end type mw3_prod_setup

  interface
    function mw3_abi_version() bind(C, name="mw3_abi_version") result(v)
      import :: c_int32_t
      integer(c_int32_t) :: v
    end function mw3_abi_version

    function mw3_last_error(ctx) bind(C, name="mw3_last_error") result(p)
      import :: c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr) :: p
    end function mw3_last_error

    function mw3_create(setup, precision, out) bind(C, name="mw3_create") result(rc)
      import :: mw3_setup, c_int32_t, c_ptr
      type(mw3_setup), intent(in) :: setup
      integer(c_int32_t), value :: precision
      type(c_ptr), intent(out) :: out
      integer(c_int32_t) :: rc
    end function mw3_create

    subroutine mw3_destroy(ctx) bind(C, name="mw3_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine mw3_destroy

    function mw3_ntot(ctx) bind(C, name="mw3_ntot") result(n)
      import :: c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t) :: n
    end function mw3_ntot

    function mw3_nflux(ctx) bind(C, name="mw3_nflux") result(n)
      import :: c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t) :: n
    end function mw3_nflux

    function mw3_hphi(ctx, psi, hpsi) bind(C, name="mw3_hphi") result(rc)
      import :: c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: psi(*)
      real(c_double), intent(out) :: hpsi(*)
      integer(c_int32_t) :: rc
    end function mw3_hphi

    function mw3_cheby_begin(ctx, psi0, stats3) bind(C, name="mw3_cheby_begin") result(rc)
      import :: c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: psi0(*)
      type(c_ptr), value :: stats3
      integer(c_int32_t) :: rc
    end function mw3_cheby_begin

    function mw3_cheby_steps(ctx, nsteps, norms, autocorr) bind(C, name="mw3_cheby_steps") &
        result(rc)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
      type(c_ptr), value :: norms, autocorr
      integer(c_int32_t) :: rc
    end function mw3_cheby_steps

    function mw3_cheby_steps_batch(ctxs, n, nsteps, norms) &
        bind(C, name="mw3_cheby_steps_batch") result(rc)
      import :: c_ptr, c_int32_t
      type(c_ptr), intent(in) :: ctxs(*)
      integer(c_int32_t), value :: n, nsteps
      type(c_ptr), value :: norms
      integer(c_int32_t) :: rc
    end function mw3_cheby_steps_batch

    function mw3_cheby_spectrum(autoc, kmax, emindlt, delta2, hbr, xkgam, e, ne, spectrum) &
        bind(C, name="mw3_cheby_spectrum") result(rc)
      import :: c_int32_t, c_double
      real(c_double), intent(in) :: autoc(*)
      integer(c_int32_t), value :: kmax
      real(c_double), value :: emindlt, delta2, hbr, xkgam
      real(c_double), intent(in) :: e(*)
      integer(c_int32_t), value :: ne
      real(c_double), intent(out) :: spectrum(*)
      integer(c_int32_t) :: rc
    end function mw3_cheby_spectrum

    ! device buffers and a hipStream_t as c_ptr (no allocation in the library)
    function mw3_cheby_spectrum_noalloc(d_autoc, kmax, emindlt, delta2, hbr, xkgam, d_e, ne, &
        d_spectrum, stream) bind(C, name="mw3_cheby_spectrum_noalloc") result(rc)
      import :: c_ptr, c_int32_t, c_double
      type(c_ptr), value :: d_autoc
      integer(c_int32_t), value :: kmax
      real(c_double), value :: emindlt, delta2, hbr, xkgam
      type(c_ptr), value :: d_e
      integer(c_int32_t), value :: ne
      type(c_ptr), value :: d_spectrum, stream
      integer(c_int32_t) :: rc
    end function mw3_cheby_spectrum_noalloc

    function mw3_cheby_k(ctx) bind(C, name="mw3_cheby_k") result(k)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: k
    end function mw3_cheby_k

    function mw3_flux_result(ctx, s2prodtot, s2prodelec) bind(C, name="mw3_flux_result") &
        result(rc)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      type(c_ptr), value :: s2prodtot, s2prodelec
      integer(c_int32_t) :: rc
    end function mw3_flux_result

    function mw3_flux_energy_test(ctx, on) bind(C, name="mw3_flux_energy_test") result(rc)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
      integer(c_int32_t) :: rc
    end function mw3_flux_energy_test

    function mw3_records(ctx, on) bind(C, name="mw3_records") result(rc)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
      integer(c_int32_t) :: rc
    end function mw3_records

    function mw3_records_get(ctx, maxrec, it, cvj, cvjprod, nrec) &
        bind(C, name="mw3_records_get") result(rc)
      import :: c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: maxrec
      integer(c_int32_t), intent(out) :: it(*)
      real(c_double), intent(out) :: cvj(*), cvjprod(*)
      integer(c_int32_t), intent(out) :: nrec
      integer(c_int32_t) :: rc
    end function mw3_records_get

    function mw3_zs2_get(ctx, zs2) bind(C, name="mw3_zs2_get") result(rc)
      import :: c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: zs2(*)
      integer(c_int32_t) :: rc
    end function mw3_zs2_get

    function mw3_flux_accum_get(ctx, zcr, zcdr) bind(C, name="mw3_flux_accum_get") result(rc)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      type(c_ptr), value :: zcr, zcdr
      integer(c_int32_t) :: rc
    end function mw3_flux_accum_get

    function mw3_flux_accum_put(ctx, zcr, zcdr, zs2) bind(C, name="mw3_flux_accum_put") &
        result(rc)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      type(c_ptr), value :: zcr, zcdr, zs2
      integer(c_int32_t) :: rc
    end function mw3_flux_accum_put

    function mw3_prod_attach(ctx, prod) bind(C, name="mw3_prod_attach") result(rc)
      import :: c_ptr, c_int32_t, mw3_prod_setup
      type(c_ptr), value :: ctx
      type(mw3_prod_setup), intent(in) :: prod
      integer(c_int32_t) :: rc
    end function mw3_prod_attach

    function mw3_zs2prod_get(ctx, zs2prod) bind(C, name="mw3_zs2prod_get") result(rc)
      import :: c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: zs2prod(*)
      integer(c_int32_t) :: rc
    end function mw3_zs2prod_get

    function mw3_zs2prod_put(ctx, zs2prod) bind(C, name="mw3_zs2prod_put") result(rc)
      import :: c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: zs2prod(*)
      integer(c_int32_t) :: rc
    end function mw3_zs2prod_put

    function mw3_cvjprod_get(ctx, cvjprod) bind(C, name="mw3_cvjprod_get") result(rc)
      import :: c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: cvjprod(*)
      integer(c_int32_t) :: rc
    end function mw3_cvjprod_get

    function mw3_psie_get(ctx, zfe) bind(C, name="mw3_psie_get") result(rc)
      import :: c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: zfe(*)
      integer(c_int32_t) :: rc
    end function mw3_psie_get

    function mw3_state_get(ctx, psi_km1, psi_k, k) bind(C, name="mw3_state_get") result(rc)
      import :: c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: psi_km1(*), psi_k(*)
      integer(c_int32_t), intent(out) :: k
      integer(c_int32_t) :: rc
    end function mw3_state_get

    function mw3_state_put(ctx, psi_km1, psi_k, k, psi0) bind(C, name="mw3_state_put") &
        result(rc)
      import :: c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: psi_km1(*), psi_k(*)
      integer(c_int32_t), value :: k
      type(c_ptr), value :: psi0
      integer(c_int32_t) :: rc
    end function mw3_state_put

    function mw3_timing_enable(ctx, on) bind(C, name="mw3_timing_enable") result(rc)
      import :: c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
      integer(c_int32_t) :: rc
    end function mw3_timing_enable

    function mw3_timing_get(ctx, id, total_ms, launches) bind(C, name="mw3_timing_get") &
        result(rc)
      import :: c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: id
      real(c_double), intent(out) :: total_ms
      integer(c_int64_t), intent(out) :: launches
      integer(c_int32_t) :: rc
    end function mw3_timing_get

    function mw3_kernel_name(id) bind(C, name="mw3_kernel_name") result(p)
      import :: c_ptr, c_int32_t
      integer(c_int32_t), value :: id
      type(c_ptr) :: p
    end function mw3_kernel_name

    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function c_strlen
  end interface

contains

  ! message of mw3_last_error(ctx) (ctx may be c_null_ptr: last failed create)
  function mw3_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: i, n
    p = mw3_last_error(ctx)
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = s(i)
    end do
  end function mw3_error_message

  ! stop with the library's message when rc is not MW3_OK
  subroutine mw3_check(rc, ctx, where)
    integer(c_int32_t), intent(in) :: rc
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: where
    if (rc /= MW3_OK) then
      write(6, '(a,a,i0,a,a)') where, ': libmw3 status ', rc, ': ', mw3_error_message(ctx)
      call flush(6)
      stop 1
    end if
  end subroutine mw3_check

end module mw3_capi
