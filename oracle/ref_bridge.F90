! Test-infrastructure bridge: exposes the reference TEMPO FP64 build
! (compiled from the reference sources by oracle/Makefile into oracle/_ref)
! through a plain C ABI so that the tests and the CPU baseline can drive it
! with ctypes.  It is never linked into the product library.
!
! ref_init replays the initialisation sequence of tempo_init
! (src/module_mp_tempo_driver.F90:42-182) but takes the four lookup tables
! from caller memory instead of files, in the reference file order
! (src/module_mp_tempo_driver.F90:532-649).
!
! State and diagnostics are exchanged in FP64 and converted to the build's
! working precision wp at the boundary, so the same bridge serves the FP64
! build and the -DSINGLE_PREC build (wp = 4).
!
! ref_run calls tempo_run (src/module_mp_tempo_driver.F90:187-513) on ncol
! columns laid out (ncol, nz) column-major, i.e. element (i,k) at
! i + k*ncol -- one column per contiguous i, levels strided by ncol.
! A NULL pointer for an optional field means "not present", exactly like an
! absent optional argument of tempo_run.
module ref_bridge
  use iso_c_binding
  use module_mp_tempo_cfgs, only : ty_tempo_cfgs
  use module_mp_tempo_params
  use module_mp_tempo_utils, only : compute_efrw, compute_efsw, compute_drop_evap, qi_aut_qs, &
    calc_gamma_p, calc_rslf, calc_rsif, snow_moments
  use module_mp_tempo_driver, only : tempo_run, ty_tempo_driver_diags, tempo_aerosol_surface_emissions
  use module_mp_tempo_main, only : cloud_check_and_update, ice_check_and_update, snow_check_and_update
  use module_mp_tempo_ml, only : ty_tempo_ml_data, save_or_read_ml_data, nc_ml_nodes, nc_ml_input, &
    nc_ml_output, nc_ml_trans_mean, nc_ml_trans_var, nc_ml_w00, nc_ml_w01, nc_ml_b00, nc_ml_b01
  implicit none

  type(ty_tempo_cfgs), save :: cfgs
  type(ty_tempo_driver_diags), save :: diags
  ! boundary-layer cloud inputs of the next ref_run (ref_set_bl)
  type(c_ptr), save :: bl_qc = c_null_ptr, bl_frac = c_null_ptr

contains

  subroutine set_flag(flag, val)
    logical, intent(inout) :: flag
    integer(c_int), intent(in) :: val
    flag = (val /= 0)
  end subroutine set_flag

  ! flags(1:18): aerosolaware, hailaware, semi_sedi, refl_from_melting,
  ! turn_off_micro, cloud_condensation, refl10cm, re_cloud, re_ice, re_snow,
  ! max_hail, mvd_r, mvd_g, nc_mixing_ratio, ml_for_bl_nc, ml_for_nc,
  ! nc_opt (0 land, 1 ocean), unused
  integer(c_int) function ref_init(flags, ccn, fz, qs, qg) bind(c, name='ref_init')
    integer(c_int), intent(in) :: flags(18)
    type(c_ptr), value :: ccn, fz, qs, qg
    real(c_float), pointer :: pccn(:)
    real(c_double), pointer :: p(:)
    integer :: nfz_r, nfz_c, nqs, nqg, off

    cfgs = ty_tempo_cfgs()
    call set_flag(cfgs%aerosolaware_flag, flags(1))
    call set_flag(cfgs%hailaware_flag, flags(2))
    call set_flag(cfgs%semi_sedi_flag, flags(3))
    call set_flag(cfgs%refl10cm_from_melting_flag, flags(4))
    call set_flag(cfgs%turn_off_micro_flag, flags(5))
    call set_flag(cfgs%cloud_condensation_flag, flags(6))
    call set_flag(cfgs%refl10cm_flag, flags(7))
    call set_flag(cfgs%re_cloud_flag, flags(8))
    call set_flag(cfgs%re_ice_flag, flags(9))
    call set_flag(cfgs%re_snow_flag, flags(10))
    call set_flag(cfgs%max_hail_diameter_flag, flags(11))
    call set_flag(cfgs%rain_med_vol_diam_flag, flags(12))
    call set_flag(cfgs%graupel_med_vol_diam_flag, flags(13))
    call set_flag(cfgs%cloud_number_mixing_ratio_flag, flags(14))
    call set_flag(cfgs%ml_for_bl_nc_flag, flags(15))
    call set_flag(cfgs%ml_for_nc_flag, flags(16))
    if (flags(17) == 1) then
      cfgs%single_moment_nc_opt = 'ocean'
    else
      cfgs%single_moment_nc_opt = 'land'
    endif

    call initialize_graupel_vars(cfgs%hailaware_flag)
    call initialize_parameters()
    call initialize_bins_for_tables()
    call initialize_array_efrw()
    call compute_efrw()
    call initialize_array_efsw()
    call compute_efsw()
    call initialize_arrays_drop_evap()
    call compute_drop_evap()
    call initialize_arrays_qi_aut_qs()
    call qi_aut_qs()

    call initialize_arrays_ccn()
    call c_f_pointer(ccn, pccn, [ntb_arc*ntb_arw*ntb_art*ntb_arr*ntb_ark])
    tnccn_act = reshape(pccn, shape(tnccn_act))

    call initialize_arrays_freezewater()
    nfz_r = ntb_r*ntb_r1*ntb_t1*ntb_in
    nfz_c = ntb_c*nbc*ntb_t1*ntb_in
    call c_f_pointer(fz, p, [4*nfz_r + 2*nfz_c])
    off = 0
    tpi_qrfz = reshape(p(off+1:off+nfz_r), shape(tpi_qrfz)); off = off + nfz_r
    tni_qrfz = reshape(p(off+1:off+nfz_r), shape(tni_qrfz)); off = off + nfz_r
    tpg_qrfz = reshape(p(off+1:off+nfz_r), shape(tpg_qrfz)); off = off + nfz_r
    tnr_qrfz = reshape(p(off+1:off+nfz_r), shape(tnr_qrfz)); off = off + nfz_r
    tpi_qcfz = reshape(p(off+1:off+nfz_c), shape(tpi_qcfz)); off = off + nfz_c
    tni_qcfz = reshape(p(off+1:off+nfz_c), shape(tni_qcfz))

    call initialize_arrays_qr_acr_qs()
    nqs = ntb_s*ntb_t*ntb_r1*ntb_r
    call c_f_pointer(qs, p, [12*nqs])
    tcs_racs1 = reshape(p(0*nqs+1:1*nqs), shape(tcs_racs1))
    tmr_racs1 = reshape(p(1*nqs+1:2*nqs), shape(tmr_racs1))
    tcs_racs2 = reshape(p(2*nqs+1:3*nqs), shape(tcs_racs2))
    tmr_racs2 = reshape(p(3*nqs+1:4*nqs), shape(tmr_racs2))
    tcr_sacr1 = reshape(p(4*nqs+1:5*nqs), shape(tcr_sacr1))
    tms_sacr1 = reshape(p(5*nqs+1:6*nqs), shape(tms_sacr1))
    tcr_sacr2 = reshape(p(6*nqs+1:7*nqs), shape(tcr_sacr2))
    tms_sacr2 = reshape(p(7*nqs+1:8*nqs), shape(tms_sacr2))
    tnr_racs1 = reshape(p(8*nqs+1:9*nqs), shape(tnr_racs1))
    tnr_racs2 = reshape(p(9*nqs+1:10*nqs), shape(tnr_racs2))
    tnr_sacr1 = reshape(p(10*nqs+1:11*nqs), shape(tnr_sacr1))
    tnr_sacr2 = reshape(p(11*nqs+1:12*nqs), shape(tnr_sacr2))

    call initialize_arrays_qr_acr_qg()
    nqg = ntb_g1*ntb_g*nrhg*ntb_r1*ntb_r
    call c_f_pointer(qg, p, [5*nqg])
    tcg_racg = reshape(p(0*nqg+1:1*nqg), shape(tcg_racg))
    tmr_racg = reshape(p(1*nqg+1:2*nqg), shape(tmr_racg))
    tcr_gacr = reshape(p(2*nqg+1:3*nqg), shape(tcr_gacr))
    tnr_racg = reshape(p(3*nqg+1:4*nqg), shape(tnr_racg))
    tnr_gacr = reshape(p(4*nqg+1:5*nqg), shape(tnr_gacr))

    if (cfgs%max_hail_diameter_flag) call initialize_bins_for_hail_size()
    if (cfgs%refl10cm_from_melting_flag) call initialize_bins_for_radar()
    if (cfgs%ml_for_bl_nc_flag .or. cfgs%ml_for_nc_flag) call load_ml_data()
    ref_init = 0
  end function ref_init

  ! the network parameters handed to the reference as init_ml_data does
  ! (src/module_mp_tempo_driver.F90:703-730; that routine is private)
  subroutine load_ml_data()
    type(ty_tempo_ml_data), dimension(1) :: d
    d(1)%input_size = nc_ml_input
    d(1)%node_size = nc_ml_nodes
    d(1)%output_size = nc_ml_output
    allocate(d(1)%transform_mean(nc_ml_input), d(1)%transform_var(nc_ml_input))
    allocate(d(1)%weights00(nc_ml_nodes, nc_ml_input), d(1)%weights01(nc_ml_output, nc_ml_nodes))
    allocate(d(1)%bias00(nc_ml_nodes), d(1)%bias01(nc_ml_output))
    d(1)%transform_mean = nc_ml_trans_mean
    d(1)%transform_var = nc_ml_trans_var
    d(1)%weights00 = reshape(nc_ml_w00, [nc_ml_nodes, nc_ml_input])
    d(1)%weights01 = reshape(nc_ml_w01, [nc_ml_output, nc_ml_nodes])
    d(1)%bias00 = nc_ml_b00
    d(1)%bias01 = nc_ml_b01
    call save_or_read_ml_data(ml_data_in=d)
  end subroutine load_ml_data

  ! qc_bl and qcfrac_bl ((ncol, nz) column-major FP64, NULL = absent) for the
  ! following ref_run calls
  subroutine ref_set_bl(qc_bl, qcfrac_bl) bind(c, name='ref_set_bl')
    type(c_ptr), value :: qc_bl, qcfrac_bl
    bl_qc = qc_bl
    bl_frac = qcfrac_bl
  end subroutine ref_set_bl

  subroutine copy_in(src, dst, n1, n2)
    type(c_ptr), intent(in) :: src
    real(wp), allocatable, intent(inout) :: dst(:,:,:)
    integer, intent(in) :: n1, n2
    real(c_double), pointer :: p(:,:,:)
    if (c_associated(src)) then
      call c_f_pointer(src, p, [n1, n2, 1])
      allocate(dst(n1, n2, 1))
      dst = p
    endif
  end subroutine copy_in

  subroutine copy_out(dst, src, n1, n2)
    type(c_ptr), intent(in) :: dst
    real(wp), allocatable, intent(in) :: src(:,:,:)
    integer, intent(in) :: n1, n2
    real(c_double), pointer :: p(:,:,:)
    if (c_associated(dst) .and. allocated(src)) then
      call c_f_pointer(dst, p, [n1, n2, 1])
      p = src
    endif
  end subroutine copy_out

  subroutine copy_out2(dst, src, n1)
    type(c_ptr), intent(in) :: dst
    real(wp), allocatable, intent(in) :: src(:,:)
    integer, intent(in) :: n1
    real(c_double), pointer :: p(:,:)
    if (c_associated(dst) .and. allocated(src)) then
      call c_f_pointer(dst, p, [n1, 1])
      p = src
    endif
  end subroutine copy_out2

  ! state(1:19): t, p, w, dz, qv, qc, qr, qi, qs, qg, ni, nr, nc, nwfa, nifa,
  !              ng, qb, th, pii      (NULL = absent; p, w, dz, qv..nr required)
  ! diag2(1:8):  rain, ice, snow, graupel precip, frozen fraction,
  !              freezing rain, max hail sfc, max hail column
  ! diag3(1:7):  refl10cm, re_cloud, re_ice, re_snow, mvd_r, mvd_g, nc mixing ratio
  integer(c_int) function ref_run(ncol, nz, dt, itimestep, state, land, diag2, diag3) &
      bind(c, name='ref_run')
    integer(c_int), value :: ncol, nz, itimestep
    real(c_double), value :: dt
    type(c_ptr), intent(in) :: state(19), diag2(8), diag3(7)
    type(c_ptr), value :: land
    real(wp), allocatable, dimension(:,:,:) :: t, p, w, dz, qv, qc, qr, qi, qs, qg, &
      ni, nr, nc, nwfa, nifa, ng, qb, th, pii, qc_bl, qcfrac_bl
    integer(c_int), pointer :: pland(:,:)
    integer, allocatable :: land_input(:,:)
    integer :: n1

    n1 = ncol
    call copy_in(state(1), t, n1, nz)
    call copy_in(state(2), p, n1, nz)
    call copy_in(state(3), w, n1, nz)
    call copy_in(state(4), dz, n1, nz)
    call copy_in(state(5), qv, n1, nz)
    call copy_in(state(6), qc, n1, nz)
    call copy_in(state(7), qr, n1, nz)
    call copy_in(state(8), qi, n1, nz)
    call copy_in(state(9), qs, n1, nz)
    call copy_in(state(10), qg, n1, nz)
    call copy_in(state(11), ni, n1, nz)
    call copy_in(state(12), nr, n1, nz)
    call copy_in(state(13), nc, n1, nz)
    call copy_in(state(14), nwfa, n1, nz)
    call copy_in(state(15), nifa, n1, nz)
    call copy_in(state(16), ng, n1, nz)
    call copy_in(state(17), qb, n1, nz)
    call copy_in(state(18), th, n1, nz)
    call copy_in(state(19), pii, n1, nz)
    call copy_in(bl_qc, qc_bl, n1, nz)
    call copy_in(bl_frac, qcfrac_bl, n1, nz)
    if (c_associated(land)) then
      call c_f_pointer(land, pland, [n1, 1])
      allocate(land_input(n1, 1))
      land_input = pland
    endif

    call tempo_run(tempo_cfgs=cfgs, dt=real(dt, kind=wp), itimestep=int(itimestep), &
      t=t, th=th, pii=pii, p=p, w=w, dz=dz, &
      qv=qv, qc=qc, qr=qr, qi=qi, qs=qs, qg=qg, ni=ni, nr=nr, &
      nc=nc, nwfa=nwfa, nifa=nifa, ng=ng, qb=qb, land_input=land_input, &
      qc_bl=qc_bl, qcfrac_bl=qcfrac_bl, &
      ids=1, ide=n1, jds=1, jde=1, kds=1, kde=int(nz), &
      ims=1, ime=n1, jms=1, jme=1, kms=1, kme=int(nz), &
      its=1, ite=n1, jts=1, jte=1, kts=1, kte=int(nz), tempo_diags=diags)

    call copy_out(state(1), t, n1, nz)
    call copy_out(state(5), qv, n1, nz)
    call copy_out(state(6), qc, n1, nz)
    call copy_out(state(7), qr, n1, nz)
    call copy_out(state(8), qi, n1, nz)
    call copy_out(state(9), qs, n1, nz)
    call copy_out(state(10), qg, n1, nz)
    call copy_out(state(11), ni, n1, nz)
    call copy_out(state(12), nr, n1, nz)
    call copy_out(state(13), nc, n1, nz)
    call copy_out(state(14), nwfa, n1, nz)
    call copy_out(state(15), nifa, n1, nz)
    call copy_out(state(16), ng, n1, nz)
    call copy_out(state(17), qb, n1, nz)
    call copy_out(state(18), th, n1, nz)

    call copy_out2(diag2(1), diags%rain_precip, n1)
    call copy_out2(diag2(2), diags%ice_liquid_equiv_precip, n1)
    call copy_out2(diag2(3), diags%snow_liquid_equiv_precip, n1)
    call copy_out2(diag2(4), diags%graupel_liquid_equiv_precip, n1)
    call copy_out2(diag2(5), diags%frozen_fraction, n1)
    call copy_out2(diag2(6), diags%frz_rain_precip, n1)
    call copy_out2(diag2(7), diags%max_hail_diameter_sfc, n1)
    call copy_out2(diag2(8), diags%max_hail_diameter_column, n1)
    call copy_out(diag3(1), diags%refl10cm, n1, nz)
    call copy_out(diag3(2), diags%re_cloud, n1, nz)
    call copy_out(diag3(3), diags%re_ice, n1, nz)
    call copy_out(diag3(4), diags%re_snow, n1, nz)
    call copy_out(diag3(5), diags%rain_med_vol_diam, n1, nz)
    call copy_out(diag3(6), diags%graupel_med_vol_diam, n1, nz)
    call copy_out(diag3(7), diags%cloud_number_mixing_ratio, n1, nz)

    ! tempo_run allocates its diagnostics on the first call for the tile shape;
    ! release them so the next call may use another ncol / nz
    if (allocated(diags%rain_precip)) deallocate(diags%rain_precip, diags%ice_liquid_equiv_precip, &
      diags%snow_liquid_equiv_precip, diags%graupel_liquid_equiv_precip, diags%frozen_fraction, &
      diags%frz_rain_precip)
    if (allocated(diags%max_hail_diameter_sfc)) deallocate(diags%max_hail_diameter_sfc, &
      diags%max_hail_diameter_column)
    if (allocated(diags%refl10cm)) deallocate(diags%refl10cm)
    if (allocated(diags%re_cloud)) deallocate(diags%re_cloud)
    if (allocated(diags%re_ice)) deallocate(diags%re_ice)
    if (allocated(diags%re_snow)) deallocate(diags%re_snow)
    if (allocated(diags%rain_med_vol_diam)) deallocate(diags%rain_med_vol_diam)
    if (allocated(diags%graupel_med_vol_diam)) deallocate(diags%graupel_med_vol_diam)
    if (allocated(diags%cloud_number_mixing_ratio)) deallocate(diags%cloud_number_mixing_ratio)
    ref_run = 0
  end function ref_run

  ! The public check-and-update routines (src/module_mp_tempo_main.F90:18-20,
  ! 891-1134) on n points, FP64 arrays; logical flags exchanged as 0/1;
  ! has_nc / has_ncsave select the optional nc1d / ncsave arguments.
  subroutine ref_cloud_check(n, rho, l_qc, qc1d, nc1d, ncsave, rc, nc, qcten, ncten, ilamc, mvd_c, dt, has_nc, &
      has_ncsave) bind(c, name='ref_cloud_check')
    integer(c_int), value :: n, has_nc, has_ncsave
    real(c_double), value :: dt
    real(c_double), intent(in) :: rho(n), ncsave(n)
    real(c_double), intent(inout) :: qc1d(n), nc1d(n), rc(n), nc(n), qcten(n), ncten(n), ilamc(n), mvd_c(n)
    integer(c_int), intent(inout) :: l_qc(n)
    real(wp) :: w_rho(n), w_qc(n), w_nc(n), w_ncs(n), w_rc(n), w_ncv(n), w_qt(n), w_nt(n), w_mvd(n)
    real(dp) :: w_ilam(n)
    logical :: l(n)
    w_rho = real(rho, wp); w_qc = real(qc1d, wp); w_nc = real(nc1d, wp); w_ncs = real(ncsave, wp)
    w_rc = real(rc, wp); w_ncv = real(nc, wp); w_qt = real(qcten, wp); w_nt = real(ncten, wp)
    l = l_qc /= 0
    if (has_nc /= 0) then
      call cloud_check_and_update(rho=w_rho, l_qc=l, qc1d=w_qc, nc1d=w_nc, rc=w_rc, nc=w_ncv, qcten=w_qt, &
        ncten=w_nt, ilamc=w_ilam, mvd_c=w_mvd, dt=real(dt, wp), odt=1._wp/real(dt, wp))
    else if (has_ncsave /= 0) then
      call cloud_check_and_update(rho=w_rho, l_qc=l, qc1d=w_qc, ncsave=w_ncs, rc=w_rc, nc=w_ncv, qcten=w_qt, &
        ncten=w_nt, ilamc=w_ilam, mvd_c=w_mvd, dt=real(dt, wp), odt=1._wp/real(dt, wp))
    else
      call cloud_check_and_update(rho=w_rho, l_qc=l, qc1d=w_qc, rc=w_rc, nc=w_ncv, qcten=w_qt, &
        ncten=w_nt, ilamc=w_ilam, mvd_c=w_mvd, dt=real(dt, wp), odt=1._wp/real(dt, wp))
    endif
    l_qc = merge(1, 0, l)
    qc1d = w_qc; nc1d = w_nc; rc = w_rc; nc = w_ncv; qcten = w_qt; ncten = w_nt; ilamc = w_ilam; mvd_c = w_mvd
  end subroutine ref_cloud_check

  subroutine ref_ice_check(n, rho, l_qi, qi1d, ni1d, ri, ni, qiten, niten, ilami, dt) bind(c, name='ref_ice_check')
    integer(c_int), value :: n
    real(c_double), value :: dt
    real(c_double), intent(in) :: rho(n)
    real(c_double), intent(inout) :: qi1d(n), ni1d(n), ri(n), ni(n), qiten(n), niten(n), ilami(n)
    integer(c_int), intent(inout) :: l_qi(n)
    real(wp) :: w_rho(n), w_q(n), w_n(n), w_r(n), w_nv(n), w_qt(n), w_nt(n)
    real(dp) :: w_ilam(n)
    logical :: l(n)
    w_rho = real(rho, wp); w_q = real(qi1d, wp); w_n = real(ni1d, wp); w_r = real(ri, wp); w_nv = real(ni, wp)
    w_qt = real(qiten, wp); w_nt = real(niten, wp)
    l = l_qi /= 0
    call ice_check_and_update(w_rho, l, w_q, w_n, w_r, w_nv, w_qt, w_nt, w_ilam, real(dt, wp), 1._wp/real(dt, wp))
    l_qi = merge(1, 0, l)
    qi1d = w_q; ni1d = w_n; ri = w_r; ni = w_nv; qiten = w_qt; niten = w_nt; ilami = w_ilam
  end subroutine ref_ice_check

  subroutine ref_snow_check(n, rho, l_qs, qs1d, rs, qsten, dt) bind(c, name='ref_snow_check')
    integer(c_int), value :: n
    real(c_double), value :: dt
    real(c_double), intent(in) :: rho(n)
    real(c_double), intent(inout) :: qs1d(n), rs(n), qsten(n)
    integer(c_int), intent(inout) :: l_qs(n)
    real(wp) :: w_rho(n), w_q(n), w_r(n), w_qt(n)
    logical :: l(n)
    w_rho = real(rho, wp); w_q = real(qs1d, wp); w_r = real(rs, wp); w_qt = real(qsten, wp)
    l = l_qs /= 0
    call snow_check_and_update(w_rho, l, w_q, w_r, w_qt, real(dt, wp), 1._wp/real(dt, wp))
    l_qs = merge(1, 0, l)
    qs1d = w_q; rs = w_r; qsten = w_qt
  end subroutine ref_snow_check

  ! Known-answer hooks into public reference routines
  ! (src/module_mp_tempo_utils.F90:27-51, 148-331).
  ! tempo_aerosol_surface_emissions (src/module_mp_tempo_driver.F90:516-529)
  ! on nwfa(ni, nk, nj) and nwfa2d(ni, nj), both column-major FP64, with
  ! ims = jms = kms = 1; kts is 1-based
  subroutine ref_surface_emissions(dt, nwfa, nwfa2d, ni, nk, nj, kts) bind(c, name='ref_surface_emissions')
    real(c_double), value :: dt
    integer(c_int), value :: ni, nk, nj, kts
    real(c_double), intent(inout) :: nwfa(ni, nk, nj)
    real(c_double), intent(in) :: nwfa2d(ni, nj)
    real(wp), allocatable :: w3(:,:,:), w2(:,:)
    allocate(w3(ni, nk, nj), w2(ni, nj))
    w3 = real(nwfa, wp)
    w2 = real(nwfa2d, wp)
    call tempo_aerosol_surface_emissions(real(dt, wp), w3, w2, 1, ni, 1, nj, 1, nk, kts)
    nwfa = real(w3, c_double)
  end subroutine ref_surface_emissions

  real(c_double) function ref_gamma_p(a, x) bind(c, name='ref_gamma_p')
    real(c_double), value :: a, x
    ref_gamma_p = calc_gamma_p(real(a, wp), real(x, wp))
  end function ref_gamma_p

  real(c_double) function ref_rslf(p, t) bind(c, name='ref_rslf')
    real(c_double), value :: p, t
    ref_rslf = calc_rslf(real(p, wp), real(t, wp))
  end function ref_rslf

  real(c_double) function ref_rsif(p, t) bind(c, name='ref_rsif')
    real(c_double), value :: p, t
    ref_rsif = calc_rsif(real(p, wp), real(t, wp))
  end function ref_rsif

  ! out(1:10) = smob, smoc, ns, smo0, smo1, smo2, smoe, smof, smog, smoz
  subroutine ref_snow_moments(rs, tc, out) bind(c, name='ref_snow_moments')
    real(c_double), value :: rs, tc
    real(c_double), intent(out) :: out(10)
    real(dp) :: v(10)
    call snow_moments(real(rs, wp), real(tc, wp), v(1), v(2), v(3), v(4), v(5), v(6), &
      v(7), v(8), v(9), v(10))
    out = v
  end subroutine ref_snow_moments

  ! copies of init-time state for known-answer checks of the restatement:
  ! which = 1 t_efrw(100,100), 2 t_efsw(100,100), 3 tnc_wev(100,37,100),
  ! 4 tps_iaus(64,55), 5 tni_iaus, 6 tpi_ide, 7 bins [dc dtc di dti dr dtr ds dts dg dtg t_nc]
  integer(c_int) function ref_get_array(which, out, n) bind(c, name='ref_get_array')
    integer(c_int), value :: which, n
    real(c_double), intent(out) :: out(n)
    select case (which)
    case (1); out(1:size(t_efrw)) = reshape(t_efrw, [size(t_efrw)])
    case (2); out(1:size(t_efsw)) = reshape(t_efsw, [size(t_efsw)])
    case (3); out(1:size(tnc_wev)) = reshape(tnc_wev, [size(tnc_wev)])
    case (4); out(1:size(tps_iaus)) = reshape(tps_iaus, [size(tps_iaus)])
    case (5); out(1:size(tni_iaus)) = reshape(tni_iaus, [size(tni_iaus)])
    case (6); out(1:size(tpi_ide)) = reshape(tpi_ide, [size(tpi_ide)])
    case (7)
      out(1:100) = dc; out(101:200) = dtc; out(201:300) = di; out(301:400) = dti
      out(401:500) = dr; out(501:600) = dtr; out(601:700) = ds; out(701:800) = dts
      out(801:900) = dg; out(901:1000) = dtg; out(1001:1100) = t_nc
    case (8)
      ! scalar constants: ccg(5,15), cig(7), crg(13), csg(17), cgg(12,9)
      out(1:75) = reshape(ccg, [75]); out(76:82) = cig; out(83:95) = crg
      out(96:112) = csg; out(113:220) = reshape(cgg, [108])
    case default
      ref_get_array = 1
      return
    end select
    ref_get_array = 0
  end function ref_get_array

end module ref_bridge
