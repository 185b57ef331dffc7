! Fortran host binding of the MI355X TEMPO column step (libtempo_hip.so).
!
! A host model that calls the reference driver
!     use module_mp_tempo_driver, only : tempo_init, tempo_run, ty_tempo_driver_diags
! switches to this library with
!     use tempo_hip_mod, only : tempo_init, tempo_run, ty_tempo_driver_diags, ty_tempo_cfgs
! and keeps its calls: tempo_init takes the same optional flags
! (src/module_mp_tempo_driver.F90:42-54), tempo_run the same argument list
! and memory/tile bounds (src/module_mp_tempo_driver.F90:187-279), and the
! diagnostics come back in a ty_tempo_driver_diags with the reference's
! component names (:18-34), allocated on the first call as the reference
! does (:304-386).  Arrays are the host's real(8) arrays; the library stages
! them through device memory (tempo_hip_run_host), without host copies for
! arrays allocated with tempo_alloc_page_locked.  Hosts whose state already
! lives in device memory call the C entry point tempo_hip_run directly (the
! interface block below).
!
! The reference's unused tempo_run arguments (qc_bl, qcfrac_bl, qcfrac,
! qifrac, thten_*, q*ten_bl, thten_*rad; never read by tempo_main,
! src/module_mp_tempo_main.F90:76-79,109-118) are accepted and ignored.
module tempo_hip_mod
  use iso_c_binding
  implicit none
  private

  public :: tempo_init, tempo_run, tempo_finalize, ty_tempo_cfgs, ty_tempo_driver_diags
  public :: tempo_aerosol_surface_emissions
  public :: cloud_check_and_update, ice_check_and_update, snow_check_and_update
  public :: tempo_hip_cfg, tempo_hip_fields, tempo_hip_diags, tempo_hip_dims
  public :: tempo_hip_cfg_default, tempo_hip_init, tempo_hip_run, tempo_hip_run_host, tempo_hip_run_host_prec
  public :: tempo_hip_last_error
  public :: tempo_hip_host_alloc, tempo_hip_host_free, tempo_alloc_page_locked, tempo_free_page_locked

  integer, parameter :: wp = c_double

  ! ---- C ABI (include/tempo_hip.h) -------------------------------------
  type, bind(c) :: tempo_hip_cfg
    integer(c_int) :: aerosolaware_flag, hailaware_flag, ml_for_bl_nc_flag, ml_for_nc_flag, semi_sedi_flag
    integer(c_int) :: refl10cm_from_melting_flag, turn_off_micro_flag, cloud_condensation_flag, verbose
    integer(c_int) :: check_tables, nc_option
    real(c_double) :: nc_value
    integer(c_int) :: refl10cm_flag, re_cloud_flag, re_ice_flag, re_snow_flag, max_hail_diameter_flag
    integer(c_int) :: rain_med_vol_diam_flag, graupel_med_vol_diam_flag, cloud_number_mixing_ratio_flag
  end type

  type, bind(c) :: tempo_hip_fields
    type(c_ptr) :: t = c_null_ptr, th = c_null_ptr, pii = c_null_ptr, p = c_null_ptr, w = c_null_ptr
    type(c_ptr) :: dz = c_null_ptr, qv = c_null_ptr, qc = c_null_ptr, qr = c_null_ptr, qi = c_null_ptr
    type(c_ptr) :: qs = c_null_ptr, qg = c_null_ptr, ni = c_null_ptr, nr = c_null_ptr, nc = c_null_ptr
    type(c_ptr) :: nwfa = c_null_ptr, nifa = c_null_ptr, ng = c_null_ptr, qb = c_null_ptr
    type(c_ptr) :: land = c_null_ptr
    type(c_ptr) :: qc_bl = c_null_ptr, qcfrac_bl = c_null_ptr
  end type

  type, bind(c) :: tempo_hip_diags
    type(c_ptr) :: rain_precip = c_null_ptr, ice_liquid_equiv_precip = c_null_ptr
    type(c_ptr) :: snow_liquid_equiv_precip = c_null_ptr, graupel_liquid_equiv_precip = c_null_ptr
    type(c_ptr) :: frozen_fraction = c_null_ptr, frz_rain_precip = c_null_ptr
    type(c_ptr) :: max_hail_diameter_sfc = c_null_ptr, max_hail_diameter_column = c_null_ptr
    type(c_ptr) :: refl10cm = c_null_ptr, re_cloud = c_null_ptr, re_ice = c_null_ptr, re_snow = c_null_ptr
    type(c_ptr) :: rain_med_vol_diam = c_null_ptr, graupel_med_vol_diam = c_null_ptr
    type(c_ptr) :: cloud_number_mixing_ratio = c_null_ptr
  end type

  type, bind(c) :: tempo_hip_dims
    integer(c_int) :: ni, nj, nz
    integer(c_long_long) :: ldk, ldj, dldk, dldj, d2ldj
    type(c_ptr) :: nlev = c_null_ptr
  end type

  interface
    subroutine tempo_hip_cfg_default(cfg) bind(c, name='tempo_hip_cfg_default')
      import :: tempo_hip_cfg
      type(tempo_hip_cfg), intent(out) :: cfg
    end subroutine
    integer(c_int) function tempo_hip_init(cfg, table_dir, ccn_path, ctx) bind(c, name='tempo_hip_init')
      ! table_dir = c_null_ptr: build the tables on the GPU
      import :: tempo_hip_cfg, c_int, c_ptr
      type(tempo_hip_cfg), intent(in) :: cfg
      type(c_ptr), value :: table_dir, ccn_path
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function tempo_hip_run(ctx, f, d, dims, dt, itimestep, precision, stream) &
        bind(c, name='tempo_hip_run')
      import :: tempo_hip_fields, tempo_hip_diags, tempo_hip_dims, c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, stream
      type(tempo_hip_fields), intent(in) :: f
      type(tempo_hip_diags), intent(in) :: d
      type(tempo_hip_dims), intent(in) :: dims
      real(c_double), value :: dt
      integer(c_int), value :: itimestep, precision
    end function
    integer(c_int) function tempo_hip_run_host(ctx, f, d, dims, dt, itimestep) bind(c, name='tempo_hip_run_host')
      import :: tempo_hip_fields, tempo_hip_diags, tempo_hip_dims, c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      type(tempo_hip_fields), intent(in) :: f
      type(tempo_hip_diags), intent(in) :: d
      type(tempo_hip_dims), intent(in) :: dims
      real(c_double), value :: dt
      integer(c_int), value :: itimestep
    end function
    integer(c_int) function tempo_hip_run_host_prec(ctx, f, d, dims, dt, itimestep, precision) &
        bind(c, name='tempo_hip_run_host_prec')
      import :: tempo_hip_fields, tempo_hip_diags, tempo_hip_dims, c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      type(tempo_hip_fields), intent(in) :: f
      type(tempo_hip_diags), intent(in) :: d
      type(tempo_hip_dims), intent(in) :: dims
      real(c_double), value :: dt
      integer(c_int), value :: itimestep, precision
    end function
    integer(c_int) function tempo_hip_host_alloc(bytes, ptr) bind(c, name='tempo_hip_host_alloc')
      import :: c_int, c_size_t, c_ptr
      integer(c_size_t), value :: bytes
      type(c_ptr), intent(out) :: ptr
    end function
    integer(c_int) function tempo_hip_host_free(ptr) bind(c, name='tempo_hip_host_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: ptr
    end function
    type(c_ptr) function tempo_hip_last_error_c() bind(c, name='tempo_hip_last_error')
      import :: c_ptr
    end function
    integer(c_int) function tempo_hip_set_run_cfg(ctx, cfg) bind(c, name='tempo_hip_set_run_cfg')
      import :: c_ptr, c_int, tempo_hip_cfg
      type(c_ptr), value :: ctx
      type(tempo_hip_cfg), intent(in) :: cfg
    end function
    subroutine tempo_hip_finalize(ctx) bind(c, name='tempo_hip_finalize')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    integer(c_int) function tempo_hip_cloud_check_and_update_host(ctx, n, rho, l_qc, qc1d, nc1d, ncsave, rc, nc, &
        qcten, ncten, ilamc, mvd_c, dt) bind(c, name='tempo_hip_cloud_check_and_update_host')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx, rho, l_qc, qc1d, nc1d, ncsave, rc, nc, qcten, ncten, ilamc, mvd_c
      integer(c_int), value :: n
      real(c_double), value :: dt
    end function
    integer(c_int) function tempo_hip_ice_check_and_update_host(ctx, n, rho, l_qi, qi1d, ni1d, ri, ni, qiten, &
        niten, ilami, dt) bind(c, name='tempo_hip_ice_check_and_update_host')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx, rho, l_qi, qi1d, ni1d, ri, ni, qiten, niten, ilami
      integer(c_int), value :: n
      real(c_double), value :: dt
    end function
    integer(c_int) function tempo_hip_snow_check_and_update_host(ctx, n, rho, l_qs, qs1d, rs, qsten, dt) &
        bind(c, name='tempo_hip_snow_check_and_update_host')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx, rho, l_qs, qs1d, rs, qsten
      integer(c_int), value :: n
      real(c_double), value :: dt
    end function
    integer(c_int) function tempo_hip_aerosol_surface_emissions_host(ctx, dt, nwfa, nwfa2d, dims, k) &
        bind(c, name='tempo_hip_aerosol_surface_emissions_host')
      import :: c_ptr, c_int, c_double, tempo_hip_dims
      type(c_ptr), value :: ctx, nwfa, nwfa2d
      real(c_double), value :: dt
      type(tempo_hip_dims), intent(in) :: dims
      integer(c_int), value :: k
    end function
  end interface

  ! context of the last tempo_init, for tempo_aerosol_surface_emissions, whose
  ! reference signature carries no configuration
  type(c_ptr), save :: last_ctx = c_null_ptr

  ! ---- reference-shaped types ------------------------------------------
  ! ty_tempo_cfgs (src/module_mp_tempo_cfgs.F90:14-38) plus the library
  ! context it owns after tempo_init
  type :: ty_tempo_cfgs
    logical :: aerosolaware_flag = .true.
    logical :: hailaware_flag = .true.
    logical :: ml_for_bl_nc_flag = .false.
    logical :: ml_for_nc_flag = .false.
    logical :: semi_sedi_flag = .false.
    logical :: refl10cm_from_melting_flag = .false.
    logical :: turn_off_micro_flag = .false.
    logical :: cloud_condensation_flag = .true.
    logical :: verbose = .false.
    logical :: check_tables = .false.
    character(len=32) :: single_moment_nc_opt = 'land'
    logical :: refl10cm_flag = .true.
    logical :: re_cloud_flag = .true.
    logical :: re_ice_flag = .true.
    logical :: re_snow_flag = .true.
    logical :: max_hail_diameter_flag = .true.
    logical :: rain_med_vol_diam_flag = .false.
    logical :: graupel_med_vol_diam_flag = .false.
    logical :: cloud_number_mixing_ratio_flag = .false.
    ! the step on host arrays in single precision on the device
    ! (tempo_hip_run_host_prec; the arrays stay real(8))
    logical :: single_precision_step = .false.
    character(len=256) :: table_dir = ''   ! '' = build the tables on the GPU
    character(len=256) :: ccn_path = ''    ! '' = <table_dir>/ccn_activate.bin
    type(c_ptr) :: ctx = c_null_ptr
  end type

  ! ty_tempo_driver_diags (src/module_mp_tempo_driver.F90:18-34)
  type :: ty_tempo_driver_diags
    real(wp), dimension(:,:), allocatable :: rain_precip
    real(wp), dimension(:,:), allocatable :: ice_liquid_equiv_precip
    real(wp), dimension(:,:), allocatable :: snow_liquid_equiv_precip
    real(wp), dimension(:,:), allocatable :: graupel_liquid_equiv_precip
    real(wp), dimension(:,:), allocatable :: frozen_fraction
    real(wp), dimension(:,:), allocatable :: frz_rain_precip
    real(wp), dimension(:,:), allocatable :: max_hail_diameter_sfc
    real(wp), dimension(:,:), allocatable :: max_hail_diameter_column
    real(wp), dimension(:,:,:), allocatable :: refl10cm
    real(wp), dimension(:,:,:), allocatable :: re_cloud
    real(wp), dimension(:,:,:), allocatable :: re_ice
    real(wp), dimension(:,:,:), allocatable :: re_snow
    real(wp), dimension(:,:,:), allocatable :: rain_med_vol_diam
    real(wp), dimension(:,:,:), allocatable :: graupel_med_vol_diam
    real(wp), dimension(:,:,:), allocatable :: cloud_number_mixing_ratio
  end type

contains

  function tempo_hip_last_error() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: n
    type(c_ptr) :: p
    p = tempo_hip_last_error_c()
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function

  subroutine fail(what, rc)
    character(len=*), intent(in) :: what
    integer(c_int), intent(in) :: rc
    write(*, '(a,a,i0,a,a)') what, ' failed with status ', rc, ': ', tempo_hip_last_error()
    error stop 1
  end subroutine

  ! NUL-terminated copy of a Fortran string for the C ABI
  function cstr(str) result(a)
    character(len=*), intent(in) :: str
    character(kind=c_char), allocatable :: a(:)
    integer :: i
    allocate(a(len(str) + 1))
    do i = 1, len(str)
      a(i) = str(i:i)
    end do
    a(len(str) + 1) = c_null_char
  end function

  integer(c_int) function flag(l)
    logical, intent(in) :: l
    flag = merge(1_c_int, 0_c_int, l)
  end function

  !> ty_tempo_cfgs -> tempo_hip_cfg, with resolve_nc_value
  !! (src/module_mp_tempo_cfgs.F90:50-73)
  subroutine to_c_cfg(tempo_cfgs, c)
    type(ty_tempo_cfgs), intent(in) :: tempo_cfgs
    type(tempo_hip_cfg), intent(out) :: c
    call tempo_hip_cfg_default(c)
    c%aerosolaware_flag = flag(tempo_cfgs%aerosolaware_flag)
    c%hailaware_flag = flag(tempo_cfgs%hailaware_flag)
    c%ml_for_bl_nc_flag = flag(tempo_cfgs%ml_for_bl_nc_flag)
    c%ml_for_nc_flag = flag(tempo_cfgs%ml_for_nc_flag)
    c%semi_sedi_flag = flag(tempo_cfgs%semi_sedi_flag)
    c%refl10cm_from_melting_flag = flag(tempo_cfgs%refl10cm_from_melting_flag)
    c%turn_off_micro_flag = flag(tempo_cfgs%turn_off_micro_flag)
    c%cloud_condensation_flag = flag(tempo_cfgs%cloud_condensation_flag)
    c%verbose = flag(tempo_cfgs%verbose)
    c%check_tables = flag(tempo_cfgs%check_tables)
    select case (trim(tempo_cfgs%single_moment_nc_opt))
    case ('land')
      c%nc_option = 0
    case ('ocean')
      c%nc_option = 1
    case default
      c%nc_option = 2
      read(tempo_cfgs%single_moment_nc_opt, *) c%nc_value
    end select
    c%refl10cm_flag = flag(tempo_cfgs%refl10cm_flag)
    c%re_cloud_flag = flag(tempo_cfgs%re_cloud_flag)
    c%re_ice_flag = flag(tempo_cfgs%re_ice_flag)
    c%re_snow_flag = flag(tempo_cfgs%re_snow_flag)
    c%max_hail_diameter_flag = flag(tempo_cfgs%max_hail_diameter_flag)
    c%rain_med_vol_diam_flag = flag(tempo_cfgs%rain_med_vol_diam_flag)
    c%graupel_med_vol_diam_flag = flag(tempo_cfgs%graupel_med_vol_diam_flag)
    c%cloud_number_mixing_ratio_flag = flag(tempo_cfgs%cloud_number_mixing_ratio_flag)
  end subroutine to_c_cfg

  !> tempo_init (src/module_mp_tempo_driver.F90:42-182): flags, then the
  !! library context (constants, init-time tables, collection/freezing tables
  !! read from tempo_cfgs%table_dir or built on the GPU).  force_init_flag
  !! re-creates the context.
  subroutine tempo_init(aerosolaware_flag, hailaware_flag, semi_sedi_flag, cloud_condensation_flag, &
      refl10cm_from_melting_flag, ml_for_bl_nc_flag, ml_for_nc_flag, force_init_flag, tempo_cfgs)
    logical, intent(in), optional :: aerosolaware_flag, hailaware_flag, refl10cm_from_melting_flag, &
      ml_for_bl_nc_flag, ml_for_nc_flag, force_init_flag, semi_sedi_flag, cloud_condensation_flag
    type(ty_tempo_cfgs), intent(inout) :: tempo_cfgs
    type(tempo_hip_cfg) :: c
    integer(c_int) :: rc
    character(len=:), allocatable :: ccn
    character(kind=c_char), allocatable, target :: cdir(:), cccn(:)

    if (c_associated(tempo_cfgs%ctx)) then
      if (.not. present(force_init_flag)) return
      if (.not. force_init_flag) return
      call tempo_hip_finalize(tempo_cfgs%ctx)
      tempo_cfgs%ctx = c_null_ptr
    endif
    if (present(aerosolaware_flag)) tempo_cfgs%aerosolaware_flag = aerosolaware_flag
    if (present(hailaware_flag)) tempo_cfgs%hailaware_flag = hailaware_flag
    if (present(ml_for_bl_nc_flag)) tempo_cfgs%ml_for_bl_nc_flag = ml_for_bl_nc_flag
    if (present(ml_for_nc_flag)) tempo_cfgs%ml_for_nc_flag = ml_for_nc_flag
    if (present(semi_sedi_flag)) tempo_cfgs%semi_sedi_flag = semi_sedi_flag
    if (present(cloud_condensation_flag)) tempo_cfgs%cloud_condensation_flag = cloud_condensation_flag
    if (present(refl10cm_from_melting_flag)) tempo_cfgs%refl10cm_from_melting_flag = refl10cm_from_melting_flag

    call to_c_cfg(tempo_cfgs, c)

    if (len_trim(tempo_cfgs%ccn_path) > 0) then
      ccn = trim(tempo_cfgs%ccn_path)
    else
      ccn = trim(tempo_cfgs%table_dir) // '/ccn_activate.bin'
    endif
    cccn = cstr(ccn)
    if (len_trim(tempo_cfgs%table_dir) > 0) then
      cdir = cstr(trim(tempo_cfgs%table_dir))
      rc = tempo_hip_init(c, c_loc(cdir), c_loc(cccn), tempo_cfgs%ctx)
    else
      rc = tempo_hip_init(c, c_null_ptr, c_loc(cccn), tempo_cfgs%ctx)
    endif
    if (rc /= 0) call fail('tempo_init', rc)
    last_ctx = tempo_cfgs%ctx
  end subroutine

  ! A real(8) array with bounds lb:ub in page-locked host memory
  ! (tempo_hip_host_alloc; no counterpart in the reference).  tempo_run moves
  ! state and diagnostic arrays held in such memory between host and device
  ! without staging copies.  Release with tempo_free_page_locked.
  subroutine tempo_alloc_page_locked(a, lb, ub)
    real(wp), pointer, intent(out) :: a(:,:,:)
    integer, intent(in) :: lb(3), ub(3)
    real(wp), pointer :: flat(:,:,:)
    type(c_ptr) :: p
    integer(c_size_t) :: n
    n = int(ub(1) - lb(1) + 1, c_size_t) * int(ub(2) - lb(2) + 1, c_size_t) * int(ub(3) - lb(3) + 1, c_size_t)
    if (tempo_hip_host_alloc(max(n, 1_c_size_t) * 8_c_size_t, p) /= 0) &
      error stop 'tempo_alloc_page_locked: tempo_hip_host_alloc failed'
    call c_f_pointer(p, flat, ub - lb + 1)
    a(lb(1):, lb(2):, lb(3):) => flat
  end subroutine tempo_alloc_page_locked

  subroutine tempo_free_page_locked(a)
    real(wp), pointer, intent(inout) :: a(:,:,:)
    if (.not. associated(a)) return
    if (tempo_hip_host_free(c_loc(a)) /= 0) error stop 'tempo_free_page_locked: tempo_hip_host_free failed'
    nullify(a)
  end subroutine tempo_free_page_locked

  subroutine tempo_finalize(tempo_cfgs)
    type(ty_tempo_cfgs), intent(inout) :: tempo_cfgs
    if (c_associated(tempo_cfgs%ctx)) call tempo_hip_finalize(tempo_cfgs%ctx)
    if (c_associated(last_ctx, tempo_cfgs%ctx)) last_ctx = c_null_ptr
    tempo_cfgs%ctx = c_null_ptr
  end subroutine

  !> tempo_aerosol_surface_emissions (src/module_mp_tempo_driver.F90:516-529):
  !! nwfa(i, kts, j) += nwfa2d(i, j) * dt over the memory extent, on the
  !! device of the last tempo_init.
  subroutine tempo_aerosol_surface_emissions(dt, nwfa, nwfa2d, ims, ime, jms, jme, kms, kme, kts)
    real(wp), intent(in) :: dt
    integer, intent(in) :: ims, ime, jms, jme, kms, kme, kts
    real(wp), dimension(ims:ime, kms:kme, jms:jme), intent(inout), target :: nwfa
    real(wp), dimension(ims:ime, jms:jme), intent(in), target :: nwfa2d
    type(tempo_hip_dims) :: dm
    integer(c_int) :: rc
    if (.not. c_associated(last_ctx)) call fail('tempo_aerosol_surface_emissions (call tempo_init first)', -1)
    dm%ni = ime - ims + 1
    dm%nj = jme - jms + 1
    dm%nz = kme - kms + 1
    dm%ldk = dm%ni
    dm%ldj = int(dm%ni, c_long_long) * dm%nz
    dm%dldk = dm%ldk
    dm%dldj = dm%ldj
    dm%d2ldj = dm%ni
    rc = tempo_hip_aerosol_surface_emissions_host(last_ctx, real(dt, c_double), c_loc(nwfa), c_loc(nwfa2d), &
                                                  dm, int(kts - kms, c_int))
    if (rc /= 0) call fail('tempo_aerosol_surface_emissions', rc)
  end subroutine

  !> cloud_check_and_update (src/module_mp_tempo_main.F90:891-973) with the
  !! reference's arguments (odt unused: the library forms 1/dt), on the
  !! context of the last tempo_init
  subroutine cloud_check_and_update(rho, l_qc, qc1d, nc1d, ncsave, rc, nc, qcten, ncten, ilamc, mvd_c, dt, odt)
    real(wp), intent(in) :: dt, odt
    real(wp), dimension(:), intent(in), target :: rho
    real(wp), dimension(:), intent(inout), target :: qc1d, qcten, ncten, rc, nc
    real(wp), dimension(:), intent(out), target :: mvd_c, ilamc
    real(wp), dimension(:), intent(inout), optional, target :: nc1d, ncsave
    logical, dimension(:), intent(inout) :: l_qc
    integer(c_int), allocatable, target :: l(:)
    type(c_ptr) :: pn, ps
    integer(c_int) :: rc_
    if (.not. c_associated(last_ctx)) call fail('cloud_check_and_update (call tempo_init first)', -1)
    allocate(l(size(qc1d)))
    pn = c_null_ptr
    ps = c_null_ptr
    if (present(nc1d)) pn = c_loc(nc1d)
    if (present(ncsave)) ps = c_loc(ncsave)
    rc_ = tempo_hip_cloud_check_and_update_host(last_ctx, int(size(qc1d), c_int), c_loc(rho), c_loc(l), c_loc(qc1d), &
      pn, ps, c_loc(rc), c_loc(nc), c_loc(qcten), c_loc(ncten), c_loc(ilamc), c_loc(mvd_c), real(dt, c_double))
    if (rc_ /= 0) call fail('cloud_check_and_update', rc_)
    l_qc = l /= 0
  end subroutine

  !> ice_check_and_update (src/module_mp_tempo_main.F90:1045-1107); odt unused
  subroutine ice_check_and_update(rho, l_qi, qi1d, ni1d, ri, ni, qiten, niten, ilami, dt, odt)
    real(wp), intent(in) :: dt, odt
    real(wp), dimension(:), intent(in), target :: rho
    real(wp), dimension(:), intent(inout), target :: qi1d, ni1d, qiten, niten, ri, ni
    real(wp), dimension(:), intent(out), target :: ilami
    logical, dimension(:), intent(inout) :: l_qi
    integer(c_int), allocatable, target :: l(:)
    integer(c_int) :: rc_
    if (.not. c_associated(last_ctx)) call fail('ice_check_and_update (call tempo_init first)', -1)
    allocate(l(size(qi1d)))
    rc_ = tempo_hip_ice_check_and_update_host(last_ctx, int(size(qi1d), c_int), c_loc(rho), c_loc(l), c_loc(qi1d), &
      c_loc(ni1d), c_loc(ri), c_loc(ni), c_loc(qiten), c_loc(niten), c_loc(ilami), real(dt, c_double))
    if (rc_ /= 0) call fail('ice_check_and_update', rc_)
    l_qi = l /= 0
  end subroutine

  !> snow_check_and_update (src/module_mp_tempo_main.F90:1110-1134); odt unused
  subroutine snow_check_and_update(rho, l_qs, qs1d, rs, qsten, dt, odt)
    real(wp), intent(in) :: dt, odt
    real(wp), dimension(:), intent(in), target :: rho
    real(wp), dimension(:), intent(inout), target :: qs1d, rs, qsten
    logical, dimension(:), intent(inout) :: l_qs
    integer(c_int), allocatable, target :: l(:)
    integer(c_int) :: rc_
    if (.not. c_associated(last_ctx)) call fail('snow_check_and_update (call tempo_init first)', -1)
    allocate(l(size(qs1d)))
    rc_ = tempo_hip_snow_check_and_update_host(last_ctx, int(size(qs1d), c_int), c_loc(rho), c_loc(l), c_loc(qs1d), &
      c_loc(rs), c_loc(qsten), real(dt, c_double))
    if (rc_ /= 0) call fail('snow_check_and_update', rc_)
    l_qs = l /= 0
  end subroutine

  !> tempo_run (src/module_mp_tempo_driver.F90:187-513) over the tile
  !! its:ite, kts:kte, jts:jte of the host's (ims:ime, kms:kme, jms:jme)
  !! arrays.
  subroutine tempo_run(tempo_cfgs, dt, itimestep, &
      t, th, pii, p, w, dz, &
      qv, qc, qr, qi, qs, qg, ni, nr, &
      nc, nwfa, nifa, ng, qb, &
      qc_bl, qcfrac_bl, &
      qcfrac, qifrac, &
      thten_bl, qvten_bl, qcten_bl, qiten_bl, &
      thten_lwrad, thten_swrad, &
      land_input, &
      ids, ide, jds, jde, kds, kde, &
      ims, ime, jms, jme, kms, kme, &
      its, ite, jts, jte, kts, kte, tempo_diags)
    type(ty_tempo_cfgs), intent(in) :: tempo_cfgs
    real(wp), intent(in) :: dt
    integer, intent(in) :: itimestep
    integer, intent(in) :: ids, ide, jds, jde, kds, kde
    integer, intent(in) :: ims, ime, jms, jme, kms, kme
    integer, intent(in) :: its, ite, jts, jte, kts, kte
    real(wp), dimension(ims:ime, kms:kme, jms:jme), intent(inout), optional, target :: t, th
    real(wp), dimension(ims:ime, kms:kme, jms:jme), intent(in), target :: p, w, dz
    real(wp), dimension(ims:ime, kms:kme, jms:jme), intent(in), optional, target :: pii
    real(wp), dimension(ims:ime, kms:kme, jms:jme), intent(inout), target :: qv, qc, qr, qi, qs, qg, ni, nr
    real(wp), dimension(ims:ime, kms:kme, jms:jme), intent(inout), optional, target :: nc, nwfa, nifa, qb, ng
    integer, dimension(ims:ime, jms:jme), intent(in), optional :: land_input
    real(wp), dimension(ims:ime, kms:kme, jms:jme), intent(inout), optional :: qcfrac, qifrac
    real(wp), dimension(ims:ime, kms:kme, jms:jme), intent(in), optional, target :: qc_bl, qcfrac_bl
    real(wp), dimension(ims:ime, kms:kme, jms:jme), intent(in), optional :: thten_bl, &
      qvten_bl, qcten_bl, qiten_bl, thten_lwrad, thten_swrad
    type(ty_tempo_driver_diags), intent(inout), target :: tempo_diags

    type(tempo_hip_fields) :: f
    type(tempo_hip_diags) :: d
    type(tempo_hip_dims) :: dm
    type(tempo_hip_cfg) :: cc
    integer(c_int), allocatable, target :: land(:)
    integer(c_int) :: rc
    integer :: i, j, n

    if (.not. c_associated(tempo_cfgs%ctx)) then
      write(*, '(a)') 'tempo_run: tempo_init has not been called'
      error stop 1
    endif
    if (.not. (present(t) .or. (present(th) .and. present(pii)))) then
      write(*, '(a)') 'tempo_run: need t, or th and pii'
      error stop 1
    endif
    ! tempo_run reads the run flags of tempo_cfgs on every call
    call to_c_cfg(tempo_cfgs, cc)
    rc = tempo_hip_set_run_cfg(tempo_cfgs%ctx, cc)
    if (rc /= 0) call fail('tempo_run', rc)
    call allocate_diags(tempo_cfgs, tempo_diags, its, ite, jts, jte, kts, kte)

    dm%ni = ite - its + 1
    dm%nj = jte - jts + 1
    dm%nz = kte - kts + 1
    dm%ldk = ime - ims + 1
    dm%ldj = dm%ldk * (kme - kms + 1)
    ! the diagnostics are tile-shaped (its:ite, kts:kte, jts:jte), as the
    ! reference allocates them
    dm%dldk = dm%ni
    dm%dldj = int(dm%ni, c_long_long) * dm%nz
    dm%d2ldj = dm%ni

    if (present(t)) f%t = c_loc(t(its, kts, jts))
    if (present(th)) f%th = c_loc(th(its, kts, jts))
    if (present(pii)) f%pii = c_loc(pii(its, kts, jts))
    f%p = c_loc(p(its, kts, jts))
    f%w = c_loc(w(its, kts, jts))
    f%dz = c_loc(dz(its, kts, jts))
    f%qv = c_loc(qv(its, kts, jts))
    f%qc = c_loc(qc(its, kts, jts))
    f%qr = c_loc(qr(its, kts, jts))
    f%qi = c_loc(qi(its, kts, jts))
    f%qs = c_loc(qs(its, kts, jts))
    f%qg = c_loc(qg(its, kts, jts))
    f%ni = c_loc(ni(its, kts, jts))
    f%nr = c_loc(nr(its, kts, jts))
    if (present(nc)) f%nc = c_loc(nc(its, kts, jts))
    if (present(nwfa)) f%nwfa = c_loc(nwfa(its, kts, jts))
    if (present(nifa)) f%nifa = c_loc(nifa(its, kts, jts))
    if (present(ng)) f%ng = c_loc(ng(its, kts, jts))
    if (present(qb)) f%qb = c_loc(qb(its, kts, jts))
    if (present(qc_bl) .and. present(qcfrac_bl)) then
      f%qc_bl = c_loc(qc_bl(its, kts, jts))
      f%qcfrac_bl = c_loc(qcfrac_bl(its, kts, jts))
    endif
    if (present(land_input)) then
      allocate(land(dm%ni * dm%nj))
      n = 0
      do j = jts, jte
        do i = its, ite
          n = n + 1
          land(n) = land_input(i, j)
        end do
      end do
      f%land = c_loc(land)
    endif

    d%rain_precip = c_loc(tempo_diags%rain_precip(its, jts))
    d%ice_liquid_equiv_precip = c_loc(tempo_diags%ice_liquid_equiv_precip(its, jts))
    d%snow_liquid_equiv_precip = c_loc(tempo_diags%snow_liquid_equiv_precip(its, jts))
    d%graupel_liquid_equiv_precip = c_loc(tempo_diags%graupel_liquid_equiv_precip(its, jts))
    d%frozen_fraction = c_loc(tempo_diags%frozen_fraction(its, jts))
    d%frz_rain_precip = c_loc(tempo_diags%frz_rain_precip(its, jts))
    if (allocated(tempo_diags%max_hail_diameter_sfc)) then
      d%max_hail_diameter_sfc = c_loc(tempo_diags%max_hail_diameter_sfc(its, jts))
      d%max_hail_diameter_column = c_loc(tempo_diags%max_hail_diameter_column(its, jts))
    endif
    if (allocated(tempo_diags%refl10cm)) d%refl10cm = c_loc(tempo_diags%refl10cm(its, kts, jts))
    if (allocated(tempo_diags%re_cloud)) d%re_cloud = c_loc(tempo_diags%re_cloud(its, kts, jts))
    if (allocated(tempo_diags%re_ice)) d%re_ice = c_loc(tempo_diags%re_ice(its, kts, jts))
    if (allocated(tempo_diags%re_snow)) d%re_snow = c_loc(tempo_diags%re_snow(its, kts, jts))
    if (allocated(tempo_diags%rain_med_vol_diam)) &
      d%rain_med_vol_diam = c_loc(tempo_diags%rain_med_vol_diam(its, kts, jts))
    if (allocated(tempo_diags%graupel_med_vol_diam)) &
      d%graupel_med_vol_diam = c_loc(tempo_diags%graupel_med_vol_diam(its, kts, jts))
    if (allocated(tempo_diags%cloud_number_mixing_ratio)) &
      d%cloud_number_mixing_ratio = c_loc(tempo_diags%cloud_number_mixing_ratio(its, kts, jts))

    if (tempo_cfgs%single_precision_step) then
      rc = tempo_hip_run_host_prec(tempo_cfgs%ctx, f, d, dm, real(dt, c_double), int(itimestep, c_int), 4_c_int)
    else
      rc = tempo_hip_run_host(tempo_cfgs%ctx, f, d, dm, real(dt, c_double), int(itimestep, c_int))
    endif
    if (rc /= 0) call fail('tempo_run', rc)
  end subroutine

  ! allocation and reset of the driver diagnostics on the tile
  ! (its:ite, kts:kte, jts:jte), as src/module_mp_tempo_driver.F90:304-386
  ! does; an array the caller allocated with other bounds is reallocated
  subroutine allocate_diags(cfg, dg, its, ite, jts, jte, kts, kte)
    type(ty_tempo_cfgs), intent(in) :: cfg
    type(ty_tempo_driver_diags), intent(inout) :: dg
    integer, intent(in) :: its, ite, jts, jte, kts, kte
    call alloc2(.true., dg%rain_precip)
    call alloc2(.true., dg%ice_liquid_equiv_precip)
    call alloc2(.true., dg%snow_liquid_equiv_precip)
    call alloc2(.true., dg%graupel_liquid_equiv_precip)
    call alloc2(.true., dg%frozen_fraction)
    call alloc2(.true., dg%frz_rain_precip)
    call alloc2(cfg%max_hail_diameter_flag, dg%max_hail_diameter_sfc)
    call alloc2(cfg%max_hail_diameter_flag, dg%max_hail_diameter_column)
    call alloc3(cfg%refl10cm_flag, dg%refl10cm, -35._wp)
    call alloc3(cfg%re_cloud_flag, dg%re_cloud, 0._wp)
    call alloc3(cfg%re_ice_flag, dg%re_ice, 0._wp)
    call alloc3(cfg%re_snow_flag, dg%re_snow, 0._wp)
    call alloc3(cfg%rain_med_vol_diam_flag, dg%rain_med_vol_diam, 0._wp)
    call alloc3(cfg%graupel_med_vol_diam_flag, dg%graupel_med_vol_diam, 0._wp)
    call alloc3(cfg%cloud_number_mixing_ratio_flag, dg%cloud_number_mixing_ratio, 0._wp)
  contains
    subroutine alloc2(on, a)
      logical, intent(in) :: on
      real(wp), allocatable, intent(inout) :: a(:,:)
      if (.not. on) return
      if (allocated(a)) then
        if (lbound(a, 1) /= its .or. ubound(a, 1) /= ite .or. lbound(a, 2) /= jts .or. ubound(a, 2) /= jte) &
          deallocate(a)
      endif
      if (.not. allocated(a)) allocate(a(its:ite, jts:jte))
      a = 0._wp
    end subroutine
    subroutine alloc3(on, a, v)
      logical, intent(in) :: on
      real(wp), allocatable, intent(inout) :: a(:,:,:)
      real(wp), intent(in) :: v
      if (.not. on) return
      if (allocated(a)) then
        if (lbound(a, 1) /= its .or. ubound(a, 1) /= ite .or. lbound(a, 2) /= kts .or. ubound(a, 2) /= kte .or. &
            lbound(a, 3) /= jts .or. ubound(a, 3) /= jte) deallocate(a)
      endif
      if (.not. allocated(a)) allocate(a(its:ite, kts:kte, jts:jte))
      a = v
    end subroutine
  end subroutine

end module tempo_hip_mod
