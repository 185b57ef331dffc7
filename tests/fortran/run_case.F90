! Drives the Fortran binding (tempo_amd/fortran/tempo_hip_mod.F90) the way a
! host model drives the reference: tempo_init, then tempo_run on a tile of
! arrays with a one-point halo in i and k and two j rows.
!
!   run_case <input.bin> <output.bin> <ccn_activate.bin> [locked]
!
! With "locked" the state arrays live in page-locked host memory
! (tempo_hip_host_alloc), which tempo_run moves without staging copies.
!
! input.bin (native endian): int32 ni, nj, nz, itimestep; float64 dt; then
! 17 float64 fields t p w dz qv qc qr qi qs qg ni nr nc nwfa nifa ng qb, each
! ni*nz*nj values in Fortran order (i, k, j); then int32 land(ni, nj).
! output.bin: the 14 updated fields (t qv qc qr qi qs qg ni nr nc nwfa nifa
! ng qb), the 8 two-dimensional and the 4 default three-dimensional
! diagnostics over the tile, then int32 1 if every halo point kept its
! sentinel value and the diagnostics have the tile's bounds, else 0.
program run_case
  use iso_c_binding
  use tempo_hip_mod
  implicit none
  integer, parameter :: nf = 17
  integer(c_int) :: hdr(4)
  real(c_double) :: dt
  integer :: ni, nj, nz, its, f, u, ios
  real(c_double), pointer :: st(:,:,:,:), st_flat(:,:,:,:), probe(:,:,:)
  type(c_ptr) :: blk
  character(len=16) :: mode
  logical :: locked
  integer, allocatable :: land(:,:)
  real(c_double), allocatable :: buf(:,:,:)
  type(ty_tempo_cfgs) :: cfg
  type(ty_tempo_driver_diags) :: dg
  character(len=512) :: fin, fout, fccn
  real(c_double), parameter :: sentinel = -7.25_c_double
  logical :: halo_ok
  integer, parameter :: out_fields(14) = [1, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17]

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, fccn)
  mode = ''
  if (command_argument_count() >= 4) call get_command_argument(4, mode)
  locked = trim(mode) == 'locked'
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', iostat=ios)
  if (ios /= 0) error stop 2
  read(u) hdr, dt
  ni = hdr(1); nj = hdr(2); nz = hdr(3)
  if (locked) then
    if (tempo_hip_host_alloc(int(ni + 2, c_size_t) * (nz + 2) * nj * nf * 8, blk) /= 0) error stop 3
    call c_f_pointer(blk, st_flat, [ni + 2, nz + 2, nj, nf])
    st(0:, 0:, 1:, 1:) => st_flat
    ! the 3-D helper: bounds as asked, usable, released
    call tempo_alloc_page_locked(probe, [0, 0, 1], [ni + 1, nz + 1, nj])
    if (any(lbound(probe) /= [0, 0, 1]) .or. any(ubound(probe) /= [ni + 1, nz + 1, nj])) error stop 4
    probe = 1.0_c_double
    if (sum(probe) /= real((ni + 2) * (nz + 2) * nj, c_double)) error stop 5
    call tempo_free_page_locked(probe)
    if (associated(probe)) error stop 6
  else
    allocate(st(0:ni + 1, 0:nz + 1, 1:nj, nf))
  end if
  allocate(buf(ni, nz, nj), land(0:ni + 1, 1:nj))
  st = sentinel
  do f = 1, nf
    read(u) buf
    st(1:ni, 1:nz, 1:nj, f) = buf
  end do
  land = 0
  read(u) land(1:ni, 1:nj)
  close(u)

  cfg%ccn_path = trim(fccn)
  call tempo_init(aerosolaware_flag=.true., hailaware_flag=.true., tempo_cfgs=cfg)
  ! diagnostics the host allocated over its memory extent: tempo_run gives
  ! them the tile's bounds, as the reference allocates them
  allocate(dg%rain_precip(0:ni + 1, 1:nj), dg%refl10cm(0:ni + 1, 0:nz + 1, 1:nj))
  dg%rain_precip = sentinel
  dg%refl10cm = sentinel
  its = 1
  call tempo_run(cfg, dt, int(hdr(4)), &
    t=st(:,:,:,1), p=st(:,:,:,2), w=st(:,:,:,3), dz=st(:,:,:,4), &
    qv=st(:,:,:,5), qc=st(:,:,:,6), qr=st(:,:,:,7), qi=st(:,:,:,8), qs=st(:,:,:,9), qg=st(:,:,:,10), &
    ni=st(:,:,:,11), nr=st(:,:,:,12), nc=st(:,:,:,13), nwfa=st(:,:,:,14), nifa=st(:,:,:,15), &
    ng=st(:,:,:,16), qb=st(:,:,:,17), land_input=land, &
    ids=1, ide=ni, jds=1, jde=nj, kds=1, kde=nz, &
    ims=0, ime=ni + 1, jms=1, jme=nj, kms=0, kme=nz + 1, &
    its=its, ite=ni, jts=1, jte=nj, kts=1, kte=nz, tempo_diags=dg)

  halo_ok = all(st(0, :, :, :) == sentinel) .and. all(st(ni + 1, :, :, :) == sentinel) .and. &
    all(st(:, 0, :, :) == sentinel) .and. all(st(:, nz + 1, :, :) == sentinel)
  ! tile-shaped diagnostics
  halo_ok = halo_ok .and. all(lbound(dg%rain_precip) == [1, 1]) .and. all(ubound(dg%rain_precip) == [ni, nj]) .and. &
    all(lbound(dg%refl10cm) == [1, 1, 1]) .and. all(ubound(dg%refl10cm) == [ni, nz, nj]) .and. &
    all(lbound(dg%re_ice) == [1, 1, 1]) .and. all(ubound(dg%re_ice) == [ni, nz, nj])
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  do f = 1, size(out_fields)
    write(u) st(1:ni, 1:nz, 1:nj, out_fields(f))
  end do
  write(u) dg%rain_precip(1:ni, :), dg%ice_liquid_equiv_precip(1:ni, :), dg%snow_liquid_equiv_precip(1:ni, :), &
    dg%graupel_liquid_equiv_precip(1:ni, :), dg%frozen_fraction(1:ni, :), dg%frz_rain_precip(1:ni, :), &
    dg%max_hail_diameter_sfc(1:ni, :), dg%max_hail_diameter_column(1:ni, :)
  write(u) dg%refl10cm(1:ni, 1:nz, :), dg%re_cloud(1:ni, 1:nz, :), dg%re_ice(1:ni, 1:nz, :), dg%re_snow(1:ni, 1:nz, :)
  write(u) merge(1_c_int, 0_c_int, halo_ok)
  close(u)
  call tempo_finalize(cfg)
  if (locked) then
    if (tempo_hip_host_free(blk) /= 0) error stop 7
  else
    deallocate(st)
  end if
end program run_case
