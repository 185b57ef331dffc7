! Test-infrastructure driver for the reference table builder.
!
! Runs tempo_build_tables (src/module_mp_tempo_tables.F90:22-78) for one
! (rank, nproc) slice of the rain dimension ntb_r, exactly as the reference's
! MPI build splits it (get_index_for_rank, src/module_mp_tempo_tables.F90:464-490),
! but as independent processes: each process then dumps the module arrays it
! computed (native endian, stream) so that oracle/ref_tables.py can assemble
! the full tables in the reference file format.
!
! usage: ref_tables <rank> <nproc> <outdir>
program ref_tables
  use iso_c_binding
  use module_mp_tempo_tables, only : tempo_build_tables
  use module_mp_tempo_cfgs, only : ty_tempo_cfgs
  use module_mp_tempo_params
  implicit none
  type(ty_tempo_cfgs) :: cfgs
  integer :: rank, nproc, u
  character(len=256) :: arg, outdir

  call get_command_argument(1, arg); read(arg, *) rank
  call get_command_argument(2, arg); read(arg, *) nproc
  call get_command_argument(3, outdir)

  call tempo_build_tables(rank, nproc, cfgs)

  open(newunit=u, file=trim(outdir)//'/qs_slice_'//trim(itoa(rank))//'.bin', &
    form='unformatted', access='stream', status='replace')
  write(u) tcs_racs1, tmr_racs1, tcs_racs2, tmr_racs2, tcr_sacr1, tms_sacr1, &
    tcr_sacr2, tms_sacr2, tnr_racs1, tnr_racs2, tnr_sacr1, tnr_sacr2
  close(u)
  open(newunit=u, file=trim(outdir)//'/qg_slice_'//trim(itoa(rank))//'.bin', &
    form='unformatted', access='stream', status='replace')
  write(u) tcg_racg, tmr_racg, tcr_gacr, tnr_racg, tnr_gacr
  close(u)
  if (rank == 0) then
    open(newunit=u, file=trim(outdir)//'/fz.bin', form='unformatted', access='stream', &
      status='replace')
    write(u) tpi_qrfz, tni_qrfz, tpg_qrfz, tnr_qrfz, tpi_qcfz, tni_qcfz
    close(u)
  endif
  print '(A,I3,A,I3,A,I8)', 'rank ', rank, ' of ', nproc, ' done; qg slice entries ', size(tcg_racg)

contains

  function itoa(i) result(s)
    integer, intent(in) :: i
    character(len=12) :: s
    write(s, '(I0)') i
  end function itoa

end program ref_tables
