"""Build the reference lookup tables with the reference's own builder.

Test infrastructure.  Runs oracle/_ref/ref_tables (tempo_build_tables,
src/module_mp_tempo_tables.F90:22-78, compiled from the reference sources by
``make -C oracle ref``) as NPROC independent processes, one per slice of the
rain dimension ntb_r exactly as the reference's MPI build splits it
(get_index_for_rank, :464-490), then concatenates the slices into the
reference file format (big-endian stream, src/module_mp_tempo_tables.F90:96-444).

    python -m oracle.ref_tables OUTDIR [NPROC]

Takes about 8 minutes on 8 cores; the output (975 MB) stays outside git.
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NQS_ARR, NQG_ARR = 12, 5


def slice_bounds(idx, rank, nproc):
    """get_index_for_rank (src/module_mp_tempo_tables.F90:464-490), 1-based."""
    per = idx // nproc
    rem = idx % nproc
    if rank < rem:
        per += 1
        return rank * per + 1, (rank + 1) * per
    return rem + rank * per + 1, rem + (rank + 1) * per


def run_slices(rawdir, nproc):
    exe = os.path.join(HERE, "_ref", "ref_tables")
    procs = []
    for r in range(nproc):
        wd = os.path.join(rawdir, f"rank{r}")
        os.makedirs(wd, exist_ok=True)
        procs.append(subprocess.Popen([exe, str(r), str(nproc), rawdir], cwd=wd,
                                      stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(out.decode())


def assemble(rawdir, outdir, nproc):
    from tempo_amd import tables_io as tio
    os.makedirs(outdir, exist_ok=True)
    for kind, narr, per_r in (("qs", NQS_ARR, 37 * 9 * 37), ("qg", NQG_ARR, 37 * 37 * 9 * 37)):
        full = [[] for _ in range(narr)]
        for r in range(nproc):
            lo, hi = slice_bounds(37, r, nproc)
            n = per_r * (hi - lo + 1)
            raw = np.fromfile(os.path.join(rawdir, f"{kind}_slice_{r}.bin"), dtype=np.float64)
            assert raw.size == narr * n, (kind, r, raw.size, narr * n)
            for a in range(narr):
                full[a].append(raw[a * n:(a + 1) * n])
        flat = np.concatenate([np.concatenate(parts) for parts in full])
        tio.write_stream(os.path.join(outdir, tio.FILE_NAMES[kind]), flat)
    fz = np.fromfile(os.path.join(rawdir, "fz.bin"), dtype=np.float64)
    assert fz.size == tio.N_FZ
    tio.write_stream(os.path.join(outdir, tio.FILE_NAMES["fz"]), fz)


def main():
    outdir = sys.argv[1]
    nproc = int(sys.argv[2]) if len(sys.argv) > 2 else 8
    rawdir = os.path.join(outdir, "raw")
    if not os.path.exists(os.path.join(rawdir, "fz.bin")):
        run_slices(rawdir, nproc)
    assemble(rawdir, outdir, nproc)


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(HERE))
    main()
