"""The LDS staging of the wavefront sedimentation kernel (SedTile in
tempo_amd/csrc/tempo_sedi.hpp), replayed on the CPU: a block of
64 x kSedCols threads reads a column group's run of the column-tiled scratch
(element e = level * kSedCols + column) in kSedChunks passes, clamping
indices past the run's end to its last element, and writes each element to
its (column, level) row of LDS; each wavefront then reads its own row, and
the stores write the rows back.  For every level count the kernel takes
(2..128) and for column groups cut short at the batch's end, with and
without the block-uniform test that skips passes past the run (the kernel's
form and its check form, sed_wave 2), every LDS cell is written exactly once
with its own element, no load leaves the group's scratch, and the round trip
returns the run unchanged."""
import numpy as np

K_SED_COLS, K_SED_CHUNKS, K_SCR_LEVELS = 8, 2, 16


def replay(nz, cols_in_batch, every_pass):
    nzs = (nz + K_SCR_LEVELS - 1) // K_SCR_LEVELS * K_SCR_LEVELS
    row, n, threads = nz + 1, K_SED_COLS * nz, 64 * K_SED_COLS
    run = np.arange(K_SED_COLS * nzs, dtype=np.float64) + 1000.0
    lds = np.full(K_SED_COLS * row, np.nan)
    writes = np.zeros(K_SED_COLS * row, int)
    for tid in range(threads):
        v = {}
        for it in range(K_SED_CHUNKS):
            if not every_pass and it * threads >= n:
                break
            e = min(tid + it * threads, n - 1)
            assert 0 <= e < K_SED_COLS * nzs
            v[it] = run[e]
        for it in range(K_SED_CHUNKS):
            e = tid + it * threads
            if e < n:
                cc, k = e % K_SED_COLS, e // K_SED_COLS
                lds[cc * row + k] = v[it]
                writes[cc * row + k] += 1
    for cc in range(K_SED_COLS):
        for k in range(nz):
            assert writes[cc * row + k] == 1 and lds[cc * row + k] == run[k * K_SED_COLS + cc], (nz, cc, k)
    assert writes.sum() == n
    # each wavefront's row read level by level (lane k % 64 of chunk k // 64)
    for w in range(K_SED_COLS):
        got = [lds[w * row + j * 64 + lane] for j in range(K_SED_CHUNKS) for lane in range(64)
               if j * 64 + lane < nz]
        assert got == [run[k * K_SED_COLS + w] for k in range(nz)]
    out = run.copy()
    out[:n] = -1.0
    for tid in range(threads):
        for e in range(tid, n, threads):
            cc, k = e % K_SED_COLS, e // K_SED_COLS
            out[e] = lds[cc * row + k] if cc < cols_in_batch else run[e]
    assert np.array_equal(out, run)


def test_staging_is_a_bijection_in_both_forms():
    for nz in range(2, 129):
        for cols in (1, 5, K_SED_COLS):
            replay(nz, cols, False)
            replay(nz, cols, True)
