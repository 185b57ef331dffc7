"""Lookup-table files in the reference formats.

All tables are Fortran column-major (first index fastest) and big-endian:

  ccn_activate.bin             sequential unformatted, one record with 4-byte
                               length markers: tnccn_act(7,9,7,5,4) real32
                               (src/module_mp_tempo_driver.F90:625-649)
  freeze_water_data_tempo_v3   stream: tpi_qrfz, tni_qrfz, tpg_qrfz, tnr_qrfz
                               (37,37,45,55), tpi_qcfz, tni_qcfz (37,100,45,55)
                               real64 (src/module_mp_tempo_driver.F90:532-558)
  qr_acr_qs_data_tempo_v3      stream: 12 arrays (37,9,37,37) real64
                               (src/module_mp_tempo_driver.F90:561-594)
  qr_acr_qg_data_tempo_v3      stream: 5 arrays (37,37,9,37,37) real64
                               (src/module_mp_tempo_driver.F90:597-622)

In memory a table set is kept as one flat float64 vector per file, in file
order, native endian -- the layout the C ABI and the device copies use.
"""
import os

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")

N_CCN = 7 * 9 * 7 * 5 * 4
N_FZ_R = 37 * 37 * 45 * 55
N_FZ_C = 37 * 100 * 45 * 55
N_FZ = 4 * N_FZ_R + 2 * N_FZ_C
N_QS1 = 37 * 9 * 37 * 37
N_QS = 12 * N_QS1
N_QG1 = 37 * 37 * 9 * 37 * 37
N_QG = 5 * N_QG1

FILE_NAMES = dict(ccn="ccn_activate.bin", fz="freeze_water_data_tempo_v3",
                  qs="qr_acr_qs_data_tempo_v3", qg="qr_acr_qg_data_tempo_v3")
SIZES = dict(fz=N_FZ, qs=N_QS, qg=N_QG)
QS_NAMES = ("tcs_racs1", "tmr_racs1", "tcs_racs2", "tmr_racs2", "tcr_sacr1", "tms_sacr1",
            "tcr_sacr2", "tms_sacr2", "tnr_racs1", "tnr_racs2", "tnr_sacr1", "tnr_sacr2")
QG_NAMES = ("tcg_racg", "tmr_racg", "tcr_gacr", "tnr_racg", "tnr_gacr")
FZ_NAMES = ("tpi_qrfz", "tni_qrfz", "tpg_qrfz", "tnr_qrfz", "tpi_qcfz", "tni_qcfz")


class TableFileError(RuntimeError):
    pass


def read_ccn(path=None):
    """tnccn_act as a flat float32 vector (Fortran order)."""
    path = path or os.path.join(DATA_DIR, FILE_NAMES["ccn"])
    raw = open(path, "rb").read()
    if len(raw) != 4 * N_CCN + 8:
        raise TableFileError(f"{path}: {len(raw)} bytes, expected {4 * N_CCN + 8}")
    head, tail = int.from_bytes(raw[:4], "big"), int.from_bytes(raw[-4:], "big")
    if head != 4 * N_CCN or tail != 4 * N_CCN:
        raise TableFileError(f"{path}: bad record markers {head}/{tail}")
    return np.frombuffer(raw[4:-4], dtype=">f4").astype(np.float32)


def read_stream(path, kind):
    """A big-endian stream table as a flat native float64 vector."""
    n = SIZES[kind]
    size = os.path.getsize(path)
    if size != 8 * n:
        raise TableFileError(f"{path}: {size} bytes, expected {8 * n}")
    return np.fromfile(path, dtype=">f8").astype(np.float64)


def write_stream(path, flat):
    np.asarray(flat, dtype=np.float64).astype(">f8").tofile(path)


def read_dir(path):
    """Read the three stream tables (and the CCN table) from a directory."""
    out = dict(ccn=read_ccn(os.path.join(path, FILE_NAMES["ccn"]))
               if os.path.exists(os.path.join(path, FILE_NAMES["ccn"])) else read_ccn())
    for kind in ("fz", "qs", "qg"):
        out[kind] = read_stream(os.path.join(path, FILE_NAMES[kind]), kind)
    return out


def write_dir(path, tables):
    os.makedirs(path, exist_ok=True)
    for kind in ("fz", "qs", "qg"):
        write_stream(os.path.join(path, FILE_NAMES[kind]), tables[kind])


def view(tables, kind):
    """Named Fortran-shaped views of a flat table vector."""
    flat = tables[kind]
    if kind == "qs":
        return {n: flat[i * N_QS1:(i + 1) * N_QS1].reshape((37, 37, 9, 37)).transpose(3, 2, 1, 0)
                for i, n in enumerate(QS_NAMES)}
    if kind == "qg":
        return {n: flat[i * N_QG1:(i + 1) * N_QG1].reshape((37, 37, 9, 37, 37)).transpose(4, 3, 2, 1, 0)
                for i, n in enumerate(QG_NAMES)}
    if kind == "fz":
        out = {}
        for i, n in enumerate(FZ_NAMES[:4]):
            out[n] = flat[i * N_FZ_R:(i + 1) * N_FZ_R].reshape((55, 45, 37, 37)).transpose(3, 2, 1, 0)
        base = 4 * N_FZ_R
        for i, n in enumerate(FZ_NAMES[4:]):
            out[n] = flat[base + i * N_FZ_C:base + (i + 1) * N_FZ_C].reshape(
                (55, 45, 100, 37)).transpose(3, 2, 1, 0)
        return out
    raise KeyError(kind)
