"""Restart files of the time stepper.

The reference writes restart_<cdate>.nc holding 29 fields and the dates
cdate / pdate (mod_io.f90:59-124, written by mod_mympi.f90:133-176 and read
back by mod_mympi.f90:460-547); a restart run continues from cdate with
tctr%ist = (cdate - bdate)/dtbc + 1 and without the AB3 bootstrap
(main.f90:128-141, mod_int.f90:122).  netCDF is not part of this build, so
the same fields are kept in a raw little-endian stream of the interior
(glo_ni, glo_nj[, nk]) arrays, which the Fortran driver reads and writes
with access='stream':

    int32  magic 0x50434f4d ("PCOM"), version 1
    int32  glo_ni, glo_nj, nk
    int64  step index of the last completed step (tctr%i)
    int32  cdate[6], pdate[6]          (y, m, d, h, mi, s)
    float64 fields in RESTART_FIELDS order, Fortran order, no halo

Halo columns are rebuilt on load by the reference's one-rank swap (column
1 = column ni-1, column ni = column 2); wall rows 1 and nj are kept from
the halo'd state the fields had, since the reference does not write them
either (merge_out_* gathers var(2:ni-1, 2:nj-1, :)).
"""
import struct

import numpy as np

MAGIC = 0x50434F4D
VERSION = 1

# mod_io.f90:80-113, in the reference's definition order
RESTART_FIELDS = [
    "chc", "chp",
    "bnd_taux", "bnd_tauy", "bnd_t", "bnd_s", "bnd_pa", "bnd_fw",
    "buc", "bvc", "bup", "bvp",
    "tc", "sc", "tp", "sp",
    "prho_t", "prho_s",
    "uc", "vc", "up", "vp",
    "auc", "avc", "aup", "avp", "aupp", "avpp",
    "am",
]

_HEAD = struct.Struct("<iiiiiq6i6i")


def write(path, model, cdate, pdate, step):
    """Write the restart file of `model` (a pcom_amd.lib.Model)."""
    g_ni, g_nj = model.ni - 2, model.nj - 2
    with open(path, "wb") as f:
        f.write(_HEAD.pack(MAGIC, VERSION, g_ni, g_nj, model.nk, int(step), *cdate, *pdate))
        for name in RESTART_FIELDS:
            a = model.get(name)
            inner = a[1:-1, 1:-1] if a.ndim == 2 else a[1:-1, 1:-1, :]
            f.write(np.asfortranarray(inner, dtype="<f8").tobytes(order="F"))


def read(path):
    """Return (header dict, {field: interior array})."""
    with open(path, "rb") as f:
        head = _HEAD.unpack(f.read(_HEAD.size))
        magic, ver, g_ni, g_nj, nk, step = head[:6]
        if magic != MAGIC or ver != VERSION:
            raise ValueError(f"{path}: not a PCOM restart file of version {VERSION}")
        cdate, pdate = tuple(head[6:12]), tuple(head[12:18])
        fields = {}
        for name in RESTART_FIELDS:
            shp = (g_ni, g_nj, nk) if name in _NK_FIELDS else (g_ni, g_nj)
            n = int(np.prod(shp))
            fields[name] = np.frombuffer(f.read(8 * n), dtype="<f8").reshape(shp, order="F").copy()
    return {"glo_ni": g_ni, "glo_nj": g_nj, "nk": nk, "step": step, "cdate": cdate, "pdate": pdate}, fields


_NK_FIELDS = {"tc", "sc", "tp", "sp", "prho_t", "prho_s", "uc", "vc", "up", "vp",
              "auc", "avc", "aup", "avp", "aupp", "avpp", "am"}


def load(path, model):
    """Load a restart file into a context created with restart_run = 1 and
    set its clock; fields keep their wall rows and get periodic halos."""
    head, fields = read(path)
    if (head["glo_ni"], head["glo_nj"], head["nk"]) != (model.ni - 2, model.nj - 2, model.nk):
        raise ValueError(f"{path}: grid {head['glo_ni']}x{head['glo_nj']}x{head['nk']} does not match the context")
    for name in RESTART_FIELDS:
        a = model.get(name)
        if a.ndim == 2:
            a[1:-1, 1:-1] = fields[name]
        else:
            a[1:-1, 1:-1, :] = fields[name]
        a[0] = a[-2]
        a[-1] = a[1]
        model.set(name, a)
    model.set_clock(head["step"], head["cdate"], head["pdate"])
    return head
