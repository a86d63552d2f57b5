"""ctypes mirror of include/pcom_gpu.h (the C ABI of the MI355X library).

The structs and field ids here must match the header exactly; tests check
that every symbol the header declares is exported by the built library.
"""
import ctypes as C

PCOM_OK = 0
PCOM_ERR_ARG = -1
PCOM_ERR_HIP = -2
PCOM_ERR_NAN = -3
PCOM_ERR_STATE = -4
PCOM_ERR_NOGPU = -5

PCOM_FP64 = 0
PCOM_MIXED = 1
PCOM_MIXED_EOS32 = 2

PCOM_TROP_AUTO = 0
PCOM_TROP_PERSISTENT = 1
PCOM_TROP_LAUNCHES = 2
FORMS = {"fp64": PCOM_FP64, "mixed": PCOM_MIXED, "mixed_eos32": PCOM_MIXED_EOS32}

_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)


class PcomInputs(C.Structure):
    _fields_ = [
        ("glo_ni", C.c_int), ("glo_nj", C.c_int), ("nk", C.c_int),
        ("lon", _DP), ("lat", _DP), ("z", _DP), ("itn", _IP),
        ("tmn", _DP), ("smn", _DP), ("pt", _DP), ("sa", _DP), ("vmix", _DP),
        ("taux", _DP), ("tauy", _DP), ("bct", _DP), ("bcs", _DP),
        ("pa", _DP), ("fw", _DP),
    ]


class PcomOpts(C.Structure):
    _fields_ = [
        ("dtbt", C.c_int), ("dtbc", C.c_int),
        ("year", C.c_int), ("month", C.c_int), ("day", C.c_int), ("hour", C.c_int),
        ("precision", C.c_int), ("restart_run", C.c_int), ("device", C.c_int),
        ("trop_tile_i", C.c_int), ("trop_tile_j", C.c_int), ("trop_substeps", C.c_int),
        ("trop_mode", C.c_int),
    ]


class PcomCounters(C.Structure):
    _fields_ = [
        ("step", C.c_longlong),
        ("n_ts", C.c_int), ("n_uv", C.c_int), ("n_w", C.c_int),
        ("n_ch", C.c_int), ("n_ssh", C.c_int),
        ("year", C.c_int), ("month", C.c_int), ("day", C.c_int),
        ("hour", C.c_int), ("minute", C.c_int), ("second", C.c_int),
    ]


class PcomForcing(C.Structure):
    _fields_ = [("taux", _DP), ("tauy", _DP), ("pa", _DP), ("bct", _DP)]


def forcing(taux=None, tauy=None, pa=None, bct=None):
    """A PcomForcing of (glo_ni, glo_nj) records (None = keep int_bnd's
    value); the arrays are kept alive on the returned struct."""
    f = PcomForcing()
    keep = []
    for name, a in (("taux", taux), ("tauy", tauy), ("pa", pa), ("bct", bct)):
        if a is not None:
            a = __import__("numpy").asfortranarray(a, dtype="float64")
            keep.append(a)
            setattr(f, name, a.ctypes.data_as(_DP))
    f._keep = keep
    return f


PCOM_MISSING_FLOAT = float(__import__("numpy").float32(-9.99999e30))

MEAN_NAMES = ["pt", "sa", "rho", "u", "v", "w", "ch", "ssh"]


class PcomMeans(C.Structure):
    _fields_ = [(n, _DP) for n in MEAN_NAMES]


class PcomNnStats(C.Structure):
    _fields_ = [
        ("recorded", C.c_longlong), ("surrogate_steps", C.c_longlong), ("plain_steps", C.c_longlong),
        ("checks", C.c_longlong), ("fallbacks", C.c_longlong), ("in_use", C.c_int),
        ("last_check_rmse", C.c_double * 4), ("max_check_rmse", C.c_double * 4),
        ("last_persist_rmse", C.c_double * 4),
    ]


# Field ids, in header order.
FIELD_NAMES = [
    "tc", "sc", "tp", "sp",
    "uc", "vc", "up", "vp",
    "auc", "avc", "aup", "avp", "aupp", "avpp",
    "fx", "fy", "am",
    "alpha", "adp",
    "wm",
    "prho_t", "prho_s",
    "act", "acs", "acr", "acu", "acv", "acw",
    "chc", "chp", "mch", "mch2",
    "acch", "acssh",
    "buc", "bvc", "bup", "bvp", "but", "bvt",
    "pax", "pay",
    "bphi_xn", "bphi_xs", "bphi_ye", "bphi_yw",
    "bgra_xn", "bgra_xs", "bgra_ye", "bgra_yw",
    "bnd_taux", "bnd_tauy", "bnd_t", "bnd_s", "bnd_pa", "bnd_fw",
    "ssh",
    "graphihx", "graphihy", "cor",
    "phih_t", "prh_t", "phih_u", "prh_u",
    "km", "kh",
]
FIELD = {n: i for i, n in enumerate(FIELD_NAMES)}
FIELD_COUNT = len(FIELD_NAMES)

# Fields with nk+1 levels and fields with nk levels; the rest are 2-D.
LEVELS_NKP = {"wm", "km", "kh"}
LEVELS_NK = {
    "tc", "sc", "tp", "sp", "uc", "vc", "up", "vp", "auc", "avc", "aup", "avp",
    "aupp", "avpp", "fx", "fy", "am", "alpha", "adp", "prho_t", "prho_s",
    "act", "acs", "acr", "acu", "acv", "acw",
}


def field_levels(name, nk):
    if name in LEVELS_NKP:
        return nk + 1
    if name in LEVELS_NK:
        return nk
    return 1


def dptr(a):
    """Pointer to a C-contiguous float64 numpy array, or NULL for None."""
    if a is None:
        return _DP()
    return a.ctypes.data_as(_DP)


def iptr(a):
    return a.ctypes.data_as(_IP)
