"""ctypes wrapper of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never as the thing
measured or shipped.
"""
import ctypes as C
import os

import numpy as np

from pcom_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(path)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(abi.PcomInputs), C.POINTER(abi.PcomOpts)]
        L.orc_destroy.argtypes = [vp]
        L.orc_step.argtypes = [vp, C.c_int]
        for n in ("orc_get", "orc_set"):
            getattr(L, n).argtypes = [vp, C.c_int, dp, C.c_size_t]
            getattr(L, n).restype = C.c_int
        L.orc_field_size.argtypes = [vp, C.c_int]
        L.orc_field_size.restype = C.c_size_t
        L.orc_dims.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_counters.argtypes = [vp, C.POINTER(abi.PcomCounters)]
        for n in ("orc_den_prho", "orc_den_alpha_main", "orc_int_pgra", "orc_int_readyc",
                  "orc_int_trop", "orc_int_clin", "orc_int_ssh", "orc_int_ts", "orc_int_bnd",
                  "orc_advance_time"):
            getattr(L, n).argtypes = [vp]
        L.orc_set_step_index.argtypes = [vp, C.c_longlong]
        L.orc_set_clock.argtypes = [vp, C.c_longlong, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_output_means.argtypes = [vp, C.POINTER(abi.PcomMeans)]
        L.orc_set_forcing.argtypes = [vp, C.POINTER(abi.PcomForcing), C.c_int]
        L.orc_set_forcing.restype = C.c_int
        L.orc_forcing_due.argtypes = [vp, C.POINTER(C.c_int)]
        L.orc_output_means.restype = C.c_int
        L.orc_op_ter.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double, dp, dp]
        L.orc_op_px1.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp]
        L.orc_op_px2.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp]
        L.orc_op_div.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, dp]
        L.orc_op_lap.argtypes = [vp, C.c_int, C.c_int, dp, dp]
        L.orc_op_adv.argtypes = [vp, dp, dp, dp]
        L.orc_op_adv_ts.argtypes = [vp, dp, dp, dp]
        L.orc_op_dif.argtypes = [vp, dp, dp, dp]
        L.orc_op_fri.argtypes = [vp, dp]
        L.orc_upwelling.argtypes = [vp]
        L.orc_smag.argtypes = [vp, dp, dp]
        L.orc_metric.argtypes = [vp, C.c_int, C.c_int, dp, C.c_size_t]
        L.orc_metric.restype = C.c_int
        L.orc_vgrid.argtypes = [vp, C.c_int, C.c_int, dp, C.c_size_t]
        L.orc_vgrid.restype = C.c_int
        L.orc_igrid.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.c_size_t]
        L.orc_igrid.restype = C.c_int
        L.orc_den_rho.restype = C.c_double
        L.orc_den_rho.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_den_rho_n.argtypes = [C.c_size_t, dp, dp, dp, dp]
        L.orc_constants.restype = C.c_int
        L.orc_constants.argtypes = [C.POINTER(C.POINTER(C.c_char_p)), C.POINTER(dp)]
        L.orc_convect_column.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, dp]
        L.orc_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def threads(n=None):
    """Host threads the oracle's loops run on (OpenMP); sets them if n."""
    if n:
        lib().orc_set_threads(int(n))
    return lib().orc_threads()


def den_rho(t, s, p):
    t, s, p = np.broadcast_arrays(np.asarray(t, np.float64), np.asarray(s, np.float64),
                                  np.asarray(p, np.float64))
    t, s, p = (np.ascontiguousarray(x) for x in (t, s, p))
    out = np.empty_like(t)
    lib().orc_den_rho_n(t.size, _dp(t), _dp(s), _dp(p), _dp(out))
    return out


def constants():
    names = C.POINTER(C.c_char_p)()
    vals = C.POINTER(C.c_double)()
    n = lib().orc_constants(C.byref(names), C.byref(vals))
    return {names[i].decode(): vals[i] for i in range(n)}


def convect_column(nlev, dpr, prho_t, prho_s, t, s):
    nk = len(dpr)
    t = np.array(t, np.float64)
    s = np.array(s, np.float64)
    dpr, prho_t, prho_s = (np.ascontiguousarray(x, np.float64) for x in (dpr, prho_t, prho_s))
    lib().orc_convect_column(nk, nlev, _dp(dpr), _dp(prho_t), _dp(prho_s), _dp(t), _dp(s))
    return t, s


class Oracle:
    """The FP64 CPU restatement on one configuration."""

    def __init__(self, cfg, restart_run=0):
        self.cfg = cfg
        self._inp = cfg.c_inputs()
        self._opt = cfg.c_opts(precision=abi.PCOM_FP64, restart_run=restart_run)
        self.h = lib().orc_create(C.byref(self._inp), C.byref(self._opt))
        if not self.h:
            raise ValueError("orc_create rejected the configuration")
        ni, nj, nk = C.c_int(), C.c_int(), C.c_int()
        lib().orc_dims(self.h, C.byref(ni), C.byref(nj), C.byref(nk))
        self.ni, self.nj, self.nk = ni.value, nj.value, nk.value

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _LIB is not None:
            _LIB.orc_destroy(h)
            self.h = None

    def step(self, n=1):
        lib().orc_step(self.h, n)

    def shape(self, name):
        nd3 = abi.field_levels(name, self.nk)
        return (self.ni, self.nj) if nd3 == 1 else (self.ni, self.nj, nd3)

    def get(self, name):
        shp = self.shape(name)
        out = np.empty(shp, dtype=np.float64, order="F")
        rc = lib().orc_get(self.h, abi.FIELD[name], _dp(out), out.size)
        if rc != 0:
            raise KeyError(name)
        return out

    def set(self, name, arr):
        a = np.asfortranarray(arr, dtype=np.float64)
        if a.shape != self.shape(name):
            raise ValueError(f"{name}: shape {a.shape} != {self.shape(name)}")
        rc = lib().orc_set(self.h, abi.FIELD[name], _dp(a), a.size)
        if rc != 0:
            raise KeyError(name)

    def counters(self):
        c = abi.PcomCounters()
        lib().orc_counters(self.h, C.byref(c))
        return c

    def run_stage(self, name):
        getattr(lib(), "orc_" + name)(self.h)

    def set_step_index(self, i):
        lib().orc_set_step_index(self.h, i)

    def set_clock(self, step, cdate, pdate):
        c = (C.c_int * 6)(*cdate)
        p = (C.c_int * 6)(*pdate)
        lib().orc_set_clock(self.h, C.c_longlong(step), c, p)

    def forcing_due(self):
        d = (C.c_int * 2)()
        lib().orc_forcing_due(self.h, d)
        return bool(d[0]), bool(d[1])

    def set_forcing(self, interpolate, taux=None, tauy=None, pa=None, bct=None):
        f = abi.forcing(taux, tauy, pa, bct)
        if lib().orc_set_forcing(self.h, C.byref(f), int(bool(interpolate))) != 0:
            raise ValueError("orc_set_forcing: the daily refresh of the next step comes first")

    def output_means(self, names=abi.MEAN_NAMES):
        g = (self.ni - 2, self.nj - 2)
        out, m = {}, abi.PcomMeans()
        for n in names:
            shp = g if n in ("ch", "ssh") else g + (self.nk,)
            out[n] = np.empty(shp, order="F")
            setattr(m, n, _dp(out[n]))
        if lib().orc_output_means(self.h, C.byref(m)) != 0:
            raise ValueError("orc_output_means: a record count is zero")
        return out

    def metric(self, g, which):
        out = np.empty((self.ni, self.nj), order="F")
        assert lib().orc_metric(self.h, g, which, _dp(out), out.size) == 0
        return out

    def vgrid(self, gi, which):
        n = self.nk + 1 if gi == 1 else self.nk
        out = np.empty(n)
        assert lib().orc_vgrid(self.h, gi, which, _dp(out), n) == 0
        return out

    def igrid(self, which):
        shp = (self.ni, self.nj, self.nk) if which < 2 else (self.ni, self.nj)
        out = np.empty(shp, dtype=np.int32, order="F")
        assert lib().orc_igrid(self.h, which, out.ctypes.data_as(C.POINTER(C.c_int)), out.size) == 0
        return out

    def _out(self, nd3):
        return np.zeros((self.ni, self.nj, nd3) if nd3 > 1 else (self.ni, self.nj), order="F")

    def op_ter(self, ga, gb, var, dft=0.0):
        var = np.asfortranarray(var, np.float64)
        nd3 = 1 if var.ndim == 2 else var.shape[2]
        out = self._out(nd3)
        lib().orc_op_ter(self.h, ga, gb, nd3, dft, _dp(var), _dp(out))
        return out

    def op_px1(self, ga, gb, var):
        var = np.asfortranarray(var, np.float64)
        nd3 = 1 if var.ndim == 2 else var.shape[2]
        out = self._out(nd3)
        lib().orc_op_px1(self.h, ga, gb, nd3, _dp(var), _dp(out))
        return out

    def op_px2(self, ga, gb, var):
        var = np.asfortranarray(var, np.float64)
        nd3 = 1 if var.ndim == 2 else var.shape[2]
        out = self._out(nd3)
        lib().orc_op_px2(self.h, ga, gb, nd3, _dp(var), _dp(out))
        return out

    def op_div(self, ga, gb, gc, va, vb):
        va = np.asfortranarray(va, np.float64)
        vb = np.asfortranarray(vb, np.float64)
        nd3 = 1 if va.ndim == 2 else va.shape[2]
        out = self._out(nd3)
        lib().orc_op_div(self.h, ga, gb, gc, nd3, _dp(va), _dp(vb), _dp(out))
        return out

    def op_lap(self, ga, gb, var):
        var = np.asfortranarray(var, np.float64)
        out = self._out(1)
        lib().orc_op_lap(self.h, ga, gb, _dp(var), _dp(out))
        return out

    # composite operators on the context's fields (set them with set())
    def op_adv(self, var, sch):
        """op_adv (mod_op.f90:224-278) of var on U points with the
        context's uc, vc, wm, chc."""
        var = np.asfortranarray(var, np.float64)
        sch = np.asfortranarray(sch, np.float64)
        out = self._out(self.nk)
        lib().orc_op_adv(self.h, _dp(var), _dp(sch), _dp(out))
        return out

    def op_adv_ts(self, var, wm):
        """op_adv_ts (mod_op.f90:280-325) of var on T points with the
        context's uc, vc, chc and the given wm (nk+1 levels)."""
        var = np.asfortranarray(var, np.float64)
        wm = np.asfortranarray(wm, np.float64)
        out = self._out(self.nk)
        lib().orc_op_adv_ts(self.h, _dp(var), _dp(wm), _dp(out))
        return out

    def op_dif(self, ch, var):
        """op_dif (mod_op.f90:186-222) of var on T points with the
        context's kh."""
        var = np.asfortranarray(var, np.float64)
        ch = np.asfortranarray(ch, np.float64)
        out = self._out(self.nk)
        lib().orc_op_dif(self.h, _dp(ch), _dp(var), _dp(out))
        return out

    def op_fri(self, sch):
        """op_fri (mod_op.f90:169-184): returns (fx, fy) from the context's
        uc, vc, am, km, wind stress and chc."""
        sch = np.asfortranarray(sch, np.float64)
        lib().orc_op_fri(self.h, _dp(sch))
        return self.get("fx"), self.get("fy")

    def smag(self, u, v):
        """smag (mod_int.f90:403-440): returns am from unweighted u, v (g3);
        am keeps its previous values where smag does not write it."""
        u = np.asfortranarray(u, np.float64)
        v = np.asfortranarray(v, np.float64)
        lib().orc_smag(self.h, _dp(u), _dp(v))
        return self.get("am")

    def upwelling(self):
        """upwelling (mod_int.f90:626-647): returns wm from the context's
        uc, vc, chp and wm(1)."""
        lib().orc_upwelling(self.h)
        return self.get("wm")


_LIB_R8 = None


def den_rho_r8(t, s, p):
    """den_rho with FP64 literals (the reference's -r8 build variant); used
    only to check the EOS against the notes.rst:20-26 KAT statement."""
    global _LIB_R8
    if _LIB_R8 is None:
        path = os.path.join(_HERE, "liboracle_r8.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle R8 variant not built: run `make -C oracle`")
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        L.orc_den_rho_n.argtypes = [C.c_size_t, dp, dp, dp, dp]
        _LIB_R8 = L
    t, s, p = np.broadcast_arrays(np.asarray(t, np.float64), np.asarray(s, np.float64),
                                  np.asarray(p, np.float64))
    t, s, p = (np.ascontiguousarray(x) for x in (t, s, p))
    out = np.empty_like(t)
    _LIB_R8.orc_den_rho_n(t.size, _dp(t), _dp(s), _dp(p), _dp(out))
    return out
