"""Loader and Python mirror of the C ABI in include/pcom_gpu.h.

``Model`` mirrors the reference's driver loop for the time-stepping path
(main.f90:158-238): create a context from raw inputs, ``step(n)``, read or
write fields.  The HIP library is required; there is no CPU fallback, and
a missing or unloadable library raises at import of ``lib()``.
"""
import atexit
import ctypes as C
import os
import sys
import weakref

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpcom_gpu.so")
_LIB = None
# contexts and ensembles not closed yet: closed by _close_all at interpreter
# exit, while the HIP runtime is still fully up (ensembles before members)
_LIVE_ENSEMBLES = weakref.WeakSet()
_LIVE_MODELS = weakref.WeakSet()

# every symbol include/pcom_gpu.h declares
EXPORTS = [
    "pcom_create", "pcom_step", "pcom_get_dims", "pcom_field_size", "pcom_get_field",
    "pcom_set_field", "pcom_get_counters", "pcom_global_sums", "pcom_check_finite", "pcom_sync",
    "pcom_phase_times", "pcom_destroy", "pcom_last_error", "pcom_op_ter", "pcom_op_px1",
    "pcom_op_px2", "pcom_op_div", "pcom_op_lap", "pcom_den_rho", "pcom_get_metric",
    "pcom_set_clock", "pcom_output_means", "pcom_forcing_due", "pcom_set_forcing",
    "pcom_output_means_noalloc",
    "pcom_nn_create", "pcom_nn_record", "pcom_nn_augment", "pcom_nn_train", "pcom_nn_use", "pcom_nn_resume", "pcom_nn_eval", "pcom_nn_get_stats",
    "pcom_nn_param_count", "pcom_nn_get_params", "pcom_nn_set_params", "pcom_nn_forward_backward",
    "pcom_ensemble_create", "pcom_ensemble_step", "pcom_ensemble_packs", "pcom_ensemble_destroy",
    "pcom_tune_subcycle", "pcom_subcycle_tiling", "pcom_subcycle_wait_limit",
]


class PcomError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pcom error {code}: {msg}")
        self.code = code


def lib():
    """Load libpcom_gpu.so (built by __graft_entry__.build())."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C pcom_amd/csrc` "
                               "or __graft_entry__.build(); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        vp, dp, sz = C.c_void_p, C.POINTER(C.c_double), C.c_size_t
        L.pcom_create.argtypes = [C.POINTER(abi.PcomInputs), C.POINTER(abi.PcomOpts), C.POINTER(vp)]
        L.pcom_step.argtypes = [vp, C.c_int]
        L.pcom_get_dims.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.pcom_field_size.argtypes = [vp, C.c_int]
        L.pcom_field_size.restype = sz
        L.pcom_get_field.argtypes = [vp, C.c_int, dp, sz]
        L.pcom_set_field.argtypes = [vp, C.c_int, dp, sz]
        L.pcom_get_counters.argtypes = [vp, C.POINTER(abi.PcomCounters)]
        L.pcom_global_sums.argtypes = [vp, dp]
        L.pcom_check_finite.argtypes = [vp]
        L.pcom_sync.argtypes = [vp]
        L.pcom_phase_times.argtypes = [vp, C.c_int, dp, C.POINTER(C.c_int)]
        L.pcom_destroy.argtypes = [vp]
        L.pcom_last_error.restype = C.c_char_p
        L.pcom_op_ter.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double, dp, dp]
        L.pcom_op_px1.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp]
        L.pcom_op_px2.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp]
        L.pcom_op_div.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, dp]
        L.pcom_op_lap.argtypes = [vp, C.c_int, C.c_int, dp, dp]
        L.pcom_den_rho.argtypes = [vp, sz, dp, dp, dp, dp]
        L.pcom_get_metric.argtypes = [vp, C.c_int, C.c_int, dp, sz]
        L.pcom_set_clock.argtypes = [vp, C.c_longlong, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.pcom_output_means.argtypes = [vp, C.POINTER(abi.PcomMeans)]
        L.pcom_output_means_noalloc.argtypes = [vp, C.POINTER(abi.PcomMeans)]
        L.pcom_forcing_due.argtypes = [vp, C.POINTER(C.c_int)]
        L.pcom_set_forcing.argtypes = [vp, C.POINTER(abi.PcomForcing), C.c_int]
        fp = C.POINTER(C.c_float)
        L.pcom_nn_create.argtypes = [vp, C.c_int, C.c_ulonglong]
        L.pcom_nn_record.argtypes = [vp, C.c_int]
        L.pcom_nn_augment.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_ulonglong]
        L.pcom_nn_train.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_ulonglong, dp]
        L.pcom_nn_use.argtypes = [vp, C.c_int, C.c_double]
        L.pcom_nn_resume.argtypes = [vp, C.c_int]
        L.pcom_nn_eval.argtypes = [vp, C.c_int, C.c_int, dp]
        L.pcom_nn_get_stats.argtypes = [vp, C.POINTER(abi.PcomNnStats)]
        L.pcom_nn_param_count.argtypes = [vp]
        L.pcom_nn_param_count.restype = sz
        L.pcom_nn_get_params.argtypes = [vp, fp, sz]
        L.pcom_nn_set_params.argtypes = [vp, fp, sz]
        L.pcom_nn_forward_backward.argtypes = [vp, C.c_int, fp, fp, fp, fp, dp]
        ip = C.POINTER(C.c_int)
        L.pcom_ensemble_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(vp)]
        L.pcom_ensemble_step.argtypes = [vp, C.c_int]
        L.pcom_ensemble_packs.argtypes = [vp, ip, ip]
        L.pcom_ensemble_destroy.argtypes = [vp]
        L.pcom_tune_subcycle.argtypes = [vp, C.c_int, ip, dp]
        L.pcom_subcycle_tiling.argtypes = [vp, C.c_int, ip]
        L.pcom_subcycle_wait_limit.argtypes = [vp, C.c_ulonglong]
        for name in EXPORTS:
            if name not in ("pcom_field_size", "pcom_last_error", "pcom_nn_param_count"):
                getattr(L, name).restype = C.c_int
        _LIB = L
        atexit.register(_close_all)
    return _LIB


def _close_all():
    """Destroy every live ensemble, then every live context, in the order
    pcom_destroy needs (its streams drained, its device memory freed) and
    before Python's finalisation and the C runtime's exit handlers run:
    nothing of the library is left for object finalisers at shutdown or for
    the HIP runtime's own teardown (or a profiler's) to race with."""
    for e in list(_LIVE_ENSEMBLES):
        e.close()
    for m in list(_LIVE_MODELS):
        m.close()


def _check(rc):
    if rc != abi.PCOM_OK:
        raise PcomError(rc, lib().pcom_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Ensemble:
    """Several Models stepped together (pcom_ensemble_*)."""

    def __init__(self, members):
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = C.c_void_p()
        _check(lib().pcom_ensemble_create(arr, len(self.members), C.byref(h)))
        self.h = h
        _LIVE_ENSEMBLES.add(self)

    def step(self, n=1):
        _check(lib().pcom_ensemble_step(self.h, int(n)))

    def packs(self):
        npk = C.c_int()
        of = (C.c_int * len(self.members))()
        _check(lib().pcom_ensemble_packs(self.h, C.byref(npk), of))
        return npk.value, list(of)

    def close(self):
        if getattr(self, "h", None) is not None and _LIB is not None:
            _LIB.pcom_ensemble_destroy(self.h)
            self.h = None
        _LIVE_ENSEMBLES.discard(self)

    def __del__(self):
        # at interpreter shutdown _close_all has already run; a finaliser
        # that reaches here then must not call into the HIP runtime
        if not sys.is_finalizing():
            self.close()


def den_rho(t, s, p):
    """den_rho (mod_den.f90:57-118) evaluated by the HIP kernel."""
    t, s, p = np.broadcast_arrays(np.asarray(t, np.float64), np.asarray(s, np.float64),
                                  np.asarray(p, np.float64))
    t, s, p = (np.ascontiguousarray(x) for x in (t, s, p))
    out = np.empty_like(t)
    _check(lib().pcom_den_rho(None, t.size, _dp(t), _dp(s), _dp(p), _dp(out)))
    return out


class Model:
    """One device-resident PCOM time stepper (a pcom_ctx)."""

    def __init__(self, cfg, precision=abi.PCOM_FP64, device=0, restart_run=0, trop_tile=(0, 0, 0),
                 trop_mode=abi.PCOM_TROP_AUTO):
        self.cfg = cfg
        self.precision = precision
        self._inp = cfg.c_inputs()
        self._opt = cfg.c_opts(precision=precision, device=device, restart_run=restart_run,
                               trop_tile=trop_tile, trop_mode=trop_mode)
        h = C.c_void_p()
        _check(lib().pcom_create(C.byref(self._inp), C.byref(self._opt), C.byref(h)))
        self.h = h
        _LIVE_MODELS.add(self)
        ni, nj, nk = C.c_int(), C.c_int(), C.c_int()
        _check(lib().pcom_get_dims(self.h, C.byref(ni), C.byref(nj), C.byref(nk)))
        self.ni, self.nj, self.nk = ni.value, nj.value, nk.value

    def close(self):
        if getattr(self, "h", None) is not None and _LIB is not None:
            _LIB.pcom_destroy(self.h)
            self.h = None
        _LIVE_MODELS.discard(self)

    def __del__(self):
        if not sys.is_finalizing():
            self.close()

    def step(self, n=1):
        _check(lib().pcom_step(self.h, int(n)))

    def sync(self):
        _check(lib().pcom_sync(self.h))

    def shape(self, name):
        nd3 = abi.field_levels(name, self.nk)
        return (self.ni, self.nj) if nd3 == 1 else (self.ni, self.nj, nd3)

    def get(self, name):
        out = np.empty(self.shape(name), dtype=np.float64, order="F")
        _check(lib().pcom_get_field(self.h, abi.FIELD[name], _dp(out), out.size))
        return out

    def set(self, name, arr):
        a = np.asfortranarray(arr, dtype=np.float64)
        if a.shape != self.shape(name):
            raise ValueError(f"{name}: shape {a.shape} != {self.shape(name)}")
        _check(lib().pcom_set_field(self.h, abi.FIELD[name], _dp(a), a.size))

    def counters(self):
        c = abi.PcomCounters()
        _check(lib().pcom_get_counters(self.h, C.byref(c)))
        return c

    def set_clock(self, step, cdate, pdate):
        """Restart clock (main.f90:128-141): last completed step index and
        the current / previous dates as (y, m, d, h, mi, s)."""
        c = (C.c_int * 6)(*cdate)
        p = (C.c_int * 6)(*pdate)
        _check(lib().pcom_set_clock(self.h, int(step), c, p))

    def forcing_due(self):
        """(daily, six_hourly): the forcing refreshes the next step performs
        (main.f90:164, 184-185)."""
        d = (C.c_int * 2)()
        _check(lib().pcom_forcing_due(self.h, d))
        return bool(d[0]), bool(d[1])

    def set_forcing(self, interpolate, taux=None, tauy=None, pa=None, bct=None):
        """Forcing records (glo_ni, glo_nj) for the next step (main.f90:
        164-202): interpolate=True is the daily refresh (int_bnd, then the
        records), False the 6-hourly one (records only); den_prho follows."""
        f = abi.forcing(taux, tauy, pa, bct)
        _check(lib().pcom_set_forcing(self.h, C.byref(f), int(bool(interpolate))))

    def output_means(self, names=abi.MEAN_NAMES, noalloc=False):
        """Time means of the output interval (main.f90:726-760); resets the
        accumulators.  Returns {name: array (glo_ni, glo_nj[, nk])}.
        noalloc=True stages them in the library's static device buffer
        (pcom_output_means_noalloc)."""
        g = (self.ni - 2, self.nj - 2)
        out, m = {}, abi.PcomMeans()
        for n in names:
            shp = g if n in ("ch", "ssh") else g + (self.nk,)
            out[n] = np.empty(shp, order="F")
            setattr(m, n, _dp(out[n]))
        fn = lib().pcom_output_means_noalloc if noalloc else lib().pcom_output_means
        _check(fn(self.h, C.byref(m)))
        return out

    def tune_subcycle(self, trials=8):
        """Autotune the subcycle tiling; returns ((tile_i, tile_j, substeps), ms)."""
        t = (C.c_int * 3)()
        ms = C.c_double()
        _check(lib().pcom_tune_subcycle(self.h, int(trials), t, C.byref(ms)))
        return tuple(t), ms.value

    def wait_limit(self, polls=0):
        """Bound each neighbour wait of the persistent subcycle to `polls`
        unsuccessful polls (pcom_subcycle_wait_limit; 0: the default)."""
        _check(lib().pcom_subcycle_wait_limit(self.h, int(polls)))

    def tiling(self, tile=None, mode=None):
        """The subcycle tiling (tile_i, tile_j, substeps); sets it if given
        (with `mode`, a PCOM_TROP_* value: default the current mode)."""
        if tile is None and mode is None:
            t = (C.c_int * 4)()
            _check(lib().pcom_subcycle_tiling(self.h, 0, t))
            return tuple(t)[:3]
        cur = (C.c_int * 4)()
        _check(lib().pcom_subcycle_tiling(self.h, 0, cur))
        t = (C.c_int * 4)(*(tuple(tile or (0, 0, 0)) + (cur[3] if mode is None else mode,)))
        _check(lib().pcom_subcycle_tiling(self.h, 1, t))
        return tuple(t)[:3]

    def subcycle_mode(self):
        """PCOM_TROP_PERSISTENT or PCOM_TROP_LAUNCHES: how the subcycle runs."""
        t = (C.c_int * 4)()
        _check(lib().pcom_subcycle_tiling(self.h, 0, t))
        return t[3]

    # ---- learned barotropic surrogate
    def nn_create(self, capacity=0, seed=1):
        _check(lib().pcom_nn_create(self.h, int(capacity), int(seed)))

    def nn_record(self, on=True):
        _check(lib().pcom_nn_record(self.h, int(bool(on))))

    def nn_augment(self, fraction, amp_ch, amp_u, seed=1):
        """Record perturbed snapshots beside the plain ones (pcom_nn_augment)."""
        _check(lib().pcom_nn_augment(self.h, float(fraction), float(amp_ch), float(amp_u), int(seed)))

    def nn_train(self, iters, batch=4, lr=1e-3, seed=1):
        loss = np.zeros(max(int(iters), 1))
        _check(lib().pcom_nn_train(self.h, int(iters), int(batch), float(lr), int(seed), _dp(loss)))
        return loss[:iters]

    def nn_use(self, check_every, tol=0.0, resume=None):
        _check(lib().pcom_nn_use(self.h, int(check_every), float(tol)))
        if resume is not None:
            _check(lib().pcom_nn_resume(self.h, int(resume)))

    def nn_eval(self, first, count):
        """(network RMSE[4], persistence RMSE[4]) on recorded snapshots."""
        r = np.zeros(8)
        _check(lib().pcom_nn_eval(self.h, int(first), int(count), _dp(r)))
        return r[:4], r[4:]

    def nn_stats(self):
        st = abi.PcomNnStats()
        _check(lib().pcom_nn_get_stats(self.h, C.byref(st)))
        return st

    def nn_params(self, values=None):
        n = lib().pcom_nn_param_count(self.h)
        fp = C.POINTER(C.c_float)
        if values is None:
            out = np.empty(n, np.float32)
            _check(lib().pcom_nn_get_params(self.h, out.ctypes.data_as(fp), n))
            return out
        v = np.ascontiguousarray(values, np.float32)
        if v.size != n:
            raise ValueError(f"{v.size} parameters given, the network has {n}")
        _check(lib().pcom_nn_set_params(self.h, v.ctypes.data_as(fp), n))

    def nn_forward_backward(self, x, target):
        fp = C.POINTER(C.c_float)
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        out = np.empty(t.shape, np.float32)
        grads = np.empty(lib().pcom_nn_param_count(self.h), np.float32)
        loss = C.c_double()
        _check(lib().pcom_nn_forward_backward(self.h, x.shape[0], x.ctypes.data_as(fp), t.ctypes.data_as(fp),
                                              out.ctypes.data_as(fp), grads.ctypes.data_as(fp), C.byref(loss)))
        return loss.value, out, grads

    def global_sums(self):
        out = np.zeros(3)
        _check(lib().pcom_global_sums(self.h, _dp(out)))
        return out

    def check_finite(self):
        rc = lib().pcom_check_finite(self.h)
        return rc == abi.PCOM_OK

    def phase_times(self, enable=True):
        out = np.zeros(3)
        n = C.c_int()
        _check(lib().pcom_phase_times(self.h, int(enable), _dp(out), C.byref(n)))
        return out, n.value

    def metric(self, g, which):
        out = np.empty((self.ni, self.nj), order="F")
        _check(lib().pcom_get_metric(self.h, g, which, _dp(out), out.size))
        return out

    def _out(self, nd3):
        return np.zeros((self.ni, self.nj, nd3) if nd3 > 1 else (self.ni, self.nj), order="F")

    def op_ter(self, ga, gb, var, dft=0.0):
        var = np.asfortranarray(var, np.float64)
        nd3 = 1 if var.ndim == 2 else var.shape[2]
        out = self._out(nd3)
        _check(lib().pcom_op_ter(self.h, ga, gb, nd3, dft, _dp(var), _dp(out)))
        return out

    def op_px1(self, ga, gb, var):
        var = np.asfortranarray(var, np.float64)
        nd3 = 1 if var.ndim == 2 else var.shape[2]
        out = self._out(nd3)
        _check(lib().pcom_op_px1(self.h, ga, gb, nd3, _dp(var), _dp(out)))
        return out

    def op_px2(self, ga, gb, var):
        var = np.asfortranarray(var, np.float64)
        nd3 = 1 if var.ndim == 2 else var.shape[2]
        out = self._out(nd3)
        _check(lib().pcom_op_px2(self.h, ga, gb, nd3, _dp(var), _dp(out)))
        return out

    def op_div(self, ga, gb, gc, va, vb):
        va = np.asfortranarray(va, np.float64)
        vb = np.asfortranarray(vb, np.float64)
        nd3 = 1 if va.ndim == 2 else va.shape[2]
        out = self._out(nd3)
        _check(lib().pcom_op_div(self.h, ga, gb, gc, nd3, _dp(va), _dp(vb), _dp(out)))
        return out

    def op_lap(self, ga, gb, var):
        var = np.asfortranarray(var, np.float64)
        out = self._out(1)
        _check(lib().pcom_op_lap(self.h, ga, gb, _dp(var), _dp(out)))
        return out
