"""Error of the mixed-precision forms against the FP64 form.

Runs the FP64 form and one or more mixed forms side by side from the same
initial state and reports, at chosen simulated times, the area-weighted RMSE of
every prognostic field over sea points (weights rh*dx1*dx2 of the T grid,
mask from lev; the U-grid fields use the U mask).  The FP64 form is
bit-identical to the FP64 oracle (tests/test_gpu_parity.py), so it stands in
for the oracle over runs the CPU cannot finish.

    python -m pcom_amd.rmse --days 365 --every 30 [--forms mixed,mixed+surrogate]

A form "<form>+surrogate" fits the learned barotropic surrogate on a
scratch model of the form (pcom_amd.surrogate.fit: --spin-days, then
--snapshots recorded plain subcycles, training, on-policy rounds); its own
model steps plainly, alongside the other forms, to the first whole day
after the fit's steps, and from then on runs with the fitted network under
the policy pcom_amd.surrogate.USE (the plain subcycle judging every 10th
step).  The form "fp64+f32init" is the FP64 form started
from the initial T and S rounded to FP32: its distance from FP64 is the
model's own sensitivity to perturbations at FP32 rounding level, the floor
any non-bit-identical form meets.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import abi, configs, surrogate
from .lib import Model

# the prognostic state a restart file holds (mod_io.f90:80-113), less the
# copies a step leaves equal (tp = tc, up = uc, chp = chc, bup = buc) and
# the forcing, plus the diagnosed SSH and the step-mean free surface
FIELDS_T = ["tc", "sc", "prho_t", "prho_s"]
FIELDS_U = ["uc", "vc", "auc", "avc", "aup", "avp", "aupp", "avpp", "am"]
FIELDS_2D = ["chc", "mch", "mch2", "ssh", "buc", "bvc"]


def weights(m, cfg):
    area = m.metric(1, 0) * m.metric(1, 2) * m.metric(1, 3)
    lev_t = np.zeros((m.ni, m.nj), dtype=np.int32)
    lev_t[1:-1, 1:-1] = cfg.itn
    inner = np.zeros((m.ni, m.nj), dtype=bool)
    inner[1:-1, 1:-1] = True
    prh_u = m.get("prh_u")
    wet_u = (prh_u != np.float64(np.float32(0.1))) & inner
    return area, lev_t, inner, wet_u


def rmse_report(ma, mb, cfg):
    area, lev_t, inner, wet_u = weights(ma, cfg)
    out = {}
    sea1 = (lev_t >= 1) & inner
    for name in FIELDS_T:
        a, b = ma.get(name), mb.get(name)
        w = area[..., None] * ((lev_t[..., None] >= np.arange(1, ma.nk + 1)) & inner[..., None])
        out[name] = float(np.sqrt(np.sum(w * (a - b) ** 2) / np.sum(w)))
        if name == "tc":
            d = (a[:, :, 0] - b[:, :, 0])[sea1]
            out["sst"] = float(np.sqrt(np.sum(area[sea1] * d * d) / np.sum(area[sea1])))
    for name in FIELDS_U:
        a, b = ma.get(name), mb.get(name)
        w = area[..., None] * wet_u[..., None]
        out[name] = float(np.sqrt(np.sum(w * (a - b) ** 2) / np.sum(w)))
    for name in FIELDS_2D:
        a, b = ma.get(name), mb.get(name)
        msk = sea1 if name in ("chc", "mch", "mch2", "ssh") else wet_u
        d = (a - b)[msk]
        out[name] = float(np.sqrt(np.sum(area[msk] * d * d) / np.sum(area[msk])))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--config", default="global", choices=list(configs.CONFIGS))
    ap.add_argument("--days", type=int, default=365)
    ap.add_argument("--every", type=int, default=30, help="report every N days")
    ap.add_argument("--forms", default="mixed", help="comma-separated forms to compare with FP64")
    ap.add_argument("--spin-days", type=int, default=10, help="surrogate forms: plain days before recording")
    ap.add_argument("--snapshots", type=int, default=1000, help="surrogate forms: steps recorded for training")
    ap.add_argument("--iters", type=int, default=6000, help="surrogate forms: training iterations")
    ap.add_argument("--onpolicy-rounds", type=int, default=3, help="surrogate forms: on-policy rounds")
    ap.add_argument("--onpolicy-steps", type=int, default=240, help="surrogate forms: steps per on-policy round")
    ap.add_argument("--onpolicy-iters", type=int, default=2000, help="surrogate forms: iterations after each round")
    ap.add_argument("--check-every", type=int, default=surrogate.USE["check_every"])
    ap.add_argument("--tol", type=float, default=surrogate.USE["tol"],
                    help="surrogate check bound (pcom_nn_use; <0: relative to persistence)")
    ap.add_argument("--resume", type=int, default=surrogate.USE["resume"],
                    help="plain steps after a failed check (pcom_nn_resume)")
    ap.add_argument("--out", default=None, help="write the JSON report here")
    args = ap.parse_args(argv)
    cfg = configs.CONFIGS[args.config]()
    forms = args.forms.split(",")
    ref = Model(cfg, precision=abi.PCOM_FP64)
    models = {}
    steps_per_day = 86400 // cfg.dtbc
    for f in forms:
        base = f.split("+")[0]
        models[f] = Model(cfg, precision=abi.FORMS[base])
        if f.endswith("+f32init"):
            for name in ("tc", "sc"):
                models[f].set(name, models[f].get(name).astype(np.float32).astype(np.float64))
    report = {"config": cfg.name, "dims": [cfg.glo_ni, cfg.glo_nj, cfg.nk], "records": {f: [] for f in forms},
              "surrogate": {}}
    t0 = time.time()
    # surrogate forms: fit on a scratch model of the form (its on-policy
    # rounds leave that model off the plain trajectory), then every model
    # steps plainly to the next whole day after the fit, from which the
    # fitted network is in use
    taken, params = {}, {}
    for f, m in models.items():
        if f.endswith("+surrogate"):
            scratch = Model(cfg, precision=m.precision)
            rep = surrogate.fit(scratch, cfg, spin_days=args.spin_days, snapshots=args.snapshots, iters=args.iters,
                                onpolicy_rounds=args.onpolicy_rounds, onpolicy_steps=args.onpolicy_steps,
                                onpolicy_iters=args.onpolicy_iters, log=lambda s: print(s, flush=True))
            params[f] = scratch.nn_params()
            scratch.close()
            taken[f] = rep["steps"]
            report["surrogate"][f] = {k: rep[k] for k in ("steps", "recorded", "holdout_gain", "train_s", "onpolicy")}
    day = 0
    if taken:
        day = -(-max(taken.values()) // steps_per_day)
        ref.step(day * steps_per_day)
        for f, m in models.items():
            m.step(day * steps_per_day)
            if f in taken:
                m.nn_create(capacity=0, seed=1)
                m.nn_params(params[f])
                m.nn_use(args.check_every, args.tol, args.resume)
                report["surrogate"][f]["in_use_from_day"] = day
    while day < args.days:
        n = min(args.every, args.days - day)
        ref.step(n * steps_per_day)
        for m in models.values():
            m.step(n * steps_per_day)
        day += n
        for f, m in models.items():
            rec = {"form": f, "day": day, "finite": bool(ref.check_finite() and m.check_finite())}
            rec.update(rmse_report(m, ref, cfg))
            if f.endswith("+surrogate"):
                st = m.nn_stats()
                rec["surrogate_steps"] = st.surrogate_steps
                rec["in_use"] = st.in_use
                rec["fallbacks"] = st.fallbacks
                rec["check_rmse"] = list(st.last_check_rmse)
                rec["persist_rmse"] = list(st.last_persist_rmse)
            report["records"][f].append(rec)
            print(json.dumps(rec), flush=True)
    report["wall_s"] = time.time() - t0
    report["gate"] = {}
    for f in forms:
        last = report["records"][f][-1]
        report["gate"][f] = {"sst_K": 1e-3, "ssh_m": 1e-3, "sst": last["sst"], "ssh": last["ssh"],
                             "pass": bool(last["finite"] and last["sst"] <= 1e-3 and last["ssh"] <= 1e-3)}
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(report, fh, indent=1)
    print(json.dumps(report["gate"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
