"""The learned barotropic surrogate workflow on one configuration.

1. spin the model up for `spin_days` with the plain subcycle;
2. record `snapshots` consecutive steps (inputs and outputs of the plain
   subcycle) into the device store;
3. train the BF16 network on them (pcom_nn_train);
3b. a few on-policy rounds: the network in use while recording, each of
   its steps beside the plain subcycle's answer from the same state (the
   states the network itself leads to), then more training on all;
4. record `holdout` further steps and measure, on those, the network's
   RMSE against the plain subcycle beside the RMSE of persistence (the
   step without the network: no change of the barotropic mode);
5. time plain steps and surrogate steps (plain judging subcycle every
   `check_every`-th step, the policy USE);
6. with --closed-loop-days D: run D days with the network in use (on a
   model that took the fit's steps plainly) beside a plain twin started
   from the same state, and report the RMSE of SSH,
   SST, the free surface and the barotropic velocity between them daily.

    python -m pcom_amd.surrogate --config global --spin-days 10 \\
        --snapshots 1000 --iters 6000 --closed-loop-days 30 --out report.json
"""
import argparse
import json
import sys
import time

import numpy as np

from . import abi, configs
from .lib import Model

OUTPUTS = ["mch2", "mch", "ubt", "vbt"]


def step_scales(m, cfg):
    """Sea RMS of one step's change of ch and of the barotropic velocity
    (the model takes that step): the size of the perturbations the
    augmented snapshots use."""
    from . import rmse
    area, lev_t, inner, wet_u = rmse.weights(m, cfg)
    sea = (lev_t > 0) & inner
    c0, u0, v0 = m.get("chc"), m.get("buc"), m.get("bvc")
    m.step(1)
    dc = m.get("chc") - c0
    du = np.concatenate([(m.get("buc") - u0)[wet_u], (m.get("bvc") - v0)[wet_u]])
    return float(np.sqrt(np.mean(dc[sea] ** 2))), float(np.sqrt(np.mean(du ** 2)))


def train(m, iters, batch=8, lr=1e-3, seed=2, log=None):
    """`iters` Adam iterations on the recorded snapshots, the learning rate
    on a cosine from `lr` to 0 over the run (pcom_nn_train); returns the
    loss of every iteration."""
    loss = m.nn_train(iters, batch=batch, lr=lr, seed=seed)
    if log:
        for k in range(1000, iters + 1, 1000):
            log(f"  {k} iterations, loss {loss[k - 50:k].mean():.3e}")
    return loss


# how the network is used once trained: a plain checking subcycle every
# 10th step; a check fails when an output's error exceeds twice that of
# persistence (now, or at the first check); a failure hands 10 steps to
# the plain subcycle, after which the network resumes
USE = {"check_every": 10, "tol": -2.0, "resume": 10}


def prepare(cfg, precision=abi.PCOM_MIXED, **kw):
    """A new model of `cfg` put through fit(); returns (model, report)."""
    m = Model(cfg, precision=precision)
    return m, fit(m, cfg, **kw)


def fit(m, cfg, spin_days=10, snapshots=1000, holdout=48, iters=6000, batch=8, lr=1e-3, seed=1, augment=0.5,
        amp=1.0, onpolicy_rounds=3, onpolicy_steps=240, onpolicy_iters=2000, onpolicy_lr=3e-4, log=print):
    """On model m: spin up, record (with a fraction `augment` of perturbed
    snapshots beside the plain ones), train; then `onpolicy_rounds` rounds
    of `onpolicy_steps` steps with the network in use (policy USE) while
    recording, each step's state with the plain subcycle's answer, each
    round followed by `onpolicy_iters` more training iterations on all
    snapshots; then evaluate on `holdout` new plain steps.  Returns the
    report; report["steps"] is the number of steps the model took (the
    on-policy steps leave it off the plain trajectory)."""
    spd = 86400 // cfg.dtbc
    m.step(spin_days * spd)
    amp_ch, amp_u = step_scales(m, cfg)
    # perturbed snapshots come two at a time before a random fraction
    # `augment` of the steps: room for six standard deviations more
    extra = 2 * (snapshots * augment + 6.0 * np.sqrt(snapshots * augment * (1.0 - augment)))
    cap = snapshots + int(np.ceil(extra)) + onpolicy_rounds * onpolicy_steps + holdout + 8
    m.nn_create(capacity=cap, seed=seed)
    m.nn_augment(augment, amp * amp_ch, amp * amp_u, seed=seed + 100)
    m.nn_record(True)
    m.step(snapshots)
    m.nn_record(False)
    m.nn_augment(0.0, 0.0, 0.0)
    t0 = time.time()
    loss = train(m, iters, batch, lr, seed + 1, log)
    train_s = time.time() - t0
    log(f"trained {iters} iterations in {train_s:.1f} s, loss {loss[:20].mean():.3e} -> {loss[-100:].mean():.3e}")
    onpolicy = []
    for r in range(onpolicy_rounds):
        m.nn_use(**USE)
        m.nn_record(True)
        m.step(onpolicy_steps)
        m.nn_record(False)
        st = m.nn_stats()
        m.nn_use(0)
        t0 = time.time()
        ol = train(m, onpolicy_iters, batch, onpolicy_lr, seed + 2 + r, log)
        train_s += time.time() - t0
        rec = {"round": r + 1, "surrogate_steps": st.surrogate_steps, "fallbacks": st.fallbacks,
               "check_rmse": list(st.last_check_rmse), "persist_rmse": list(st.last_persist_rmse),
               "loss_last": float(ol[-100:].mean())}
        onpolicy.append(rec)
        log(json.dumps(rec))
    first = m.nn_stats().recorded
    m.nn_record(True)
    m.step(holdout)
    m.nn_record(False)
    net, per = m.nn_eval(first, holdout)
    report = {
        "config": cfg.name, "spin_days": spin_days, "snapshots": snapshots, "holdout": holdout,
        "augment": augment, "perturbation_scales": [amp_ch, amp_u], "recorded": first + holdout,
        "steps": spin_days * spd + 1 + snapshots + onpolicy_rounds * onpolicy_steps + holdout,
        "iters": iters, "batch": batch, "lr": lr, "train_s": train_s,
        "onpolicy": onpolicy, "onpolicy_iters": onpolicy_iters, "onpolicy_lr": onpolicy_lr,
        "loss_first": float(loss[:20].mean()), "loss_last": float(loss[-100:].mean()),
        "holdout_rmse_network": dict(zip(OUTPUTS, map(float, net))),
        "holdout_rmse_persistence": dict(zip(OUTPUTS, map(float, per))),
        "holdout_gain": dict(zip(OUTPUTS, map(float, per / np.maximum(net, 1e-300)))),
    }
    log("held-out gain over persistence: " + ", ".join(f"{k} {v:.3f}" for k, v in report["holdout_gain"].items()))
    return report


def closed_loop(m, cfg, steps_done, days, check_every=USE["check_every"], tol=USE["tol"], resume=USE["resume"],
                log=print):
    """The network of model m (fitted in `steps_done` steps) in use for
    `days` days on a new model of the same form that took `steps_done`
    plain steps, beside a plain twin from that state; daily RMSE between
    the two."""
    from . import rmse
    params = m.nn_params()
    m = Model(cfg, precision=m.precision)
    m.step(steps_done)
    m.nn_create(capacity=0, seed=1)
    m.nn_params(params)
    twin = Model(cfg, precision=m.precision)
    twin.step(steps_done)
    spd = 86400 // cfg.dtbc
    m.nn_use(check_every, tol, resume)
    recs = []
    for d in range(days):
        m.step(spd)
        twin.step(spd)
        r = rmse.rmse_report(m, twin, cfg)
        st = m.nn_stats()
        rec = {"day": d + 1, "sst": r["sst"], "ssh": r["ssh"], "chc": r["chc"], "buc": r["buc"], "bvc": r["bvc"],
               "tc": r["tc"], "uc": r["uc"], "in_use": bool(st.in_use), "surrogate_steps": st.surrogate_steps,
               "fallbacks": st.fallbacks,
               "check_rmse": list(st.last_check_rmse), "persist_rmse": list(st.last_persist_rmse)}
        recs.append(rec)
        log(json.dumps(rec))
    twin.close()
    m.close()
    return recs


def time_steps(m, nsteps, check_every=USE["check_every"], tol=USE["tol"], resume=USE["resume"]):
    """ms per step of plain steps, then of steps with the surrogate."""
    m.nn_use(0)
    m.sync()
    t0 = time.perf_counter()
    m.step(nsteps)
    m.sync()
    plain = 1000.0 * (time.perf_counter() - t0) / nsteps
    m.nn_use(check_every, tol, resume)
    m.sync()
    t0 = time.perf_counter()
    m.step(nsteps)
    m.sync()
    sur = 1000.0 * (time.perf_counter() - t0) / nsteps
    return plain, sur


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--config", default="global", choices=list(configs.CONFIGS))
    ap.add_argument("--form", default="mixed", choices=list(abi.FORMS))
    ap.add_argument("--spin-days", type=int, default=10)
    ap.add_argument("--snapshots", type=int, default=1000)
    ap.add_argument("--holdout", type=int, default=48)
    ap.add_argument("--iters", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--check-every", type=int, default=USE["check_every"])
    ap.add_argument("--time-steps", type=int, default=240)
    ap.add_argument("--augment", type=float, default=0.5, help="fraction of recorded steps with a perturbed snapshot")
    ap.add_argument("--amp", type=float, default=1.0,
                    help="largest perturbation, in units of one step's change of ch and velocity")
    ap.add_argument("--onpolicy-rounds", type=int, default=3)
    ap.add_argument("--onpolicy-steps", type=int, default=240)
    ap.add_argument("--onpolicy-iters", type=int, default=2000)
    ap.add_argument("--onpolicy-lr", type=float, default=3e-4)
    ap.add_argument("--closed-loop-days", type=int, default=0)
    ap.add_argument("--tol", type=float, default=USE["tol"], help="check bound (pcom_nn_use)")
    ap.add_argument("--resume", type=int, default=USE["resume"], help="plain steps after a failed check (pcom_nn_resume)")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    cfg = configs.CONFIGS[args.config]()
    m, rep = prepare(cfg, abi.FORMS[args.form], spin_days=args.spin_days, snapshots=args.snapshots,
                     holdout=args.holdout, iters=args.iters, batch=args.batch, lr=args.lr, augment=args.augment,
                     amp=args.amp, onpolicy_rounds=args.onpolicy_rounds, onpolicy_steps=args.onpolicy_steps,
                     onpolicy_iters=args.onpolicy_iters, onpolicy_lr=args.onpolicy_lr)
    if args.closed_loop_days > 0:
        rep["closed_loop"] = closed_loop(m, cfg, rep["steps"], args.closed_loop_days, args.check_every, args.tol,
                                         args.resume)
    plain, sur = time_steps(m, args.time_steps, args.check_every, args.tol, args.resume)
    st = m.nn_stats()
    rep.update({"form": args.form, "check_every": args.check_every, "tol": args.tol, "resume": args.resume,
                "ms_per_step_plain": plain,
                "ms_per_step_surrogate": sur, "speedup": plain / sur,
                "last_check_rmse": list(st.last_check_rmse), "last_persist_rmse": list(st.last_persist_rmse),
                "finite": bool(m.check_finite())})
    print(json.dumps(rep), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(rep, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
