"""Ensemble workload: seeded members of uneven grid sizes stepped together.

The default ensemble has 24 members at 2 degrees (180 x 90 x 30) and 8 at
1 degree (360 x 180 x 30). All are on the seeded quasi-global geometry
(configs.global_1deg at each resolution, same bathymetry seed), in the
PCOM_MIXED form. Member m starts from the initial T plus N(0, 0.01 K) noise
with seed 100 + m.

Tile shapes come from an autotuner. For each grid size it collects a few
candidate subcycle tilings:
- the cost model's choice;
- the single-member fastest, by pcom_tune_subcycle;
- larger tiles, so that more members share one packed launch.

Every combination is timed in real ensemble steps on scratch members built
from the same inputs, and the fastest combination is applied to the
ensemble. Throughput is reported as aggregate simulated years per day over
all members, beside the same members stepped one after another.

    python -m pcom_amd.ensemble --days 365 [--n2 24 --n1 8] [--out report.json]
"""
import argparse
import itertools
import json
import sys
import time

import numpy as np

from . import abi, configs
from .lib import Ensemble, Model

YEAR_S = 31536000.0

# larger-tile options, to put more small members into one launch (a tile
# plus its halo of 2S points a side must fit one 1024-thread workgroup)
PACKING_TILES = [(26, 16, 2), (32, 12, 2), (22, 22, 2), (24, 24, 2), (20, 20, 3)]


def member_configs(n2=24, n1=8, seed0=100, sd=0.01):
    base = {2: configs.global_1deg(glo_ni=180, glo_nj=90), 1: configs.global_1deg()}
    out = []
    for m in range(n2 + n1):
        deg = 2 if m < n2 else 1
        b = base[deg]
        cfg = configs.Config(**{k: getattr(b, k) for k in b.__dataclass_fields__})
        rng = np.random.default_rng(seed0 + m)
        cfg.pt = np.asfortranarray(b.pt + sd * rng.standard_normal(b.pt.shape))
        cfg.name = f"member{m:02d}-{deg}deg"
        out.append((deg, cfg))
    return out


def member_spread(members, cfgs):
    """Per grid size, the area-weighted RMS over the sea of the members'
    SST standard deviation (K): how far the perturbed members have drifted
    apart."""
    out = {}
    for deg in sorted({d for d, _ in cfgs}):
        sst = np.stack([m.get("tc")[1:-1, 1:-1, 0] for (d, _), m in zip(cfgs, members) if d == deg])
        cfg = next(c for d, c in cfgs if d == deg)
        wet = cfg.itn > 0
        w = np.cos(np.deg2rad(cfg.lat))[None, :] * wet
        sd = sst.std(axis=0)
        out[f"{deg}deg"] = float(np.sqrt((w * sd ** 2).sum() / w.sum()))
    return out


def sypd(n_steps, dtbc, wall_s):
    return (n_steps * dtbc / YEAR_S) / (wall_s / 86400.0)


def time_ensemble(members, steps, warmup=1, chunk=0, log=None):
    ens = Ensemble(members)
    ens.step(warmup)
    for m in members:
        m.sync()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        n = min(chunk or steps, steps - done)
        ens.step(n)
        done += n
        if log and done < steps:
            members[0].sync()
            log(f"  {done} of {steps} ensemble steps, {time.perf_counter() - t0:.1f} s")
    for m in members:
        m.sync()
    wall = time.perf_counter() - t0
    npk, _ = ens.packs()
    ens.close()
    return wall, npk


def autotune(cfgs, precision, steps=3, log=print):
    """Choose one tiling per grid size by timing ensemble steps of scratch
    members; returns ({deg: tiling}, table of trials)."""
    scratch = [(deg, Model(cfg, precision=precision)) for deg, cfg in cfgs]
    degs = sorted({d for d, _ in scratch})
    cands = {}
    for d in degs:
        rep = next(m for dd, m in scratch if dd == d)
        default = rep.tiling()
        best, _ = rep.tune_subcycle(trials=6)
        opts = [default, best]
        for t in PACKING_TILES:
            try:
                rep.tiling(t)
                opts.append(tuple(t))
            except Exception:  # does not fit this grid or the device
                pass
        rep.tiling(default)
        cands[d] = list(dict.fromkeys(opts))
        log(f"  {d} degree candidates: {cands[d]}")
    table = []
    best = None
    for combo in itertools.product(*[cands[d] for d in degs]):
        choice = dict(zip(degs, combo))
        for d, m in scratch:
            m.tiling(choice[d])
        wall, npk = time_ensemble([m for _, m in scratch], steps)
        ms = 1000.0 * wall / steps
        table.append({"tiling": {str(d): list(t) for d, t in choice.items()}, "ms_per_step": ms, "launches": npk})
        log(f"  {choice}: {ms:.2f} ms per ensemble step, {npk} subcycle launches")
        if best is None or ms < best[0]:
            best = (ms, choice)
    for _, m in scratch:
        m.close()
    return best[1], table


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n2", type=int, default=24, help="members at 2 degrees")
    ap.add_argument("--n1", type=int, default=8, help="members at 1 degree")
    ap.add_argument("--form", default="mixed", choices=list(abi.FORMS))
    ap.add_argument("--steps", type=int, default=48, help="ensemble steps timed")
    ap.add_argument("--days", type=float, default=0, help="simulated days instead of --steps (365 = one year)")
    ap.add_argument("--compare-steps", type=int, default=48,
                    help="steps of the same members stepped one after another, for comparison")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    prec = abi.FORMS[args.form]
    cfgs = member_configs(args.n2, args.n1)
    dtbc = cfgs[0][1].dtbc
    if args.days > 0:
        args.steps = int(round(args.days * 86400 / dtbc))
    report = {"members": {"2deg": args.n2, "1deg": args.n1}, "form": args.form, "steps": args.steps,
              "simulated_days": args.steps * dtbc / 86400}
    choice = None
    if not args.no_tune:
        t0 = time.time()
        choice, table = autotune(cfgs, prec)
        report["autotune"] = {"choice": {str(d): list(t) for d, t in choice.items()}, "trials": table,
                              "seconds": time.time() - t0}
    members = [Model(cfg, precision=prec) for _, cfg in cfgs]
    if choice:
        for (deg, _), m in zip(cfgs, members):
            m.tiling(choice[deg])
    print(f"stepping {len(members)} members for {args.steps} steps", flush=True)
    wall, npk = time_ensemble(members, args.steps, chunk=500, log=lambda s: print(s, flush=True))
    sums = [m.global_sums() for m in members]
    report["ensemble"] = {"ms_per_step": 1000.0 * wall / args.steps, "wall_s": wall,
                          "subcycle_launches_per_step": npk,
                          "aggregate_sypd": len(members) * sypd(args.steps, dtbc, wall),
                          "finite": all(m.check_finite() for m in members),
                          "heat_salt_mass_first_member": [float(x) for x in sums[0]],
                          "sst_spread_k": member_spread(members, cfgs)}
    # the same members one after another (each pcom_step in turn)
    ns = args.compare_steps
    for m in members:
        m.sync()
    t0 = time.perf_counter()
    for _ in range(ns):
        for m in members:
            m.step(1)
    for m in members:
        m.sync()
    wall_seq = time.perf_counter() - t0
    report["one_after_another"] = {"steps": ns, "ms_per_step": 1000.0 * wall_seq / ns,
                                   "aggregate_sypd": len(members) * sypd(ns, dtbc, wall_seq)}
    report["speedup"] = (wall_seq / ns) / (wall / args.steps)
    print(json.dumps(report), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(report, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
