"""Benchmark: simulated years per day (SYPD) of the PCOM time stepper on one
MI355X at 1 degree (360 x 180 x 30, dtbt = 60 s, dtbc = 3600 s).

Prints ONE JSON line.  `value` is SYPD of the reported form over exactly
`--steps` baroclinic steps (each = daily forcing when due, EOS, pressure
gradient integrals, advection/friction, the 120-substep barotropic cycle,
baroclinic momentum, SSH, tracers, convection), state resident in HBM.
Alongside: the other precision form, a per-phase breakdown from HIP events
on the library's stream, the HBM roofline of the step, and the CPU FP64
oracle timed on a bounded sample on the host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SYPD at 1° 360x180x30 on 1 GPU, mixed+surrogate vs plain FP64, at stated RMSE"
PEAK_HBM_GBS = 8000.0
YEAR_S = 31536000.0


def sypd(ms_per_step, dtbc):
    return (dtbc / YEAR_S) / (ms_per_step / 1000.0 / 86400.0)


def step_bytes(cfg, s3):
    """Algorithmic HBM bytes per baroclinic step: 31 compulsory 3-D field
    transfers (T, S, u, v read+write; AB3 history 6; am 2; kh 1; prho 2;
    six accumulators read+write) at s3 bytes, plus 50 FP64 2-D transfers."""
    n3 = cfg.glo_ni * cfg.glo_nj * cfg.nk
    n2 = cfg.glo_ni * cfg.glo_nj
    return n3 * 31 * s3 + n2 * 50 * 8


def run_form(cfg, precision, steps, warmup):
    import ctypes as C

    from pcom_amd.lib import Model
    m = Model(cfg, precision=precision)
    m.step(warmup)
    m.sync()
    t0 = time.perf_counter()
    m.step(steps)
    m.sync()
    wall = time.perf_counter() - t0
    # per-phase device time from HIP events on the library's stream
    m.phase_times(True)
    m.step(min(steps, 20))
    ph, n = m.phase_times(False)
    ok = m.check_finite()
    sums = m.global_sums()
    m.close()
    return {"ms_per_step": 1000.0 * wall / steps, "phase_ms": {"step": ph[0], "3d": ph[1], "subcycle": ph[2]},
            "phase_steps": n, "finite": bool(ok), "heat_salt_mass": [float(x) for x in sums]}


def cpu_baseline(cfg, nsteps):
    from oracle.oracle import Oracle
    o = Oracle(cfg)
    o.step(1)   # first step includes the daily forcing; time the steady ones
    t0 = time.perf_counter()
    o.step(nsteps)
    wall = time.perf_counter() - t0
    ms = 1000.0 * wall / nsteps
    return {"value": round(sypd(ms, cfg.dtbc), 4), "unit": "SYPD", "cores": 1, "kind": "port",
            "sample": f"{nsteps} steady steps of the same 360x180x30 workload on the FP64 C oracle "
                      f"(1 thread), {ms:.0f} ms/step"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=240)
    ap.add_argument("--warmup", type=int, default=24)
    ap.add_argument("--form", choices=["fp64", "mixed"], default="fp64",
                    help="form reported as `value`; the other is measured alongside")
    ap.add_argument("--cpu-steps", type=int, default=4, help="oracle steps timed for cpu_baseline (0 = skip)")
    ap.add_argument("--config", default="global", choices=["global", "box"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU (one MI355X holds the whole 1-degree domain); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    from pcom_amd import abi, configs
    cfg = configs.CONFIGS[args.config]()
    forms = {}
    for name, prec in (("fp64", abi.PCOM_FP64), ("mixed", abi.PCOM_MIXED)):
        forms[name] = run_form(cfg, prec, args.steps, args.warmup)
        forms[name]["sypd"] = round(sypd(forms[name]["ms_per_step"], cfg.dtbc), 2)
    rep = forms[args.form]
    s3 = 8 if args.form == "fp64" else 4
    b = step_bytes(cfg, s3)
    achieved = b / (rep["ms_per_step"] / 1000.0) / 1e9
    out = {
        "metric": METRIC,
        "value": rep["sypd"],
        "unit": "SYPD",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(rep["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.form == "fp64" else "f32+f64",
        "data": "synthetic: seeded analytic quasi-global configuration (no input files)",
        "config": {"workload": f"{cfg.name} {cfg.glo_ni}x{cfg.glo_nj}x{cfg.nk}, dtbt={cfg.dtbt} s, "
                               f"dtbc={cfg.dtbc} s, 120 barotropic substeps/step",
                   "form": args.form, "parallelism": "single GPU, whole domain"},
        "forms": forms,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 5), "traffic": None,
                     "bytes_per_step": b,
                     "definition": "algorithmic bytes per baroclinic step (N3*31*s + N2*50*8) / wall time per step"},
    }
    if rank == 0 and args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_steps)
    if rank == 0:
        print(json.dumps(out))


if __name__ == "__main__":
    main()
