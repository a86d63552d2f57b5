"""Benchmark: simulated years per day (SYPD) of the PCOM time stepper on one
MI355X at 1 degree (360 x 180 x 30, dtbt = 60 s, dtbc = 3600 s).

Prints ONE JSON line.  `value` is SYPD of the reported form (by default
mixed+surrogate: FP32 tendency stencils with FP64 state, and the learned
barotropic surrogate in place of most subcycles) over exactly `--steps`
baroclinic steps (each = daily forcing when due, EOS, pressure gradient
integrals, advection/friction, the 120-substep barotropic cycle or the
network, baroclinic momentum, SSH, tracers, convection), state resident in
HBM.  The surrogate form is first fitted untimed (pcom_amd.surrogate.fit:
10 days of spin-up, 1000 recorded steps with perturbed snapshots beside
half of them, 6000 training iterations) and then used under the policy
pcom_amd.surrogate.USE.  Alongside: the plain FP64 form and the other
mixed forms, a per-phase breakdown from HIP events on the library's
stream, the roofline of the dominant kernel (the barotropic subcycle,
k_trop_tiles: FP64 flops per launch over its event-timed duration), the
HBM roofline of the whole step, each form's RMSE against FP64 after
--rmse-days days, and the CPU FP64 oracle timed on a bounded sample on
the host.

    python bench.py [--steps K] [--warmup W] [--form mixed+surrogate|fp64|mixed|mixed_eos32]
    python bench.py --calibrate-copy    # a plain 1 GiB copy, for PMC scaling
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
PEAK_FP64_TFLOPS = 78.6
YEAR_S = 31536000.0
# FP64 operations (+ - * / sqrt) of one barotropic substep at one point, as
# the reference's int_trop statements count them (mod_int.f90:184-273):
# predictor sch and fluxes 9, predictor div and ch 22, viscosity fluxes,
# sch and gradients 23, momentum with pressure gradient, Coriolis and
# corrector fluxes 104, corrector div, ch and accumulators 26.
TROP_FLOPS_PER_POINT = 184
# HBM bytes per k_trop_tiles launch at 1 degree, from two rocprofv3 --pmc
# passes over this bench (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, the
# factors being what the counters read for a plain 1 GiB copy); the passes
# and their commands are in the profile named here
TROP_TRAFFIC_BYTES = 403_400_000
TROP_TRAFFIC_SOURCE = "profiles/k_trop_tiles_global1deg_pmc.md (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated)"


def sypd(ms_per_step, dtbc):
    return (dtbc / YEAR_S) / (ms_per_step / 1000.0 / 86400.0)


def step_bytes(cfg, s3):
    """Algorithmic HBM bytes per baroclinic step: 31 compulsory 3-D field
    transfers (T, S, u, v read+write; AB3 history 6; am 2; kh 1; prho 2;
    six accumulators read+write) at s3 bytes, plus 50 FP64 2-D transfers."""
    n3 = cfg.glo_ni * cfg.glo_nj * cfg.nk
    n2 = cfg.glo_ni * cfg.glo_nj
    return n3 * 31 * s3 + n2 * 50 * 8


def _nn_counts(m):
    st = m.nn_stats()
    return st, {"surrogate_steps": st.surrogate_steps, "plain_steps": st.plain_steps, "checks": st.checks,
                "fallbacks": st.fallbacks}


FITTED = {}  # the surrogate form's trained parameters and fit steps, for accuracy()


def run_form(cfg, precision, steps, warmup, surrogate=False):
    from pcom_amd import abi
    from pcom_amd import surrogate as sm
    from pcom_amd.lib import Model
    m = Model(cfg, precision=precision)
    sur = None
    if surrogate:
        # fitted untimed, then used under the policy USE
        t0 = time.perf_counter()
        rep = sm.fit(m, cfg, log=lambda s: None)
        sur = {"fit_s": round(time.perf_counter() - t0, 2), "fit_steps": rep["steps"], "recorded": rep["recorded"],
               "train_iters": rep["iters"], "train_s": round(rep["train_s"], 2),
               "holdout_gain": {k: round(v, 3) for k, v in rep["holdout_gain"].items()}, "policy": sm.USE}
        FITTED[precision] = (m.nn_params(), rep["steps"])
        # the model time of the timed steps: the fit's steps and the warm-up
        sur["timed_from_step"] = rep["steps"] + warmup
        m.nn_use(**sm.USE)
    m.step(warmup)
    if sur is not None:
        _, before = _nn_counts(m)
    m.sync()
    t0 = time.perf_counter()
    m.step(steps)
    m.sync()
    wall = time.perf_counter() - t0
    if sur is not None:
        _, after = _nn_counts(m)
        sur["timed"] = {k: after[k] - before[k] for k in after}
    # per-phase device time from HIP events on the library's stream
    # (phase 2 brackets the subcycle kernel alone)
    m.phase_times(True)
    m.step(min(steps, 48))
    ph, n = m.phase_times(False)
    ok = m.check_finite()
    sums = m.global_sums()
    mode, tile = m.subcycle_mode(), m.tiling()
    nstep = 2 * cfg.dtbc // cfg.dtbt
    out = {"ms_per_step": 1000.0 * wall / steps,
           "phase_ms": {"step": ph[0], "other_kernels": ph[1], "subcycle_kernel": ph[2]},
           "phase_steps": n, "finite": bool(ok), "heat_salt_mass": [float(x) for x in sums],
           "subcycle_mode": mode, "subcycle_tiling": list(tile),
           "subcycle_launches": 1 if mode == abi.PCOM_TROP_PERSISTENT else -(-nstep // tile[2])}
    if sur is not None:
        st, _ = _nn_counts(m)
        sur.update({"in_use_at_end": bool(st.in_use),
                    "last_check_rmse": [float(f"{x:.4g}") for x in st.last_check_rmse],
                    "last_persistence_rmse": [float(f"{x:.4g}") for x in st.last_persist_rmse]})
        out["surrogate"] = sur
    m.close()
    return out


def time_plain_window(cfg, precision, start, steps):
    """ms per step of a plain form over steps start+1 .. start+steps of a
    run (the surrogate form is timed after its fit, later in model time,
    where the 3-D kernels cost more than in the first days)."""
    from pcom_amd.lib import Model
    m = Model(cfg, precision=precision)
    m.step(start)
    m.sync()
    t0 = time.perf_counter()
    m.step(steps)
    m.sync()
    wall = time.perf_counter() - t0
    m.close()
    return 1000.0 * wall / steps


def calibrate_copy():
    """A plain device copy of 1 GiB (reads 1 GiB, writes 1 GiB), run a few
    times: the reference kernel for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch
    a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes_read": a.numel() * 8, "calibrate_copy_bytes_written": b.numel() * 8}))


def accuracy(cfg, forms, days):
    """RMSE of SST, SSH and every prognostic field of each form against the
    FP64 form run side by side from the same state (the FP64 form is
    bit-identical to the FP64 oracle): the plain forms after `days`
    simulated days from the initial state; a surrogate form after `days`
    days of use, counted from the whole day after its fit (the FP64 run
    takes the same steps)."""
    from pcom_amd import abi, rmse
    from pcom_amd import surrogate as sm
    from pcom_amd.lib import Model
    spd = 86400 // cfg.dtbc
    res = {"days": days, "gates": {"sst_K": 1e-3, "ssh_m": 1e-3}, "forms": {}}
    ref = None
    for name in forms:
        if name == "fp64":
            continue
        base, sur = name.split("+")[0], name.endswith("+surrogate")
        m = Model(cfg, precision=abi.FORMS[base])
        r = {}
        if sur:
            # the fitted network of the timed run (recording leaves the state
            # as plain steps leave it, so plain steps reach the same state)
            if abi.FORMS[base] in FITTED:
                params, fit_steps = FITTED[abi.FORMS[base]]
                d0 = -(-fit_steps // spd)
                m.step(d0 * spd)
                m.nn_create(capacity=0, seed=1)
                m.nn_params(params)
            else:
                fit_steps = sm.fit(m, cfg, log=lambda s: None)["steps"]
                d0 = -(-fit_steps // spd)
                m.step(d0 * spd - fit_steps)
            m.nn_use(**sm.USE)
            m.step(days * spd)
            ref_s = Model(cfg, precision=abi.PCOM_FP64)
            ref_s.step((d0 + days) * spd)
            r = {k: float(f"{v:.4g}") for k, v in rmse.rmse_report(m, ref_s, cfg).items()}
            ref_s.close()
            _, cnt = _nn_counts(m)
            r.update({"in_use_from_day": d0, "to_day": d0 + days, **cnt})
        else:
            if ref is None:
                ref = Model(cfg, precision=abi.PCOM_FP64)
                ref.step(days * spd)
            m.step(days * spd)
            r = {k: float(f"{v:.4g}") for k, v in rmse.rmse_report(m, ref, cfg).items()}
        r["finite"] = bool(m.check_finite())
        res["forms"][name] = r
        m.close()
    if ref is not None:
        ref.close()
    return res


def cpu_baseline(cfg, nsteps):
    """The FP64 C oracle (the reference's algorithm restated, bit-identical
    to the FP64 GPU form) on the host's cores: its independent loops run
    under OpenMP with OMP_NUM_THREADS threads (all cores by default)."""
    from oracle.oracle import Oracle, threads
    o = Oracle(cfg)
    o.step(1)   # first step includes the daily forcing; time the steady ones
    t0 = time.perf_counter()
    o.step(nsteps)
    wall = time.perf_counter() - t0
    ms = 1000.0 * wall / nsteps
    nt = threads()
    return {"value": round(sypd(ms, cfg.dtbc), 4), "unit": "SYPD", "cores": nt, "kind": "port",
            "host_cpus": os.cpu_count(),
            "sample": f"{nsteps} steady steps of the same {cfg.glo_ni}x{cfg.glo_nj}x{cfg.nk} workload on the "
                      f"FP64 C oracle ({nt} OpenMP threads), {ms:.0f} ms/step"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=240)
    ap.add_argument("--warmup", type=int, default=24)
    ap.add_argument("--form", choices=["fp64", "mixed", "mixed_eos32", "mixed+surrogate"], default="mixed+surrogate",
                    help="form reported as `value`; the others are measured alongside")
    ap.add_argument("--cpu-steps", type=int, default=30, help="oracle steps timed for cpu_baseline (0 = skip)")
    ap.add_argument("--config", default="global", choices=["global", "box", "eddy"])
    ap.add_argument("--rmse-days", type=int, default=30,
                    help="simulated days of the side-by-side accuracy run of each form against FP64 (0 = skip)")
    ap.add_argument("--forms", default="fp64,mixed,mixed_eos32,mixed+surrogate",
                    help="forms to measure (comma-separated; must include --form)")
    ap.add_argument("--calibrate-copy", action="store_true", help=calibrate_copy.__doc__.split(":")[0])
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU (one MI355X holds the whole 1-degree domain); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    from pcom_amd import abi, configs
    cfg = configs.CONFIGS[args.config]()
    wanted = args.forms.split(",")
    if args.form not in wanted:
        wanted.insert(0, args.form)
    forms = {}
    for name in wanted:
        base, sur = name.split("+")[0], name.endswith("+surrogate")
        forms[name] = run_form(cfg, abi.FORMS[base], args.steps, args.warmup, surrogate=sur)
        forms[name]["sypd"] = round(sypd(forms[name]["ms_per_step"], cfg.dtbc), 2)
    rep = forms[args.form]
    # the 31 compulsory 3-D transfers; the mixed forms keep T, S, u, v in FP64
    s3 = 8 if args.form == "fp64" else (8 * 8 + 23 * 4) / 31
    b = step_bytes(cfg, s3)
    step_gbs = b / (rep["ms_per_step"] / 1000.0) / 1e9
    nstep = 2 * cfg.dtbc // cfg.dtbt
    flops = TROP_FLOPS_PER_POINT * cfg.glo_ni * cfg.glo_nj * nstep
    # the subcycle kernel timed where it runs every step (the surrogate form
    # runs it on its checks and holds only)
    kform = next((f for f in ("mixed", "fp64", "mixed_eos32") if f in forms), args.form)
    trop_ms = forms[kform]["phase_ms"]["subcycle_kernel"]
    trop_tflops = flops / (trop_ms / 1000.0) / 1e12
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
        "dtype": "f64" if args.form == "fp64" else "f32 stencils + f64 state"
                 + (" + bf16 surrogate" if "surrogate" in args.form else ""),
        "data": "synthetic: seeded analytic quasi-global configuration (no input files)",
        "config": {"workload": f"{cfg.name} {cfg.glo_ni}x{cfg.glo_nj}x{cfg.nk}, dtbt={cfg.dtbt} s, "
                               f"dtbc={cfg.dtbc} s, {nstep} barotropic substeps/step",
                   "form": args.form, "parallelism": "single GPU, whole domain"},
        "forms": forms,
        "plain_fp64_sypd": forms["fp64"]["sypd"] if "fp64" in forms else None,
        "surrogate_timed": rep.get("surrogate", {}).get("timed"),
        "accuracy_1yr": "1-year SST/SSH RMSE of each form against FP64: profiles/rmse_1yr_global1deg.md",
        "roofline": {"bound": "fp64-valu",
                     "limiter": "latency: five barrier-separated stages per substep on a cache-resident "
                                "2-D working set (waves wait for most cycles; see the PMC profile)",
                     "us_per_substep": round(1000.0 * trop_ms / nstep, 4),
                     "kernel": ("k_trop_tiles (barotropic subcycle, FP64, one persistent launch per step)"
                                if rep["subcycle_mode"] == abi.PCOM_TROP_PERSISTENT else
                                f"k_trop_block (barotropic subcycle, FP64, {rep['subcycle_launches']} launches "
                                "per step, timed together)"),
                     "achieved": round(trop_tflops, 4), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(trop_tflops / PEAK_FP64_TFLOPS, 5),
                     "traffic": TROP_TRAFFIC_BYTES if cfg.name == "global" else None,
                     "traffic_source": TROP_TRAFFIC_SOURCE if cfg.name == "global" else None,
                     "flops_per_launch": flops, "launch_ms": round(trop_ms, 4), "timed_in_form": kform,
                     "definition": f"{TROP_FLOPS_PER_POINT} FP64 flops per point per substep x "
                                   f"{cfg.glo_ni}x{cfg.glo_nj} points x {nstep} substeps / event-timed "
                                   "kernel duration; peak = MI355X FP64 (vector and matrix) 78.6 TFLOP/s"},
        "step_roofline": {"bound": "hbm", "achieved": round(step_gbs, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": round(step_gbs / PEAK_HBM_GBS, 5), "bytes_per_step": b,
                          "definition": "algorithmic bytes per baroclinic step (N3*31*s + N2*50*8) / wall time"},
    }
    # the surrogate form's timed steps come after its fit: the plain forms
    # over the same steps of a run, for a comparison in one window
    sur_name = next((f for f in forms if f.endswith("+surrogate")), None)
    if sur_name is not None:
        start = forms[sur_name]["surrogate"]["timed_from_step"]
        spd = 86400 // cfg.dtbc
        win = {"from_step": start, "days": [round(start / spd, 2), round((start + args.steps) / spd, 2)],
               "ms_per_step": {}, "sypd": {}}
        for f in ("fp64", sur_name.split("+")[0]):
            if f in forms:
                win["ms_per_step"][f] = round(time_plain_window(cfg, abi.FORMS[f], start, args.steps), 4)
        win["ms_per_step"][sur_name] = round(forms[sur_name]["ms_per_step"], 4)
        win["sypd"] = {f: round(sypd(ms, cfg.dtbc), 2) for f, ms in win["ms_per_step"].items()}
        if "fp64" in win["ms_per_step"]:
            win["surrogate_vs_fp64"] = round(win["ms_per_step"]["fp64"] / win["ms_per_step"][sur_name], 3)
        out["same_window"] = win
    if args.rmse_days > 0:
        out["accuracy"] = accuracy(cfg, wanted, args.rmse_days)
    if rank == 0 and args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_steps)
    if rank == 0:
        print(json.dumps(out))


if __name__ == "__main__":
    main()
