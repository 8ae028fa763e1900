"""Benchmark: Gsites/s of the 5-loop F_munu + B^2 (magnetic) measurement on a 48^3x96
lattice, FP32 links with FP64 sums versus the plain FP64 path, on one MI355X.

A step is one full measurement of B over the lattice in the variant the autotuner picks
on the benchmark field (the rectangle-first form at 48^3x96: the z-fastest copy of two
link directions, the tile kernels for the three spatial planes, their gathers with the
-ReTr F^2 sums, and the fixed-order second pass), with the links already resident in
HBM in the SoA layout.  The field is
synthetic: the seeded SU(3) generator with eps = 0.3, generated in FP64 on the GPU
and rounded to FP32 for the FP32 path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.  Beside the headline it carries the FP64 line, the roofline of the
magnetic measurement, event-timed rows of the other operations, the ensemble-flow workload
(32 configurations of 16^3 x Nt, 100 RK3 steps in one hipGraph) and, unless --no-cpu,
the CPU oracle on the same 48^3x96 field: its time is the CPU baseline and its B,
plaquette and Polyakov sums give each observable's error (rel_err_vs_oracle).
Only one GPU is used: --gpus N > 1 exits with a message.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DIMS = (96, 48, 48, 48)      # (NT, NX, NY, NZ)
SEED = 1003
EPS = 0.3
# Algorithmic work per site of the magnetic measurement: 258 SU(3) products of 198 flop
# (28 leaves x 3 products x 3 planes + 2 line products x 3 directions) and the
# compulsory read of the three spatial links.
FLOP_PER_SITE = 258 * 198
BYTES_PER_SITE = {"fp32": 3 * 72, "fp64": 3 * 144}
PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}   # MI355X vector peaks (FP32 counts packed FMA)
PEAK_HBM_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "magnetic_48x96_traffic.md")


def time_steps(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(steps):
        fn()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return ev0.elapsed_time(ev1) / 1e3 / steps, wall / steps


def timed(fn):
    t0 = time.perf_counter()
    fn()
    return time.perf_counter() - t0


def cpu_rows(U, threads):
    """The oracle on the other operations of the path, one pass each over bounded samples
    (per-site work of every operation is independent of the lattice size)."""
    import numpy as np

    import oracle as O
    V = U.shape[0] * U.shape[1] * U.shape[2] * U.shape[3]
    small = O.random_field((8, 24, 24, 24), SEED, EPS)
    Vs = 8 * 24 ** 3
    rows = {}
    for op, fn, n, lat in (
            ("plaquette", lambda: O.plaquette(U, 1, 4, 4), V, U.shape),
            ("polyakov", lambda: O.polyakov(U), V, U.shape),
            ("fmunu", lambda: O.loop5_munu(U, 2, 3), V, U.shape),
            ("wilson_loops_6x6", lambda: [O.wmunu(small, 2, 1, R, T) for R in range(1, 7) for T in range(1, 7)],
             Vs, small.shape),
            ("propagator", lambda: O.mom_scalar_d(np.fft.fftn(small, axes=(0, 1, 2, 3)), 1, 1.0), Vs, small.shape)):
        dt = timed(fn)
        rows[op] = {"gsites": n / dt / 1e9, "seconds": round(dt, 3),
                    "sample": "x".join(str(d) for d in lat[1:4]) + f"x{lat[0]}"}
    rows["threads"] = threads
    return rows


def gpu_rows(fields, torch):
    """Event-timed kernels of the other operations on the benchmark's fields (ms per call,
    mean of 5 after one warm-up, 3 for the B^2 adjoint; not part of the headline).  The
    autotuner first picks the plaquette walk length, the stencil visit order and the B^2
    adjoint form on each field (variants chosen are in `tuned`); the flow step runs on a
    copy, so the fields stay as measured."""
    V = DIMS[0] * DIMS[1] * DIMS[2] * DIMS[3]
    out = {}
    for name, g in fields:
        tuned = {op: g.autotune(op, reps=2)[0] for op in ("plaquette", "flow", "b2_grad")}
        flowed = g.clone()
        ops = {
            "plaquette": lambda: g.plaquette(1, 4, 4, sync=False),
            "polyakov": lambda: g.polyakov(sync=False),
            "fmunu": lambda: g.fmunu(2, 3),
            "wilson_loops_6x6": lambda: g.wilson_loops(2, 1, 6, 6),
            "propagator": lambda: g.gluon_propagator(mu_start=1),
            "wilson_grad": lambda: g.wilson_grad(6.0),
            "flow_step": lambda: flowed.wilson_flow(0.01, 1, use_graph=False),
            "b2_grad": lambda: g.magnetic_grad(),
        }
        r = {"tuned": tuned}
        for op, fn in ops.items():
            t_ev, _ = time_steps(fn, 3 if op == "b2_grad" else 5, 1, torch)
            r[op] = {"ms": round(t_ev * 1e3, 4), "gsites": round(V / t_ev / 1e9, 4)}
        del flowed
        torch.cuda.empty_cache()
        link = 72 if name == "fp32" else 144
        r["plaquette"]["hbm_frac"] = round(4 * link * V / (r["plaquette"]["ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)
        r["polyakov"]["hbm_frac"] = round(link * V / (r["polyakov"]["ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)
        out[name] = r
    return out


def oracle_leg(f64, torch):
    """The CPU oracle (FP64, OpenMP) on the whole benchmark workload: the 48^3x96 field the
    GPU measured, exported unchanged to the host.  Its B pass is the CPU baseline; its B,
    plaquette and Polyakov sums are what the GPU results are compared with."""
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    U = f64.to_complex().cpu().numpy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    B = O.magnetic(U)
    dt = time.perf_counter() - t0
    V = DIMS[0] * DIMS[1] * DIMS[2] * DIMS[3]
    ref = {"B": B, "plaquette": O.plaquette(U, 1, 4, 4)[0], "polyakov": O.polyakov(U)[0]}
    small = O.random_field((8,) + DIMS[1:], SEED, EPS)
    del U
    base = {"value": V / dt / 1e9, "unit": "Gsites/s", "cores": threads, "kind": "port",
            "sample": f"oracle magnetic (FP64, {threads} OpenMP threads) over the whole 48^3x96 lattice: "
                      f"{V} sites in {dt:.2f} s",
            "rows": cpu_rows(small, threads)}
    return base, ref


def ensemble_leg(FL, torch):
    """The batching workload: 32 configurations 16^3 x Nt, Nt drawn from {16, 24, 32, 48, 64,
    128} (seeded), links eps = 0.4 (seeds 4000 + j), 100 RK3 Wilson-flow steps at dt = 0.01
    with the plaquette and B of every configuration every 10 steps, all in one hipGraph.
    Timed around the whole call (capture, launch, and the copy of the measurements back)."""
    import numpy as np
    from flue_lat_amd.ensemble import Ensemble
    nts = np.random.default_rng(4000).choice([16, 24, 32, 48, 64, 128], size=32)
    dims = [(int(nt), 16, 16, 16) for nt in nts]
    seeds = [4000 + j for j in range(32)]
    sites = int(sum(np.prod(d) for d in dims))
    out = {"configs": 32, "nt": [int(v) for v in nts], "sites": sites, "steps": 100, "dt": 0.01,
           "measure_every": 10}
    for prec in ("fp32", "fp64"):
        e = Ensemble.random(dims, seeds, 0.4, prec)
        grid_variant = e.autotune(reps=2)[0]   # persistent grid, chosen on this ensemble
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plaq, B = e.flow(0.01, 100, measure_every=10, use_graph=True)
        dt = time.perf_counter() - t0
        out[prec] = {"s": round(dt, 4), "gsite_steps_per_s": round(sites * 100 / dt / 1e9, 4),
                     "grid_variant": grid_variant,
                     "mean_plaquette_final": float(np.mean(plaq[-1] / (18.0 * np.array([np.prod(d) for d in dims])))),
                     "mean_B_final": float(np.mean(B[-1]))}
        del e
        torch.cuda.empty_cache()
    return out


def hbm_traffic():
    """HBM-side bytes of one FP32 B measurement (all its dispatches), from the FETCH_SIZE /
    WRITE_SIZE counter summary kept with the profiles (None when that summary is absent)."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    m = re.search(r"^hbm_bytes_per_launch_fp32\s*=\s*(\d+)", open(TRAFFIC_FILE).read(), flags=re.M)
    return int(m.group(1)) if m else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the path runs on a single MI355X, so only --gpus 1 is measured",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import flue_lat_amd as FL

    torch.cuda.set_device(0)
    V = DIMS[0] * DIMS[1] * DIMS[2] * DIMS[3]
    f64 = FL.GaugeField.random(DIMS, SEED, EPS, "fp64")
    f32 = f64.astype("fp32")
    torch.cuda.synchronize()

    res = {}
    tuned = {}
    for name, g in (("fp32", f32), ("fp64", f64)):
        # the autotuner times the kernel variants on this field (outside the timed region);
        # g.magnetic() then runs the fastest
        tuned[name] = g.autotune("magnetic", reps=3)[0]
        holder = {}

        def step(g=g, holder=holder):
            holder["B"] = g.magnetic(sync=False)

        t_ev, t_wall = time_steps(step, a.steps, a.warmup, torch)
        B = holder["B"].item()
        res[name] = {"s_per_step": t_ev, "wall_s_per_step": t_wall, "B": B, "gsites": V / t_ev / 1e9}

    # FP64 links declared SU(3) (opt-in): the same field, two-row chains where they pay
    f64u = FL.GaugeField(DIMS, "fp64-su3", _data=f64.data)
    tuned["fp64_su3"] = f64u.autotune("magnetic", reps=3)[0]
    holder_u = {}

    def step_u():
        holder_u["B"] = f64u.magnetic(sync=False)

    t_u, _ = time_steps(step_u, a.steps, a.warmup, torch)
    res_su3 = {"gsites": V / t_u / 1e9, "s_per_step": t_u, "B": holder_u["B"].item()}

    # companion observables of the same configuration (plaquette, Polyakov), FP32 vs FP64
    extras = {}
    for name, g in (("fp32", f32), ("fp64", f64)):
        s, n = g.plaquette(1, 4, 4)
        p, npl = g.polyakov()
        extras[name] = {"plaquette": s / (3 * n), "polyakov": p / (3 * npl), "plaquette_sum": s, "polyakov_sum": p}
    rel_B = abs(res["fp32"]["B"] - res["fp64"]["B"]) / abs(res["fp64"]["B"])

    traffic = hbm_traffic()
    rows = gpu_rows((("fp32", f32), ("fp64", f64)), torch)

    t32 = res["fp32"]["s_per_step"]
    achieved_tflops = FLOP_PER_SITE * V / t32 / 1e12
    achieved_gbs = BYTES_PER_SITE["fp32"] * V / t32 / 1e9
    line = {
        "metric": "Gsites/s for 5-loop Fmunu+B^2 at 48^3x96, FP32-link vs FP64, % of HBM roofline",
        "value": round(res["fp32"]["gsites"], 6),
        "unit": "Gsites/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(t32 * 1e3, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 links, f64 sums",
        "data": "synthetic (seeded SU(3) field, eps=0.3, generated on the GPU)",
        "config": {"workload": "magnetic B (5-loop F_munu on the 3 spatial planes + sum of -ReTr F^2) on 48^3x96",
                   "lattice": "48^3x96", "sites": V, "eps": EPS, "seed": SEED},
        "fp64": {"value": round(res["fp64"]["gsites"], 6), "ms_per_step": round(res["fp64"]["s_per_step"] * 1e3, 6)},
        "fp32_over_fp64": round(res["fp32"]["gsites"] / res["fp64"]["gsites"], 4),
        "fp64_su3": {"value": round(res_su3["gsites"], 6), "ms_per_step": round(res_su3["s_per_step"] * 1e3, 6),
                     "rel_err_B_vs_fp64": abs(res_su3["B"] - res["fp64"]["B"]) / abs(res["fp64"]["B"]),
                     "note": "FP64 links declared SU(3) (opt-in precision); fp64 above makes no such assumption"},
        "kernel_variant": tuned,
        "rel_err_B_fp32_vs_fp64": rel_B,
        "B": {"fp32": res["fp32"]["B"], "fp64": res["fp64"]["B"]},
        "observables": extras,
        "rows": rows,
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved_tflops, 3),
            "peak": PEAK_TFLOPS["fp32"],
            "unit": "TFLOP/s",
            "frac": round(achieved_tflops / PEAK_TFLOPS["fp32"], 4),
            "traffic": traffic,
            "flop_per_site": FLOP_PER_SITE,
            # the canonical count (28 leaves x 3 products per plane); the rectangle-first form
            # (magnetic variant 7) executes fewer, so there frac is the fraction of the
            # canonical algorithm's VALU ceiling rather than of the VALU issue rate
            "flop_basis": "canonical",
            "variant": tuned["fp32"],
            # the code is built without packed FP32 instructions: their rate is half the peak
            "peak_unpacked": PEAK_TFLOPS["fp32"] / 2,
            "frac_of_unpacked": round(achieved_tflops / (PEAK_TFLOPS["fp32"] / 2), 4),
            "hbm": {"achieved": round(achieved_gbs, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(achieved_gbs / PEAK_HBM_GBS, 5), "bytes_per_site": BYTES_PER_SITE["fp32"]},
            "fp64": {"achieved": round(FLOP_PER_SITE * V / res["fp64"]["s_per_step"] / 1e12, 3),
                     "peak": PEAK_TFLOPS["fp64"],
                     "frac": round(FLOP_PER_SITE * V / res["fp64"]["s_per_step"] / 1e12 / PEAK_TFLOPS["fp64"], 4)},
        },
    }
    line["ensemble"] = ensemble_leg(FL, torch)
    if not a.no_cpu:
        base, ref = oracle_leg(f64, torch)
        line["cpu_baseline"] = base
        got = {"fp32": {"B": res["fp32"]["B"]}, "fp64": {"B": res["fp64"]["B"]}}
        for name in ("fp32", "fp64"):
            got[name]["plaquette"] = extras[name]["plaquette_sum"]
            got[name]["polyakov"] = extras[name]["polyakov_sum"]
        line["rel_err_vs_oracle"] = {name: {k: abs(got[name][k] - ref[k]) / abs(ref[k]) for k in ref}
                                     for name in got}
        line["oracle"] = ref
    print(json.dumps(line))


if __name__ == "__main__":
    main()
