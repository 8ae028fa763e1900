"""Profiling driver: runs one device operation (by default the magnetic 5-loop F_munu + B^2
measurement) a few times on a synthetic field and, for calibrating memory counters, one
plain device copy of known size.  Meant to sit under a profiler:

    rocprofv3 --kernel-trace --stats -d out -- python3 -m flue_lat_amd.prof
    rocprofv3 --pmc FETCH_SIZE -d out -- python3 -m flue_lat_amd.prof --prec fp32
    rocprofv3 --kernel-trace --stats -d out -- python3 -m flue_lat_amd.prof --op all --copy-mib 0
    rocprofv3 --kernel-trace --stats -d out -- python3 -m flue_lat_amd.prof --op flow --variant 7

--variant N runs variant N of the operation: the magnetic kernel form, the B^2 adjoint form,
the plaquette walk length, or the flow/gradient stencil (through a tuning table written for
the field's extents and precisions before the run).
"""
import argparse
import sys


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--dims", type=int, nargs=4, default=[96, 48, 48, 48], help="NT NX NY NZ")
    ap.add_argument("--prec", choices=["fp32", "fp64", "both"], default="both")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=1024, help="size of the calibration copy (0 = none)")
    ops = ["magnetic", "plaquette", "polyakov", "fmunu", "wloops", "propagator", "wilson_grad", "flow", "b2_grad"]
    ap.add_argument("--op", choices=ops + ["all"], default="magnetic")
    ap.add_argument("--variant", type=int, default=0, help="variant of the operation (0 = tuned or default)")
    a = ap.parse_args(argv)

    import os
    import tempfile

    import torch

    from . import _lib
    from .field import GaugeField, tune_load

    torch.cuda.set_device(0)
    tunable = {"plaquette": "plaquette", "flow": "flow", "wilson_grad": "flow", "b2_grad": "b2_grad"}
    if a.variant and a.op in tunable:
        code = _lib.OPS[tunable[a.op]]
        with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as t:
            for prec in (0, 1, 2):
                t.write(f"0 {code} {prec} {' '.join(str(d) for d in a.dims)} {a.variant} 1.0\n")
        tune_load(t.name)
        os.unlink(t.name)
    f64 = GaugeField.random(a.dims, 1003, 0.3, "fp64")
    fields = []
    if a.prec in ("fp32", "both"):
        fields.append(f64.astype("fp32"))
    if a.prec in ("fp64", "both"):
        fields.append(f64)

    def run(op, g):
        if op == "magnetic":
            return g.magnetic(variant=a.variant)
        if op == "plaquette":
            return g.plaquette()[0]
        if op == "polyakov":
            return g.polyakov()[0]
        if op == "fmunu":
            return float(g.fmunu(2, 3).abs().sum())
        if op == "wloops":
            return float(g.wilson_loops(2, 1, 6, 6).sum())
        if op == "propagator":
            return float(g.gluon_propagator(mu_start=1).real.sum())
        if op == "wilson_grad":
            return float(g.wilson_grad(6.0).data.abs().sum())
        if op == "b2_grad":
            return float(g.magnetic_grad().data.abs().sum())
        h = g.clone()
        h.wilson_flow(0.01, 1, use_graph=False)
        return h.plaquette()[0]

    for op in (ops if a.op == "all" else [a.op]):
        for g in fields:
            for _ in range(a.reps):
                v = run(op, g)
            print(f"{op} {g.precision} {tuple(a.dims)}: {v!r}", flush=True)
    if a.copy_mib > 0:
        n = a.copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copy {a.copy_mib} MiB done", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
