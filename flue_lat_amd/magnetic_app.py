"""Ensemble measurement driver: plaquettes and the magnetic B with jackknife errors
(app/magnetic.f90).

    python -m flue_lat_amd.magnetic_app --format openqcd --path DIR --list cfgs.list --nt 8 --ns 32

Each configuration is read on the host (openQCD or ILDG binary), moved to the GPU and
measured there: the average plaquette over all planes, the spatial and the temporal
planes, and B.  The reference app's spatial average uses genPlaquette(1, 2, 4), which
also takes the (t,x), (t,y), (t,z) planes; --app-compat reproduces that, the default
uses the three spatial planes (2, 4, 4).
"""
import argparse
import os
import sys

import numpy as np

from . import gaugeio
from .jack import jackknife_mean_err


def measure_configs(files, fmt, NT, NS, precision="fp64", app_compat=False):
    """Per-configuration (aplaq, splaq, tplaq, B) arrays for the listed files."""
    from .field import GaugeField
    spatial = (1, 2, 4) if app_compat else (2, 4, 4)
    out = {k: [] for k in ("aplaq", "splaq", "tplaq", "B")}
    for path in files:
        if fmt == "openqcd":
            U = gaugeio.ReadGaugeField_OpenQCD(path, NS, NS, NS, NT)
        elif fmt in ("ildg", "cssmILDG"):
            U = gaugeio.ReadGaugeField_ILDG(path, NS, NS, NS, NT)
        else:
            raise ValueError(f"gauge format {fmt} is not supported")
        g = GaugeField.from_complex(U, precision)
        for key, args in (("aplaq", (1, 4, 4)), ("splaq", spatial), ("tplaq", (1, 1, 4))):
            s, n = g.plaquette(*args)
            out[key].append(s / n)
        out["B"].append(g.magnetic())
    return {k: np.array(v) for k, v in out.items()}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--format", default="openqcd", choices=["openqcd", "ildg", "cssmILDG"])
    ap.add_argument("--path", default=".", help="directory of the configurations")
    ap.add_argument("--list", required=True, help="text file, one configuration file name per line")
    ap.add_argument("--nt", type=int, required=True)
    ap.add_argument("--ns", type=int, required=True)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--app-compat", action="store_true", help="spatial plaquette as genPlaquette(1, 2, 4)")
    a = ap.parse_args(argv)
    with open(a.list) as fh:
        names = [ln.strip() for ln in fh if ln.strip()]
    files = [os.path.join(a.path, n) for n in names]
    r = measure_configs(files, a.format, a.nt, a.ns, a.precision, a.app_compat)
    for key, label in (("B", "B"), ("aplaq", "aplaq"), ("splaq", "splaq"), ("tplaq", "tplaq")):
        m, e = jackknife_mean_err(r[key])
        print(f"{label} is {m!r} +- {e!r}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
