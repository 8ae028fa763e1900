"""Writes observables.json: seeded fields from the generator (checksums of the links) and
their observables from the CPU oracle, for small versions of the workloads.

    python tests/golden/make_golden.py

The numbers are regression fixtures: the oracle is pinned by closed forms and invariants
(tests/test_oracle_*.py); these freeze the generator and the observables on fixed inputs.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

CASES = [
    # name, dims (NT, NX, NY, NZ), seed, eps
    ("c0_8x16", (16, 8, 8, 8), 1000, 0.25),
    ("c1_small_6x12", (12, 6, 6, 6), 1001, 0.35),
    ("c2_small_6x8", (8, 6, 6, 6), 1002, 0.3),
    ("uneven_6574", (6, 5, 7, 4), 1234, 0.35),
]


# position-space propagator case (NT <= NS)
PROP_CASE = ("prop_4x4", (4, 4, 4, 4), 1005, 0.3)


def checksums(U):
    return {"sum_re": float(U.real.sum()), "sum_im": float(U.imag.sum()),
            "sum_abs2": float((np.abs(U) ** 2).sum()),
            "weighted": float((U.real * ((np.arange(U.size) % 7) + 1).reshape(U.shape)).sum())}


def main():
    out = {}
    for name, dims, seed, eps in CASES:
        U = O.random_field(dims, seed, eps)
        e = {"dims": list(dims), "seed": seed, "eps": eps, "checksums": checksums(U)}
        for key, args in (("plaq_total", (1, 4, 4)), ("plaq_spatial", (2, 4, 4)), ("plaq_temporal", (1, 1, 4))):
            s, n = O.plaquette(U, *args)
            e[key] = [s, n]
        s, n = O.polyakov(U)
        e["polyakov"] = [s, n]
        e["magnetic"] = O.magnetic(U)
        e["wmunu"] = {f"{mu}{nu}_{R}x{T}": O.wmunu(U, mu, nu, R, T)
                      for (mu, nu) in ((2, 1), (3, 4)) for (R, T) in ((1, 1), (2, 3), (3, 2))}
        F = O.loop5_munu(U, 2, 3)
        e["loop5_23_site0"] = [[F[0, 0, 0, 0, a, b].real, F[0, 0, 0, 0, a, b].imag] for a in range(3) for b in range(3)]
        e["wilson_action_beta6"] = O.wilson_action(U, 6.0)
        # calc_mom_space_scalarD of the field's transform, over the loop range E//2+1
        Uq = np.fft.fftn(U, axes=(0, 1, 2, 3))
        for key, (mu_start, xi) in (("gluon_mom_d_mu1", (1, 1.0)), ("gluon_mom_d_mu0_xi15", (0, 1.5))):
            D = O.mom_scalar_d(Uq, mu_start, xi).real
            e[key] = {"sum": float(D.sum()), "q0000": float(D[0, 0, 0, 0]), "q1111": float(D[1, 1, 1, 1]),
                      "qmax": float(D[-1, -1, -1, -1])}
        out[name] = e
    # scalarGluonProp needs NT <= NS
    U = O.random_field(PROP_CASE[1], PROP_CASE[2], PROP_CASE[3])
    D = O.scalar_gluon_prop(U, 1.0)
    out["scalar_gluon_prop"] = {"dims": list(PROP_CASE[1]), "seed": PROP_CASE[2], "eps": PROP_CASE[3],
                                "sum": float(D.sum()), "d0000": float(D[0, 0, 0, 0]), "d1231": float(D[1, 2, 3, 1])}
    with open(os.path.join(HERE, "observables.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
