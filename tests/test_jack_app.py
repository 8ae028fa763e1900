"""Jackknife statistics (src/jack.f90) and the ensemble measurement driver (app/magnetic.f90)."""
import numpy as np
import pytest

import oracle as O
from flue_lat_amd import gaugeio as G
from flue_lat_amd import jack as J


def test_first_order_jackknife_is_standard_error():
    rng = np.random.default_rng(3)
    x = rng.normal(1.0, 0.3, size=17)
    c = np.concatenate([[J.Complement(x)], J.Complement(x, order=1)])
    err, bias = J.jackknife(c)
    assert abs(err - np.std(x, ddof=1) / np.sqrt(x.size)) < 1e-14
    assert abs(bias) < 1e-14 and abs(c[0] - x.mean()) < 1e-15
    # a nonlinear estimator picks up a bias; the closed form for f = mean^2
    f = c ** 2
    err2, bias2 = J.jackknife(f)
    assert abs(bias2 - (f[0] - f[1:].mean())) < 1e-15 and err2 > 0


@pytest.mark.parametrize("icon,jcon", [(0, None), (4, None), (2, 6)])
def test_higher_complements_by_brute_force(icon, jcon):
    rng = np.random.default_rng(8)
    x = rng.normal(size=(9, 3))                 # nine configurations of a 3-vector
    got = J.Complement(x, icon, jcon)
    for k in range(9):
        drop = [k, icon] + ([jcon] if jcon is not None else [])
        # the reference subtracts data(icon) and data itself even when they coincide
        s = x.sum(axis=0) - sum(x[d] for d in drop)
        assert np.allclose(got[k], s / (9 - len(drop)), rtol=0, atol=1e-14)


@pytest.mark.gpu
@pytest.mark.parametrize("app_compat", [False, True])
def test_magnetic_app_on_openqcd_files(tmp_path, capsys, app_compat):
    from flue_lat_amd import magnetic_app
    NT, NS = 4, 4
    fields = [O.random_field((NT, NS, NS, NS), 500 + k, 0.35) for k in range(4)]
    names = []
    for k, U in enumerate(fields):
        s, n = O.plaquette(U, 1, 4, 4)
        G.WriteGaugeField_OpenQCD(tmp_path / f"cfg{k}", U, s / n / 3)
        names.append(f"cfg{k}")
    (tmp_path / "list").write_text("\n".join(names) + "\n")
    args = ["--format", "openqcd", "--path", str(tmp_path), "--list", str(tmp_path / "list"),
            "--nt", str(NT), "--ns", str(NS)] + (["--app-compat"] if app_compat else [])
    assert magnetic_app.main(args) == 0
    lines = dict(l.split(" is ", 1) for l in capsys.readouterr().out.strip().splitlines())
    spatial = (1, 2, 4) if app_compat else (2, 4, 4)
    want = {"B": [O.magnetic(U) for U in fields],
            "aplaq": [np.divide(*O.plaquette(U, 1, 4, 4)) for U in fields],
            "splaq": [np.divide(*O.plaquette(U, *spatial)) for U in fields],
            "tplaq": [np.divide(*O.plaquette(U, 1, 1, 4)) for U in fields]}
    for key, vals in want.items():
        m, e = (float(v) for v in lines[key].split("+-"))
        assert abs(m - np.mean(vals)) < 1e-12 * abs(np.mean(vals))
        assert abs(e - np.std(vals, ddof=1) / 2.0) < 1e-9 * max(e, 1e-12)
