"""Gauge-field files: openQCD and ILDG readers/writers (host side).

The formats are pinned by writers in this file that follow the reference readers loop
by loop (src/gaugeIO/openqcdfileio_standalone.f90:46-91, src/gaugeIO/ILDG_bin.f90:26-56),
written independently of the vectorised code in flue_lat_amd.gaugeio, on tiny lattices.
"""
import numpy as np
import pytest

import oracle as O
from flue_lat_amd import gaugeio as G


def modc(a, b):
    return (a - 1) % b + 1


def write_openqcd_reference_order(path, U, plaq):
    """Byte stream in the order the reference reader consumes it."""
    NT, NX, NY, NZ = U.shape[:4]
    out = [np.array([NT, NX, NY, NZ], "<i4").tobytes(), np.array([plaq], "<f8").tobytes()]
    for it in range(1, NT + 1):
        for ix in range(1, NX + 1):
            for iy in range(1, NY + 1):
                for iz in range(1, NZ + 1):
                    if (ix + iy + iz + it - 4) % 2 == 0:
                        continue
                    for idd in range(1, 5):
                        mu = modc(idd - 1, 4)          # 4, 1, 2, 3 = t, x, y, z before the CSHIFT
                        dmu = [0, 0, 0, 0]
                        dmu[mu - 1] = 1
                        jx, jy = modc(ix - dmu[0], NX), modc(iy - dmu[1], NY)
                        jz, jt = modc(iz - dmu[2], NZ), modc(it - dmu[3], NT)
                        m = mu % 4                      # FLUE index after CSHIFT(-1): t=0, x=1, y=2, z=3
                        for (a, b, c, d) in ((it, ix, iy, iz), (jt, jx, jy, jz)):
                            UTmp = U[a - 1, b - 1, c - 1, d - 1, m].T   # the reader transposes what it reads
                            out.append(UTmp.flatten(order="F").astype("<c16").tobytes())
    with open(path, "wb") as fh:
        fh.write(b"".join(out))


def write_ildg_reference_order(path, U):
    NT, NX, NY, NZ = U.shape[:4]
    nu_of = {1: 2, 2: 3, 3: 4, 4: 1}                    # ILDG mu -> FLUE nu (1-based)
    URead = np.zeros((3, 3, 4, NX, NY, NZ, NT), np.complex128)
    for it in range(NT):
        for ix in range(NX):
            for iy in range(NY):
                for iz in range(NZ):
                    for mu in range(1, 5):
                        URead[:, :, mu - 1, ix, iy, iz, it] = U[it, ix, iy, iz, nu_of[mu] - 1].T
    with open(path, "wb") as fh:
        fh.write(URead.flatten(order="F").astype(">c16").tobytes())


@pytest.fixture(scope="module")
def field():
    return O.random_field((4, 2, 4, 2), 31, 0.4)


def test_openqcd_reads_reference_order(field, tmp_path):
    p = tmp_path / "cnfg"
    write_openqcd_reference_order(p, field, 0.625)
    (dims, plaq) = G.read_openqcd_header(p)
    assert dims == (4, 2, 4, 2) and plaq == 0.625
    U = G.ReadGaugeField_OpenQCD(p)
    assert U.shape == field.shape
    assert np.abs(U - field).max() < 1e-14


def test_openqcd_writer_matches_reference_order(field, tmp_path):
    a, b = tmp_path / "a", tmp_path / "b"
    write_openqcd_reference_order(a, field, 1.5)
    G.WriteGaugeField_OpenQCD(b, field, 1.5)
    assert a.read_bytes() == b.read_bytes()


def test_openqcd_rejects_mismatch(field, tmp_path):
    p = tmp_path / "c"
    G.WriteGaugeField_OpenQCD(p, field, 0.0)
    with pytest.raises(ValueError):
        G.ReadGaugeField_OpenQCD(p, NX=3)
    with pytest.raises(ValueError):
        G.WriteGaugeField_OpenQCD(p, O.random_field((4, 2, 3, 2), 1, 0.1), 0.0)
    p.write_bytes(p.read_bytes()[:-16])
    with pytest.raises(ValueError):
        G.ReadGaugeField_OpenQCD(p)


def test_ildg_reads_reference_order(field, tmp_path):
    p = tmp_path / "ildg"
    write_ildg_reference_order(p, field)
    NT, NX, NY, NZ = field.shape[:4]
    U = G.ReadGaugeField_ILDG(p, NX, NY, NZ, NT)
    assert np.abs(U - field).max() < 1e-14
    q = tmp_path / "ildg2"
    G.WriteGaugeField_ILDG(q, field)
    assert q.read_bytes() == p.read_bytes()
    with pytest.raises(ValueError):
        G.ReadGaugeField_ILDG(p, NX, NY, NZ, NT + 1)


def test_readers_reunitarise_like_fix_su3(tmp_path):
    rng = np.random.default_rng(5)
    U = rng.normal(size=(2, 2, 2, 2, 4, 3, 3)) + 1j * rng.normal(size=(2, 2, 2, 2, 4, 3, 3))
    want = np.empty_like(U)
    for idx in np.ndindex(U.shape[:5]):
        want[idx] = O.fix_su3(U[idx])
    assert np.abs(G.fix_su3(U) - want).max() < 1e-13
    p = tmp_path / "raw"
    G.WriteGaugeField_ILDG(p, U)
    assert np.abs(G.ReadGaugeField_ILDG(p, 2, 2, 2, 2) - want).max() < 1e-13
    q = tmp_path / "raw_oq"
    G.WriteGaugeField_OpenQCD(q, U, 0.0)
    assert np.abs(G.ReadGaugeField_OpenQCD(q) - want).max() < 1e-13


@pytest.mark.gpu
def test_openqcd_header_plaquette_check(tmp_path):
    U = O.random_field((4, 4, 4, 4), 9, 0.3)
    s, n = O.plaquette(U, 1, 4, 4)
    p = tmp_path / "cfg"
    G.WriteGaugeField_OpenQCD(p, U, s / n / 3.0)
    ok, m, h = G.openqcd_plaquette_matches(p)
    assert ok and abs(m / 3.0 - h) < 1e-12
    G.WriteGaugeField_OpenQCD(p, U, 0.5)
    assert not G.openqcd_plaquette_matches(p)[0]
