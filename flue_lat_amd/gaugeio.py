"""Gauge-field files on the host: openQCD and big-endian ILDG binary.

Fields are returned as complex128 numpy arrays of shape (NT, NX, NY, NZ, 4, 3, 3) indexed
[t, x, y, z, mu, a, b] = (U_mu(x))_ab with mu = 0 time, the reference's data layout
(src/wloops.F90:21), ready for GaugeField.from_complex.  Readers reunitarise every link
as the reference readers do (FixSU3Matrix, src/gaugeIO/ILDG_bin.f90:138-151).

openQCD (src/gaugeIO/openqcdfileio_standalone.f90:29-95), little endian:
    int32 NT, NX, NY, NZ; float64 plaq; then for every odd site (t + x + y + z odd, 0-based),
    in the order t slowest ... z fastest, the links U(x, d) and U(x - d, d) for
    d = t, x, y, z; each link 9 complex128 in row-major order.
ILDG binary (src/gaugeIO/ILDG_bin.f90:14-63), big endian, no header:
    sites with t slowest, then z, y, x fastest; per site the directions x, y, z, t; each
    link 9 complex128 in row-major order.
"""
import numpy as np


def fix_su3(U):
    """FixSU3Matrix on every 3x3 of U (last two axes): normalise row 1, orthogonalise and
    normalise row 2, row 3 = conj(row1 x row2) normalised (src/gaugeIO/ILDG_bin.f90:108-151)."""
    U = np.array(U, dtype=np.complex128, copy=True)
    v1 = U[..., 0, :]
    v1 /= np.sqrt(np.sum(np.abs(v1) ** 2, axis=-1, keepdims=True))
    v2 = U[..., 1, :]
    v2 -= v1 * np.sum(np.conj(v1) * v2, axis=-1, keepdims=True)
    v2 /= np.sqrt(np.sum(np.abs(v2) ** 2, axis=-1, keepdims=True))
    v3 = np.conj(np.cross(v1, v2))
    v3 /= np.sqrt(np.sum(np.abs(v3) ** 2, axis=-1, keepdims=True))
    U[..., 0, :], U[..., 1, :], U[..., 2, :] = v1, v2, v3
    return U


def _odd_sites(NT, NX, NY, NZ):
    t, x, y, z = np.meshgrid(np.arange(NT), np.arange(NX), np.arange(NY), np.arange(NZ), indexing="ij")
    odd = ((t + x + y + z) % 2) == 1
    return t[odd], x[odd], y[odd], z[odd]


def _check_dims(dims):
    dims = tuple(int(n) for n in dims)
    if len(dims) != 4 or min(dims) < 1:
        raise ValueError("lattice dims must be four positive extents (NT, NX, NY, NZ)")
    return dims


def read_openqcd_header(filename):
    """(NT, NX, NY, NZ), plaq from an openQCD configuration file."""
    with open(filename, "rb") as fh:
        head = fh.read(24)
    if len(head) != 24:
        raise ValueError(f"{filename}: too short for an openQCD header")
    dims = tuple(int(v) for v in np.frombuffer(head[:16], dtype="<i4"))
    plaq = float(np.frombuffer(head[16:], dtype="<f8")[0])
    return dims, plaq


def ReadGaugeField_OpenQCD(filename, NX=None, NY=None, NZ=None, NT=None):
    """Field of an openQCD file (src/gaugeIO/openqcdfileio_standalone.f90:29-95).  The
    extents come from the header; given extents must agree with it."""
    (NT_, NX_, NY_, NZ_), _ = read_openqcd_header(filename)
    for given, have in zip((NT, NX, NY, NZ), (NT_, NX_, NY_, NZ_)):
        if given is not None and int(given) != have:
            raise ValueError(f"{filename}: header extents {(NT_, NX_, NY_, NZ_)} differ from the requested ones")
    NT, NX, NY, NZ = _check_dims((NT_, NX_, NY_, NZ_))
    V = NT * NX * NY * NZ
    if NT % 2 or NX % 2 or NY % 2 or NZ % 2:
        raise ValueError("openQCD lattices have even extents (the file lists odd sites only)")
    raw = np.fromfile(filename, dtype="<c16", offset=24)
    if raw.size != (V // 2) * 8 * 9:
        raise ValueError(f"{filename}: expected {(V // 2) * 72} complex numbers, found {raw.size}")
    links = raw.reshape(V // 2, 4, 2, 3, 3)   # odd site, direction (t, x, y, z), forward/backward, a, b
    U = np.empty((NT, NX, NY, NZ, 4, 3, 3), np.complex128)
    t, x, y, z = _odd_sites(NT, NX, NY, NZ)
    ext = (NT, NX, NY, NZ)
    for mu in range(4):
        U[t, x, y, z, mu] = links[:, mu, 0]
        c = [t, x, y, z]
        c[mu] = (c[mu] - 1) % ext[mu]
        U[c[0], c[1], c[2], c[3], mu] = links[:, mu, 1]
    return fix_su3(U)


def WriteGaugeField_OpenQCD(filename, U, plaq):
    """Write U (NT, NX, NY, NZ, 4, 3, 3) in the openQCD layout with the given header
    plaquette value."""
    U = np.asarray(U, dtype=np.complex128)
    NT, NX, NY, NZ = _check_dims(U.shape[:4])
    if U.shape[4:] != (4, 3, 3):
        raise ValueError("U must have shape (NT, NX, NY, NZ, 4, 3, 3)")
    if NT % 2 or NX % 2 or NY % 2 or NZ % 2:
        raise ValueError("openQCD lattices have even extents (the file lists odd sites only)")
    t, x, y, z = _odd_sites(NT, NX, NY, NZ)
    ext = (NT, NX, NY, NZ)
    links = np.empty((t.size, 4, 2, 3, 3), np.complex128)
    for mu in range(4):
        links[:, mu, 0] = U[t, x, y, z, mu]
        c = [t, x, y, z]
        c[mu] = (c[mu] - 1) % ext[mu]
        links[:, mu, 1] = U[c[0], c[1], c[2], c[3], mu]
    with open(filename, "wb") as fh:
        fh.write(np.array([NT, NX, NY, NZ], dtype="<i4").tobytes())
        fh.write(np.array([plaq], dtype="<f8").tobytes())
        fh.write(links.astype("<c16").tobytes())


def ReadGaugeField_ILDG(filename, NX, NY, NZ, NT):
    """Field of a big-endian ILDG binary file (src/gaugeIO/ILDG_bin.f90:14-63)."""
    NT, NX, NY, NZ = _check_dims((NT, NX, NY, NZ))
    raw = np.fromfile(filename, dtype=">c16")
    if raw.size != NT * NX * NY * NZ * 36:
        raise ValueError(f"{filename}: expected {NT * NX * NY * NZ * 36} complex numbers, found {raw.size}")
    f = raw.reshape(NT, NZ, NY, NX, 4, 3, 3)          # file directions x, y, z, t
    f = f[:, :, :, :, [3, 0, 1, 2]]                     # -> t, x, y, z
    return fix_su3(np.transpose(f, (0, 3, 2, 1, 4, 5, 6)).astype(np.complex128))


def WriteGaugeField_ILDG(filename, U):
    """Write U (NT, NX, NY, NZ, 4, 3, 3) as big-endian ILDG binary data."""
    U = np.asarray(U, dtype=np.complex128)
    _check_dims(U.shape[:4])
    if U.shape[4:] != (4, 3, 3):
        raise ValueError("U must have shape (NT, NX, NY, NZ, 4, 3, 3)")
    f = np.transpose(U, (0, 3, 2, 1, 4, 5, 6))[:, :, :, :, [1, 2, 3, 0]]
    with open(filename, "wb") as fh:
        fh.write(np.ascontiguousarray(f).astype(">c16").tobytes())


def openqcd_plaquette_matches(filename, U=None, tol=1e-10):
    """Check an openQCD file's header plaquette against the field: returns (ok, measured,
    header).  The measurement is the average ReTr P over all 6V plaquettes, on the GPU;
    the header is accepted in either common normalisation (ReTr P or ReTr P / 3)."""
    from .field import GaugeField
    _, header = read_openqcd_header(filename)
    if U is None:
        U = ReadGaugeField_OpenQCD(filename)
    g = GaugeField.from_complex(U)
    s, n = g.plaquette(1, 4, 4)
    m = s / n
    ok = abs(m - header) < tol * max(1.0, abs(header)) or abs(m / 3 - header) < tol * max(1.0, abs(header))
    return ok, m, header
