"""flue_magnetic_noalloc: the magnetic B with its partial sums in the library's static
device scratch instead of a workspace, for callers that cannot allocate.  Called through
the C ABI and checked against flue_magnetic_ex on the same field given a workspace of the
scratch's size (8 MiB): the kernels and the fixed-order reduction are the same, so the
results must agree bit for bit.  (The rectangle-first variant 7 needs more room than that
on most lattices and runs as the default form in both there; on 16^3 x 2 its footprints fit
and it runs in both.)  Two calls in a row, the second on a smaller lattice, show that the
shared scratch carries nothing from one call to the next.
"""
import ctypes

import pytest

import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

EWS = 10


@pytest.fixture(scope="module")
def FL():
    import flue_lat_amd as FL
    return FL


def noalloc(lib):
    return lib.flue_magnetic_noalloc


SCRATCH = 8 << 20


def run_both(FL, fn, g, variant):
    from flue_lat_amd._lib import check, lib
    out = torch.empty(1, dtype=torch.float64, device=g.data.device)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(fn(g._cdims, ctypes.c_void_p(g.data.data_ptr()), g.prec, variant, ctypes.c_void_p(out.data_ptr()), st),
          "flue_magnetic_noalloc")
    got = out.item()
    ws = torch.empty(SCRATCH, dtype=torch.uint8, device=g.data.device)
    check(lib().flue_magnetic_ex(g._cdims, ctypes.c_void_p(g.data.data_ptr()), g.prec, variant,
                                 ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()), SCRATCH, st),
          "flue_magnetic_ex")
    want = out.item()
    return got, want


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_matches_workspace_entry(FL, prec):
    from flue_lat_amd._lib import lib
    fn = noalloc(lib())
    # the larger lattice first, then a smaller one: the second call reads only what it wrote
    for dims, seed in (((24, 16, 12, 20), 61), ((6, 5, 7, 4), 62)):
        print(f"[noalloc] {prec} {dims}", flush=True)
        U = O.random_field(dims, seed, 0.3)
        g = FL.GaugeField.from_complex(U, prec)
        for variant in (0, 1, 4, 5, 7):
            got, want = run_both(FL, fn, g, variant)
            assert got == want, (dims, variant, got, want)
        assert run_both(FL, fn, g, 7)[0] == g.magnetic(variant=4 if prec == "fp64" else 5)
        tol = 1e-12 if prec == "fp64" else 1e-6
        assert abs(got - O.magnetic(U)) <= tol * O.magnetic(U)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_rectangle_form_where_it_fits(FL, prec):
    # extents divisible by 16 and footprints within the scratch (about 4.9 MB in FP64 at
    # 16^3 x 2): variant 7 runs the rectangle-first kernels on the scratch, then on a smaller
    # lattice
    from flue_lat_amd._lib import lib
    fn = noalloc(lib())
    for dims, seed in (((2, 16, 16, 16), 63), ((1, 16, 16, 16), 64)):
        print(f"[noalloc] rect {prec} {dims}", flush=True)
        U = O.random_field(dims, seed, 0.3)
        g = FL.GaugeField.from_complex(U, prec)
        assert lib().flue_magnetic_work_bytes(g._cdims, g.prec) <= SCRATCH
        got, want = run_both(FL, fn, g, 7)
        assert got == want, (dims, got, want)
        # the same kernels and fixed-order sums as with the field's own workspace
        assert got == g.magnetic(variant=7)
        tol = 1e-12 if prec == "fp64" else 1e-6
        assert abs(got - O.magnetic(U)) <= tol * O.magnetic(U)


def test_noalloc_benchmark_lattice(FL):
    from flue_lat_amd._lib import lib
    fn = noalloc(lib())
    g = FL.GaugeField.random((96, 48, 48, 48), 1003, 0.3).astype("fp32")
    for variant in (0, 7):
        got, want = run_both(FL, fn, g, variant)
        assert got == want, variant
    assert got == g.magnetic(variant=5)
    del g
    torch.cuda.empty_cache()


def test_noalloc_refuses_lattices_beyond_its_scratch(FL):
    from flue_lat_amd._lib import lib
    fn = noalloc(lib())
    # 64^4 x 3 sites: more partials than the 8 MiB scratch holds; refused before any launch
    dims = (ctypes.c_int32 * 4)(192, 64, 64, 64)
    assert fn(dims, None, 0, 0, None, None) == EWS
    dims = (ctypes.c_int32 * 4)(8, 8, 8, 8)
    assert fn(dims, None, 0, 0, None, None) == 2      # FLUE_ENULL for the missing field
    assert fn(dims, None, 0, 9, None, None) == 1      # FLUE_EINVAL for a variant out of range
