"""lbm_step_ex_noalloc (the allocation-free variant of lbm_step_ex) against
lbm_step_ex through the C ABI: the same schedule of plain, snapshot and
check steps on two handles gives bit-identical sums and populations, on a
domain and then on a smaller one (the library-wide scratch array keeps the
first call's contents), and a launch beyond the scratch capacity is
rejected."""
import numpy as np
import pytest

from lbm_amd import capi
from lbm_amd import workloads as W

pytestmark = pytest.mark.gpu

# plain, snapshot, check, check + snapshot in the order a host uses them
SCHEDULE = [0, 0, 1, 2, 0, 1, 3, 2, 0, 0, 1, 2]


def run(shape, noalloc, solid=None, faces=None):
    rho, u = W.taylor_green_noise(shape, 0.05, 1e-3, seed=5)
    if solid is not None:
        u[solid] = 0.0
    lat = capi.Lattice(shape, storage="f32", omega=1 / 0.6, solid=solid, faces=faces)
    lat.init(rho, u)
    step = lat.step_ex_noalloc if noalloc else lat.step_ex
    sums = []
    for mode in SCHEDULE:
        s, rc = step(mode)
        assert rc == 0
        if mode >= 2:
            sums.append(s.as_dict())
    f = lat.populations()
    lat.close()
    return sums, f


def sphere(shape):
    nx, ny, nz = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r2 = (x + 0.5 - nx / 3) ** 2 + (y + 0.5 - ny / 2) ** 2 + (z + 0.5 - nz / 2) ** 2
    solid = (r2 < (ny / 5) ** 2).reshape(-1)
    faces = {k: ("wall", None, 1.0) for k in ["-y", "+y", "-z", "+z"]}
    faces["-x"] = ("velocity_neem", (0.04, 0.0, 0.0), 1.0)
    faces["+x"] = ("pressure_neem", None, 1.0)
    return solid, faces


@pytest.mark.parametrize("shapes", [[(96, 40, 36), (40, 20, 12)]])
def test_noalloc_equals_step_ex_twice_second_smaller(shapes):
    for shape in shapes:
        print("noalloc periodic box", shape, flush=True)
        a_sums, a_f = run(shape, False)
        b_sums, b_f = run(shape, True)
        assert a_sums == b_sums
        assert np.array_equal(a_f, b_f)
        assert a_sums[-1]["fluid_cells"] == np.prod(shape)


def test_noalloc_equals_step_ex_with_solids_and_faces():
    for shape in [(80, 32, 28), (48, 24, 20)]:
        print("noalloc sphere channel", shape, flush=True)
        solid, faces = sphere(shape)
        a_sums, a_f = run(shape, False, solid, faces)
        b_sums, b_f = run(shape, True, solid, faces)
        assert a_sums == b_sums
        assert np.array_equal(a_f, b_f)


def test_noalloc_rejects_launch_beyond_capacity():
    # 64 x 1024 x 520 cells in 64x4 blocks: 256 * 520 = 133,120 blocks > 131,072
    shape = (64, 1024, 520)
    print("noalloc capacity", shape, flush=True)
    lat = capi.Lattice(shape, storage="f32", omega=1 / 0.6)
    with pytest.raises(capi.LBMError, match="LBM_NOALLOC_MAX_BLOCKS"):
        lat.step_ex_noalloc(2)
    s, rc = lat.step_ex_noalloc(0)  # a plain step needs no scratch
    assert rc == 0
    lat.close()
