"""CPU tests of the FOM engine's option validation (shaw_fom_create): every
shape rule refuses an invalid request with its reason before the engine's
first HIP call, so these run without a GPU. A valid request gets past the
validation: without a GPU it then fails at the first HIP call; with one it is
created and closed."""
import pytest

import shaw_amd
from tests.test_gpu_fom import LAYERED, random_case

VALIDATION = ("col_block", "cols_per_thread", "strip", "prefetch_rows", "steps_per_launch", "graph_steps",
              "LDS", "forcing column count", "steps per launch", "bad grid")


@pytest.fixture(scope="module")
def app():
    return shaw_amd.App.structured(64, 256, LAYERED)


def create(app, m, **kw):
    """None if the options pass validation, else the library's message."""
    try:
        f = shaw_amd.Fom(app, m, 0.5, **kw)
    except shaw_amd.ShawError as e:
        msg = str(e)
        return None if msg.startswith("[2]") else msg  # [2]: a HIP error after validation
    f.close()
    return None


@pytest.mark.parametrize("m,kw,reason", [
    (0, {}, "forcing column count out of range"),
    (4, dict(col_block=3), "col_block must be a power of two"),
    (4, dict(col_block=64), "col_block must be a power of two"),  # 64 FP64 columns = 512 B
    (4, dict(cols_per_thread=3), "cols_per_thread must be 1, 2, 4 or 8"),
    (4, dict(col_block=2, cols_per_thread=4), "cols_per_thread must divide the column block width"),
    (32, dict(col_block=8, cols_per_thread=1, strip=128), "strip too wide for a 512-thread block"),
    (4, dict(prefetch_rows=3), "prefetch_rows must be 1 or 2"),
    (4, dict(steps_per_launch=8), "steps_per_launch must be 1 to 7"),
    (4, dict(steps_per_launch=4, strip=6), "k steps per launch need a strip of at least 2k points"),
    (32, dict(col_block=8, steps_per_launch=5), "5 to 7 steps per launch need one thread per point"),
    (4, dict(steps_per_launch=6, strip=32), "5 to 7 steps per launch need one thread per point"),
    (4, dict(steps_per_launch=3, graph_steps=8), "graph_steps must be a multiple of 2 x steps per launch"),
    (32, dict(col_block=32, cols_per_thread=8, strip=128, steps_per_launch=4), "tile needs more LDS"),
])
def test_invalid_shapes_are_refused(app, m, kw, reason):
    msg = create(app, m, kernel=shaw_amd.KERNEL_FUSED, **kw)
    assert msg is not None and reason in msg, msg


@pytest.mark.parametrize("m,kw", [
    (32, {}), (1, {}), (32, dict(col_block=8, steps_per_launch=4)), (40, dict(steps_per_launch=6, col_block=1)),
    (32, dict(precision=shaw_amd.FP32)), (32, dict(kernel=shaw_amd.KERNEL_SPLIT, cols_per_thread=2)),
    (4, dict(steps_per_launch=4, prefetch_rows=2, col_block=2)), (7, dict(steps_per_launch=4, strip=9, segment=1)),
])
def test_valid_shapes_pass(app, m, kw):
    assert create(app, m, **kw) is None


@pytest.mark.parametrize("seed", list(range(48)) + list(range(100, 124)))
def test_random_test_shapes_are_valid(seed):
    """Every shape the random GPU parity tests draw passes the validation, in
    FP64 and in FP32."""
    (nr, nth), src, rank2, rec, every, kw = random_case(seed)
    app = shaw_amd.App.structured(nr, nth, LAYERED)
    kw = dict(kw)
    r1 = kw.pop("rank1_forcing", not rank2)
    for prec in (shaw_amd.FP64, shaw_amd.FP32):
        assert create(app, len(src), precision=prec, rank1_forcing=r1, **kw) is None, (seed, kw)
