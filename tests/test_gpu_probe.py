"""The device's test-and-set ceilings (ssw_measure_atomics), reported by the
benchmark beside the solvers' own rates: both probes run and give rates in a
sane range for one GPU."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["lds", "hbm"])
def test_atomic_ceiling(ctx, kind):
    ops = ctx.measure_atomics(kind)
    print(f"{kind} test-and-set ceiling {ops:.3e}/s", flush=True)
    assert 1e8 < ops < 1e14
