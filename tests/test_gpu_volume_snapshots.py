"""Full 3-D snapshots taken without stalling the run (pw_set_volume_snapshots):
decoded in-stream after the step, copied to pinned host memory on a second
stream.  Each volume must equal the field read back synchronously after the
same number of steps on a second handle (bit for bit, FP32 and comp16), and its
y = ny/2 plane the slice snapshot of the same step.  Two snapshots on
consecutive steps and four in a short run reuse the two staging volumes, so
the ordering between the decode and the copies is exercised."""
import numpy as np
import pytest

import cases
import mixedprecisionfaracoustic_amd as pw

pytestmark = pytest.mark.gpu

STEPS = [20, 50, 51, 52, 99]
FIELDS = ["txx", "vux", "p", "vwz", "txx"]


@pytest.mark.parametrize("prec", ["fp32", "comp16"])
def test_volume_snapshots_equal_synchronous_fields(prec):
    grid, spec, nt = cases.layered_case(nt=100)
    s = cases.make_solver(grid, nt, prec, spec=spec)
    s.set_volume_snapshots(STEPS, FIELDS)
    s.set_snapshots(STEPS, FIELDS)
    with pytest.raises(pw.PorowaveError):
        s.volume_snapshot(0)  # its step has not run
    s.run(30)
    s.run(nt - 30)
    vols = [s.volume_snapshot(k) for k in range(len(STEPS))]
    slices = s.snapshots()
    s.close()
    ref = cases.make_solver(grid, nt, prec, spec=spec)
    done = 0
    for k, (st, f) in enumerate(zip(STEPS, FIELDS)):
        ref.run(st + 1 - done)
        done = st + 1
        want = ref.field(f)
        assert np.abs(want).max() > 0
        assert np.array_equal(vols[k], want), (prec, st, f)
        assert np.array_equal(vols[k][:, grid["ny"] // 2, :], slices[k]), (prec, st, f)
    ref.close()


def test_volume_snapshots_survive_reset_and_clear():
    grid, spec, nt = cases.layered_case(nt=60)
    s = cases.make_solver(grid, nt, "fp32", spec=spec)
    s.set_volume_snapshots([10, 59], ["txx", "txx"])
    s.run(nt)
    first = s.volume_snapshot(1)
    s.reset()
    with pytest.raises(pw.PorowaveError):
        s.volume_snapshot(1)
    s.run(nt)
    assert np.array_equal(s.volume_snapshot(1), first)
    s.set_volume_snapshots([], [])
    with pytest.raises(pw.PorowaveError):
        s.volume_snapshot(0)
    s.close()
