"""The BAL runner (examples/bal.rs:93-281) end to end on the GPU backends: a BAL file written
by the native writer, parsed back, optimized in each backend mode, checked by the independent
objective inside the runner, and reported with the reference's CSV columns."""
import io

import numpy as np
import pytest

from fagra_amd import bal, synthetic
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["schur", "scaled", "block", "identity"])
def test_runner_modes_reach_the_oracle_cost(tmp_path, mode):
    p = synthetic.generate("ring16", seed=21, n_cameras=8, n_points=300)
    path = tmp_path / "problem.txt"
    bal.write(str(path), p)
    out, err = io.StringIO(), io.StringIO()
    rows = bal.run(str(path), steps=30, inner=5000, repeats=2, mode=mode, tolerance=1e-10, out=out, err=err)
    lines = out.getvalue().splitlines()
    assert lines[0] == bal.CSV_HEADER and len(lines) == 3 and rows == lines[1:]
    head = bal.CSV_HEADER.split(",")
    runs = [dict(zip(head, r.split(","))) for r in rows]
    ref, _ = oracle.BA(p).lm(options=oracle.OptimizeOptions(30, 1e-8, 0.0, 0.0), tol=1e-10)
    for r in runs:
        assert len(r) == len(head)
        final, initial = float(r["final_cost"]), float(r["initial_cost"])
        assert final < initial and int(r["accepted"]) > 0
        assert abs(final - ref.cost) <= 1e-6 * ref.cost, (mode, final, ref.cost)
        assert abs(float(r["rmse"]) - np.sqrt(2 * final / p.n_observations)) < 1e-7
        assert int(r["solves"]) >= int(r["accepted"]) and int(r["inner_iterations"]) > 0
    assert runs[0]["final_cost"] == runs[1]["final_cost"]  # repeats are deterministic
    assert "cameras=8 points=300" in err.getvalue()


def test_runner_rejects_bad_controls(tmp_path):
    p = synthetic.generate("ring16", seed=21, n_cameras=4, n_points=50)
    path = tmp_path / "problem.txt"
    bal.write(str(path), p)
    for kw in [dict(mode="cg"), dict(tolerance=1.0), dict(steps=0)]:
        with pytest.raises(ValueError):
            bal.run(str(path), out=io.StringIO(), err=io.StringIO(), **kw)
    assert bal.main([str(path), "2", "10", "1", "bogus"]) == 1
    assert bal.main([]) == 2
