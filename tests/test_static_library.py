"""The static archive libfagra_gpu.a and a C host linked against it.

CPU: the archive defines every symbol include/fagra_gpu.h declares, and the C driver
tests/c/lm_bal (tests/c/lm_bal.c, linked by fagra_amd/build.py) carries the library statically
rather than loading libfagra_gpu.so. GPU: the C driver reads a BAL file, runs fg_lm_solve_batch
and reports the same LM outcome, bit for bit, as the Python binding over the shared library on
the same parsed problem.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import fagra_amd as fa
from fagra_amd import bal, build, synthetic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "fagra_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(fg_[a-z0-9_]+)\s*\(", text)) - {"fg_on_solve_fn", "fg_on_accept_fn",
                                                               "fg_lsmr_observer_fn"}


def test_archive_defines_every_declared_symbol():
    out = subprocess.run(["nm", "-g", "--defined-only", build.ARCHIVE], capture_output=True, text=True,
                         check=True).stdout
    defined = set(re.findall(r" T (fg_[a-z0-9_]+)$", out, flags=re.M))
    missing = declared() - defined
    assert not missing, missing


def test_c_driver_links_the_archive_not_the_shared_library():
    assert os.path.exists(build.C_DRIVER_BIN), "run fagra_amd.build first"
    libs = subprocess.run(["ldd", build.C_DRIVER_BIN], capture_output=True, text=True).stdout
    assert "libfagra_gpu.so" not in libs
    syms = subprocess.run(["nm", build.C_DRIVER_BIN], capture_output=True, text=True).stdout
    assert re.search(r" T fg_lm_solve_batch$", syms, flags=re.M)


def _parse_line(line):
    return {k: float(v) for k, v in (kv.split("=") for kv in line.split())}


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [fa.FP64, fa.FP32])
def test_c_host_lm_matches_python_binding(tmp_path, precision):
    p = synthetic.generate("ring16", seed=9, n_cameras=10, n_points=500, huber=1.0)
    path = tmp_path / "problem.bal"
    path.write_text(p.to_bal_text())
    q = bal.parse(path.read_text())
    print(f"c host: {q.n_cameras} cameras, {q.n_points} points, {q.n_observations} observations, "
          f"precision {precision}", flush=True)
    res = subprocess.run([build.C_DRIVER_BIN, str(path), str(precision), "25", "1.0"], capture_output=True,
                         text=True, timeout=100)
    assert res.returncode == 0, res.stderr
    got = _parse_line(res.stdout.strip())
    ba = fa.BalProblem(q, precision=precision, huber=1.0)
    out = fa.LevenbergMarquardt(fa.Schur(precision=precision)).solve_batch(
        ba, fa.OptimizeOptions(25, 1e-8, 0.0, 0.0))
    cams, _ = ba.state()
    assert int(got["status"]) == out.status
    assert int(got["accepted"]) == out.accepted_steps and int(got["rejected"]) == out.rejected_steps
    assert got["initial"] == out.initial_cost and got["cost"] == out.cost
    assert got["camera0"] == cams[0, 0]
    assert out.accepted_steps > 0 and np.isfinite(out.cost)
