"""CPU checks of the benchmark's reporting: the printed JSON line stays compact and carries the
contract's fields, and the CPU baseline leg runs the oracle on a sample and projects it."""
import json
import types

import bench
from fagra_amd import synthetic

MODES = ["fp64_block_jacobi", "fp32_block_jacobi", "fp32_learned", "fp64_lsmr_block", "fp32_lsmr_block"]


def _summary():
    at = {"steps": 106, "attempts": 118, "linearizations": 106, "cost_evaluations": 119, "solves": 118,
          "iterations": 942}
    return {m: {"time_to_target_s": 0.1234567, "min_s": 0.12, "runs_at_target": 20, "runs": 20, "at_target": at,
                "lm_iterations_per_s": 955.9} for m in MODES}


def test_result_line_is_compact_and_complete():
    args = types.SimpleNamespace(steps=20, warmup=5, cg_tol=1e-3)
    cstar = {"cost": 1.2e7, "status": 3, "accepted_steps": 400, "cg_tolerance": 1e-10, "step_budget": 400,
             "rel_decrease_last_50_steps": 1e-9, "seconds": 12.3}
    roofline = {"bound": "hbm", "achieved": 2100.0, "peak": 8000.0, "unit": "GB/s", "frac": 0.26,
                "traffic": 8.1e7, "kernel": "k_points<float, kSchur>", "bytes_per_launch": 79150000,
                "avg_launch_us": 37.6, "launches_timed": 50, "traffic_source": bench.TRAFFIC_SOURCE,
                "schur_product": {"k_points_us": 37.6, "k_cameras_us": 33.2, "bytes": 152700000, "frac": 0.27,
                                  "stored_jacobian_design_bytes": 674000000}}
    rates = {"lm_iterations_per_s": 955.9, "schur_matvecs_per_s": {"fp32": 14000.0, "fp64": 9000.0},
             "linearizations_per_s": {"fp32": 6000.0, "fp64": 4000.0}}
    sample = {k: 1.2345 for k in ("linearize", "prepare_damping", "cost", "solve", "per_iteration",
                                  "solve_fixed", "first_lm_step")}
    cpu = {"value": 600.1, "unit": "s", "cores": 1, "kind": "port", "sample": "x" * 400,
           "single_thread": sample, "all_cores": {"value": 60.0, "threads": 16, **sample},
           "cpu": "AMD EPYC 9575F 64-Core Processor", "host_threads_allowed": 16}
    line = bench.result_line(args, "rings1500 seed 0: 1500 cameras, 500000 points, 3249262 observations, Huber 1.0",
                             "fp32_learned", "fp32", _summary(), cstar, rates, roofline, cpu)
    assert "\n" not in line and len(line) < 8192, len(line)
    d = json.loads(line)
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["higher_is_better"] is False and d["unit"] == "s" and d["n_gpus"] == 1
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in d["roofline"], k
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in d["cpu_baseline"], k
    assert d["speedup_vs_fp64"] == 1.0


def test_cpu_baseline_leg_times_the_oracle_and_projects():
    p = synthetic.generate("ring16", seed=3, n_cameras=8, n_points=400, huber=1.0)
    counts = {"linearizations": 10, "solves": 12, "iterations": 50, "cost_evaluations": 13}
    cpu = bench.cpu_baseline(p, counts, 1e-3)
    t = cpu["single_thread"]
    assert t["solve_iterations"] >= 1 and t["per_iteration"] > 0
    expect = (10 * (t["linearize"] + t["prepare_damping"]) + 12 * t["solve_fixed"] + 50 * t["per_iteration"]
              + 13 * t["cost"])
    assert abs(cpu["value"] - expect) <= 1e-4 * expect
    assert cpu["cores"] == 1 and cpu["kind"] == "port" and cpu["all_cores"]["threads"] >= 1
