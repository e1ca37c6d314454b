"""bench.py prints a line only over results inside their stated tolerance
(result_check is pure host logic: no GPU needed)."""
import math

import bench


def test_result_check_passes_good_lines():
    assert bench.result_check({"fp32_rel_l2_vs_fp64": 3.3e-8}) == ""
    assert bench.result_check({"fp32_rel_l2_vs_fp64_worst_state": 3.7e-8}) == ""
    assert bench.result_check({"batch_vs_single_grid_max_rel": 0.0}) == ""
    assert bench.result_check({"convergence_32cube": {"refined_rel_residual": 9.96e-13,
                                                      "refined_vs_fp64_rel_l2": 2.3e-16}}) == ""
    assert bench.result_check({"metric": "copy"}) == ""


def test_result_check_rejects_bad_or_nan_lines():
    assert "diffusion" in bench.result_check({"fp32_rel_l2_vs_fp64": 3e-4})
    assert "diffusion" in bench.result_check({"fp32_rel_l2_vs_fp64": math.nan})
    assert "WENO5" in bench.result_check({"fp32_rel_l2_vs_fp64_worst_state": 2e-6})
    assert "batch" in bench.result_check({"batch_vs_single_grid_max_rel": 1.0})
    assert "Jacobi" in bench.result_check({"convergence_32cube": {"refined_rel_residual": 1e-11,
                                                                  "refined_vs_fp64_rel_l2": 0.0}})


def test_result_check_looks_into_every_workload_record():
    good = {"fp32_rel_l2_vs_fp64": 3.3e-8,
            "workloads": {"weno5": {"fp32_rel_l2_vs_fp64_worst_state": 3.7e-8},
                          "batch": {"batch_vs_single_grid_max_rel": 0.0}}}
    assert bench.result_check(good) == ""
    bad = {"fp32_rel_l2_vs_fp64": 3.3e-8,
           "workloads": {"weno5": {"fp32_rel_l2_vs_fp64_worst_state": 3.7e-8},
                         "jacobi": {"convergence_32cube": {"refined_rel_residual": 2e-12,
                                                           "refined_vs_fp64_rel_l2": 0.0}}}}
    assert bench.result_check(bad).startswith("jacobi: ")


def test_cpu_timed_protocol():
    calls = []
    r = bench.cpu_timed(lambda: calls.append(1), budget_s=1e-3)
    assert r["samples"] >= 10 and len(calls) == r["samples"] + 1   # one warm-up, then the samples
    assert 0.0 <= r["min_s"] <= r["mean_s"] and r["sdev_s"] >= 0.0
    assert bench.cpu_threads() >= 1
