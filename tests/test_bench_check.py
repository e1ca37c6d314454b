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


def test_profiles_hold_every_traffic_figure_the_line_reads():
    """Each roofline `traffic` the bench line quotes comes from a committed
    profile summary; every (summary, key) it looks up is there and is at
    least the record's algorithmic bytes per launch."""
    algorithmic = {("stencil7_diffusion_256.md", "f32"): 256 ** 3 * 8,
                   ("stencil7_diffusion_256.md", "f64"): 256 ** 3 * 16,
                   ("weno5_512.md", "f32"): 512 ** 3 * 140,
                   ("weno5_512.md", "f64"): 512 ** 3 * 280,
                   ("jacobi_sweep_512.md", "f32"): 512 ** 3 * 12,
                   ("jacobi_sweep_512.md", "f64"): 512 ** 3 * 24,
                   ("batch_96.md", "batch_single_launch"): 2264686592,
                   ("batch_96.md", "batch_class_launches"): 2264686592}
    for (name, key), alg in algorithmic.items():
        t = bench.measured_traffic(name, key)
        assert t is not None, (name, key)
        assert alg <= t < 1.5 * alg, (name, key, t, alg)
