"""Extract the reference's published known-answer values into a small fixture.

Runs only where the reference checkout is mounted (it reads JSON result files
under /root/reference); the generated ``reference_results.json`` is committed
and is what the tests read.  No reference source code is read or copied: the
fixture holds numbers the reference's own validation scripts wrote.

    python tests/golden/make_golden.py
"""
import json
from pathlib import Path

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent / "reference_results.json"


def load(rel):
    return json.loads((REF / rel).read_text(encoding="utf-8"))


def main():
    fixture = {"source": "numbers published by the reference's validation scripts"}

    # fast_bp/validate.py -> results/synthetic_validation.json
    syn = load("fast_bp/results/synthetic_validation.json")
    fixture["synthetic_validation"] = {
        "file": "fast_bp/results/synthetic_validation.json",
        "point_tests": [
            {key: rec[key] for key in (
                "leaf_size", "relative_complex_error", "complex_coherence",
                "peak_position_errors_m", "leaves", "merges",
                "leaf_pixel_pulse_products", "merged_pixels", "shape")}
            for rec in syn["point_tests"]
        ],
        "lfm_peak_sample": syn["lfm_peak_sample"],
        "expected_sample": syn["expected_sample"],
    }

    # fast_bp/gpu_baseline.py:121-227 -> results/gpu_baseline/operator_validation.json
    op = load("fast_bp/results/gpu_baseline/operator_validation.json")
    cases = {}
    for c in op["cases"]:
        if c["mode"] != "fp64":
            continue
        cases[f"{c['k_sign']},{c['side']}"] = {
            "direct_peak": c["direct_vs_cpu"]["reference_peak_pixel"],
            "ffbp_peak": c["ffbp_vs_cpu"]["reference_peak_pixel"],
            "direct_phase_mask_samples": c["direct_vs_cpu"]["phase_mask_samples"],
            "ffbp_phase_mask_samples": c["ffbp_vs_cpu"]["phase_mask_samples"],
        }
    fixture["operator_validation"] = {
        "file": "fast_bp/results/gpu_baseline/operator_validation.json",
        "seed": 20260916, "order": ["-1,1", "1,-1"], "cases": cases,
        "gpu_vs_cpu_relative_error": {
            f"{c['k_sign']},{c['side']},{c['mode']}": [
                c["direct_vs_cpu"]["relative_complex_error"],
                c["ffbp_vs_cpu"]["relative_complex_error"]]
            for c in op["cases"]},
    }

    # fast_bp/resident_experiment.py:146-183 -> results/resident_20260917/validation.json
    val = load("fast_bp/results/resident_20260917/validation.json")
    cases = {}
    for c in val["cases"]:
        if c["mode"] != "fp64" or c["variant"] != "original":
            continue
        entry = cases.setdefault(f"{c['sign']},{c['side']}", {})
        entry[f"{c['algorithm']}_peak"] = c["quality"]["reference_peak_pixel"]
        entry["shape"] = c["metadata"]["shape"]
        if c["algorithm"] == "direct":
            entry["direct_kernel_counts_chunk7"] = c["metadata"]["kernel_counts"]
        else:
            entry["ffbp_kernel_counts"] = c["metadata"]["kernel_counts"]
    fixture["resident_validation"] = {
        "file": "fast_bp/results/resident_20260917/validation.json",
        "seed": 20260917, "order": ["-1,1", "-1,-1", "1,1", "1,-1"], "cases": cases,
    }

    # fast_bp/resident_followup.py:48-80 -> followup_validation.json
    fol = load("fast_bp/results/resident_20260917/followup_validation.json")
    cases = {}
    for c in fol["records"]:
        entry = cases.setdefault(f"{c['sign']},{c['side']}", {})
        if c["variant"] == "batched" and c["mode"] == "fp64":
            entry["ffbp_peak"] = c["quality"]["reference_peak_pixel"]
        if c["variant"] == "phase_lut":
            entry["direct_peak"] = c["quality"]["reference_peak_pixel"]
            entry["lut_vs_fp64_direct_error"] = c["quality"]["relative_complex_error"]
    fixture["followup_validation"] = {
        "file": "fast_bp/results/resident_20260917/followup_validation.json",
        "seed": 20260918, "order": ["-1,1", "-1,-1", "1,1", "1,-1"], "cases": cases,
    }

    # Metric self-checks (closed-form sinc).
    ms = load("sar_quality/results/metric_selfcheck.json")
    qs = load("vancouver_bp/results/quality_selfcheck.json")
    fixture["metric_selfcheck"] = {
        "impulse_metrics_sinc": {k: ms["sinc"][k] for k in ("irw_m", "pslr_db", "islr_db")},
        "profile_metrics_sinc": {k: qs["sinc"][k] for k in ("irw_m", "pslr_db", "islr_db")},
        "profile_independent_expectations": qs["independent_expectations"],
    }

    af = load("fast_bp/results/autofocus_validation.json")
    fixture["autofocus_validation"] = {
        k: af[k] for k in ("targets", "cutoff_hz", "initial_coordinate_rmse_m",
                           "corrected_coordinate_rmse_m", "delta_rms_m", "delta_max_m",
                           "median_gradient_coherence")}
    fixture["autofocus_validation"]["file"] = "fast_bp/results/autofocus_validation.json"

    OUT.write_text(json.dumps(fixture, indent=1) + "\n", encoding="utf-8")
    print("wrote", OUT)


if __name__ == "__main__":
    main()
