"""Compare two output files of the program (the .dc format: pc then qc,
float32, x fastest) the way the reference validates its backends
(validation/CompareResults.R:4-58): every float of the predicted file within
an absolute tolerance (default 1e-6) of the ground truth.

    python -m distributed-cube-average_amd.compare GROUND_TRUTH.dc PREDICTED.dc [TOLERANCE]

Prints the same report (min, max and standard deviation of the ground truth
on success; the count of differences above 1e-9, the largest and the mean
absolute difference otherwise) and exits 0 when the files agree, 1 when they
differ or their sizes differ.
"""
from __future__ import annotations

import sys
from dataclasses import dataclass

import numpy as np

DEFAULT_TOLERANCE = 1e-6


@dataclass
class Comparison:
    ok: bool
    same_size: bool
    report: str
    max_abs_diff: float = 0.0
    mean_abs_diff: float = 0.0
    n_diffs: int = 0


def compare_arrays(truth, predicted, tolerance=DEFAULT_TOLERANCE) -> Comparison:
    """The check of CompareResults.R on two float32 arrays (differences taken
    in double precision, as R reads the floats into doubles)."""
    t = np.asarray(truth, np.float32).ravel().astype(np.float64)
    p = np.asarray(predicted, np.float32).ravel().astype(np.float64)
    if t.size != p.size:
        return Comparison(False, False,
                          "Validation error: the files have different sizes.")
    diff = np.abs(t - p)
    if np.all(diff <= tolerance):
        sd = float(np.std(t, ddof=1)) if t.size > 1 else float("nan")
        return Comparison(True, True,
                          f"Min: {t.min()}\nMax: {t.max()}\nStandard deviation: "
                          f"{sd}\nValidation successful! The files are "
                          f"numerically similar within a tolerance of {tolerance}.",
                          float(diff.max()) if diff.size else 0.0,
                          float(diff.mean()) if diff.size else 0.0,
                          int(np.sum(diff > 1e-9)))
    n = int(np.sum(diff > 1e-9))
    return Comparison(False, True,
                      "Validation error: the outputs differ numerically.\n"
                      f"Number of significant differences: {n}\n"
                      f"Maximum absolute difference: {diff.max():.4g}\n"
                      f"Mean absolute difference: {diff.mean():.4g}",
                      float(diff.max()), float(diff.mean()), n)


def compare_files(truth_path, predicted_path,
                  tolerance=DEFAULT_TOLERANCE) -> Comparison:
    return compare_arrays(np.fromfile(truth_path, np.float32),
                          np.fromfile(predicted_path, np.float32), tolerance)


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) not in (2, 3):
        print("usage: python -m distributed-cube-average_amd.compare "
              "GROUND_TRUTH.dc PREDICTED.dc [TOLERANCE]", file=sys.stderr)
        return 2
    tol = float(argv[2]) if len(argv) == 3 else DEFAULT_TOLERANCE
    c = compare_files(argv[0], argv[1], tol)
    print(c.report)
    return 0 if c.ok else 1


if __name__ == "__main__":
    sys.exit(main())
