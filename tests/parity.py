"""Helpers shared by the parity tests: tie-aware label comparison and
centroid error, both against the FP64 oracle."""
import numpy as np

# Labels may differ from the oracle's only where the oracle's own distances
# to the two clusters are within this SBD tolerance (FP32 vs FP64 ties).
SBD_TIE = 1e-6
# Centroid tolerance: relative L2 error against the FP64 oracle.
CENT_RTOL = 1e-4


def label_mismatch_report(gpu_labels, oracle_labels, oracle_dmat, tie=SBD_TIE):
    """Return (n_mismatch, n_non_tie, worst_gap) where a mismatch at i is a
    tie if |d(i, gpu) - d(i, oracle)| <= tie in the oracle's distances."""
    gpu_labels = np.asarray(gpu_labels)
    oracle_labels = np.asarray(oracle_labels)
    bad = np.nonzero(gpu_labels != oracle_labels)[0]
    if bad.size == 0:
        return 0, 0, 0.0
    gaps = np.abs(oracle_dmat[bad, gpu_labels[bad]] - oracle_dmat[bad, oracle_labels[bad]])
    return int(bad.size), int(np.sum(gaps > tie)), float(gaps.max())


def assert_labels_match(gpu_labels, oracle_labels, oracle_dmat, tie=SBD_TIE, what="labels"):
    n, non_tie, gap = label_mismatch_report(gpu_labels, oracle_labels, oracle_dmat, tie)
    assert non_tie == 0, f"{what}: {non_tie} of {n} mismatches are not ties (max gap {gap:.3e})"


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    if nb == 0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / nb)


def centroid_errors(gpu_c, oracle_c):
    return [rel_l2(g, o) for g, o in zip(gpu_c, oracle_c)]
