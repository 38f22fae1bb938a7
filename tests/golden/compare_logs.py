"""Compares the log/*.dat files of a run of the 256^2 free-decay case
(tests/golden/decay2d_256) with the expected ones, record by record.  For a
maintainer who can build the reference (gfortran + FFTW-MPI + HDF5): copy
tests/golden/decay2d_256, run Bastr in it, then

  python tests/golden/compare_logs.py <case>/log tests/golden/decay2d_256/expected/log

Energies are compared relative to Es+Ed of the same record (Es of this purely
dilatational field is roundoff); in a spectrum file, whose records of one nstep
are the shells of one spectrum, relative to the largest shell's Es+Ed of that
spectrum (a shell far below the spectrum's peak holds energy at the roundoff of
the peak's, which no two summation orders share).  Every other column is
compared relative to its own largest magnitude.  Exit status 0 when all agree within --tol (default 1e-10).
"""
import argparse
import os
import sys

import numpy as np


def read_log(path):
    """(header, rows): rows of floats, the leading nstep column included."""
    lines = open(path).read().splitlines()
    return lines[0].split(), np.array([[float(x) for x in l.split()] for l in lines[1:] if l.strip()])


def compare(run_dir, exp_dir, tol):
    worst = {}
    for name in sorted(os.listdir(exp_dir)):
        if not name.endswith(".dat"):
            continue
        hdr, ref = read_log(os.path.join(exp_dir, name))
        path = os.path.join(run_dir, name)
        if not os.path.exists(path):
            worst[name] = float("inf")
            continue
        _, got = read_log(path)
        if got.shape != ref.shape or not np.array_equal(got[:, 0], ref[:, 0]):
            worst[name] = float("inf")
            continue
        err = np.abs(got - ref)
        scale = np.maximum(np.abs(ref).max(axis=0, keepdims=True), 1e-300)
        if "Es" in hdr and "Ed" in hdr:
            tot = np.abs(ref[:, [hdr.index("Es")]]) + np.abs(ref[:, [hdr.index("Ed")]])
            steps = ref[:, 0]
            if len(np.unique(steps)) < len(steps):  # shells of spectra: the spectrum's peak
                for s in np.unique(steps):
                    tot[steps == s] = tot[steps == s].max()
            for c in ("Es", "Ed"):
                j = hdr.index(c)
                scale[:, j] = 0
                err[:, j] = err[:, j] / np.maximum(tot[:, 0], 1e-300)
            rel = err / np.where(scale > 0, scale, 1.0)
            for c in ("Es", "Ed"):
                rel[:, hdr.index(c)] = err[:, hdr.index(c)]
        else:
            rel = err / scale
        worst[name] = float(rel.max())
    return worst


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("run_log_dir")
    ap.add_argument("expected_log_dir")
    ap.add_argument("--tol", type=float, default=1e-10)
    a = ap.parse_args()
    worst = compare(a.run_log_dir, a.expected_log_dir, a.tol)
    for k, v in worst.items():
        print(f"{k:24s} max relative difference {v:.3e}")
    sys.exit(0 if all(v <= a.tol for v in worst.values()) else 1)


if __name__ == "__main__":
    main()
