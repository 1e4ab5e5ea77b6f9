"""Extract the reference's published error tables into a small fixture.

Reads two output files that the reference repository ships (data, not code):

* benchmarks/multipoles_error/out/multipoles_error.txt -- the relative error of
  one particle-node interaction (monopole / quadrupole / octupole) for a
  target approaching a 3000-body Plummer node (a = 0.58, truncated at 6) along
  the diagonal (benchmarks/multipoles_error/multipoles_error.f90): kept are
  the theta = 1.0 sequences (the longest; theta only decides where a sequence
  stops) of the 20 tests at eps = 0 and eps = 0.15.
* benchmarks/forces_time_and_error/out/bigger_N/test_{monopole,quadrupole,
  octupole}.txt -- the McMillan-Aarseth error of the Barnes-Hut walk at N =
  50000 (Plummer, a = 0.58, truncated at 10, eps = 0), 21 tests per theta
  (benchmarks/forces_time_and_error/forces_time_and_error.f90:113-155).

Writes tests/golden/published_errors.npz.  Run once, where the reference is.
"""
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def multipole_sequences(path, eps):
    rows = np.loadtxt(path)
    blocks, cur = [], None
    for r in rows:
        key = (r[1], r[2])
        # a new (theta, eps) sequence, or the next test's sequence restarting far out
        if cur is None or key != cur[0] or (r[0] == 1 and r[3] > cur[1][-1][3] * 1.0001):
            cur = [key, []]
            blocks.append(cur)
        cur[1].append(r)
    out = []
    for key, b in blocks:
        if key == (1.0, eps):
            a = np.array(b)
            assert np.array_equal(a[0::3, 0], np.ones(len(a) // 3))
            out.append(np.stack([a[0::3, 3], a[0::3, 4], a[1::3, 4], a[2::3, 4]]))
    return out


def main():
    fx = {}
    mp = os.path.join(REF, "benchmarks/multipoles_error/out/multipoles_error.txt")
    for tag, eps in (("eps0", 0.0), ("eps015", 0.15)):
        seqs = multipole_sequences(mp, eps)
        assert len(seqs) == 20
        L = max(s.shape[1] for s in seqs)
        arr = np.full((20, 4, L), np.nan)
        for i, s in enumerate(seqs):
            arr[i, :, : s.shape[1]] = s
        fx["multipole_" + tag] = arr  # (test, [d, err_mono, err_quad, err_oct], row)
    for name in ("monopole", "quadrupole", "octupole"):
        rows = np.loadtxt(os.path.join(REF, "benchmarks/forces_time_and_error/out/bigger_N",
                                       "test_%s.txt" % name))
        sel = rows[(rows[:, 0] == 50000) & (rows[:, 1] > 0)]
        fx["bigger_N_" + name] = sel[:, [1, 4]]  # theta, error
    np.savez_compressed(os.path.join(HERE, "published_errors.npz"), **fx)


if __name__ == "__main__":
    main()
