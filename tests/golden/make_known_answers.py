"""Write known_answers.json: the closed-form answers that the reference's own
pair-counter tests assert, with the inputs that produce them.

Each case records the function, the reference test it comes from (file:line,
PC = halotools/mock_observables/pair_counters), how its samples are generated
(tests/helpers.py restates the reference's cf_helpers generators: a regular
grid, or a seeded tight locus), the bin edges and keywords exactly as the
reference test forms them, and the expected output the reference asserts.
Run from the repository root: python tests/golden/make_known_answers.py
"""
import json
import os

import numpy as np

SPACING = 1.0 / 10
EPS = 0.0001
# rbins of the 10^3 grid tests (PC/test_pair_counters/test_npairs_3d.py:184-193)
GRID_BINS = [SPACING / 100.0, SPACING + EPS, SPACING * np.sqrt(2) + EPS, SPACING * np.sqrt(3) + EPS]
N3 = 10 ** 3

CASES = [
    dict(name="npairs_3d_tight_locus", function="npairs_3d",
         reference="PC/test_pair_counters/test_npairs_3d.py:33-45",
         sample1=dict(kind="locus", npts=100, xc=0.1, yc=0.1, zc=0.1, seed=43),
         sample2=dict(kind="locus", npts=100, xc=0.1, yc=0.1, zc=0.2, seed=43),
         args=dict(rbins=[0.05, 0.15, 0.3]), kwargs=dict(period=1.0),
         expected=[0, 100 * 100, 100 * 100]),
    dict(name="npairs_3d_grid", function="npairs_3d",
         reference="PC/test_pair_counters/test_npairs_3d.py:180-197",
         sample1=dict(kind="grid", npts_per_dim=10), sample2=dict(kind="grid", npts_per_dim=10),
         args=dict(rbins=GRID_BINS), kwargs=dict(period=1.0, approx_cell1_size=0.1),
         expected=[N3, 7 * N3, 19 * N3, 27 * N3]),
    dict(name="npairs_projected_grid_thin", function="npairs_projected",
         reference="PC/test_pair_counters/test_npairs_projected.py:220-240",
         sample1=dict(kind="grid", npts_per_dim=10), sample2=dict(kind="grid", npts_per_dim=10),
         args=dict(rp_bins=GRID_BINS, pi_max=SPACING / 100.0), kwargs=dict(period=1.0, approx_cell1_size=0.1),
         expected=[N3, 5 * N3, 9 * N3, 9 * N3]),
    dict(name="npairs_projected_grid_thick", function="npairs_projected",
         reference="PC/test_pair_counters/test_npairs_projected.py:243-263",
         sample1=dict(kind="grid", npts_per_dim=10), sample2=dict(kind="grid", npts_per_dim=10),
         args=dict(rp_bins=GRID_BINS, pi_max=1.1 * SPACING), kwargs=dict(period=1.0, approx_cell1_size=0.1),
         expected=[3 * N3, 15 * N3, 27 * N3, 27 * N3]),
    dict(name="npairs_xy_z_tight_locus", function="npairs_xy_z",
         reference="PC/test_pair_counters/test_npairs_xy_z.py:20-35",
         sample1=dict(kind="locus", npts=100, xc=0.1, yc=0.1, zc=0.1, seed=43),
         sample2=dict(kind="locus", npts=90, xc=0.1, yc=0.2, zc=0.105, seed=43),
         args=dict(rp_bins=[0.05, 0.15, 0.3], pi_bins=[0.0, 0.15]), kwargs=dict(period=1.0),
         expected=[[0, 0], [0, 100 * 90], [0, 100 * 90]]),
    dict(name="npairs_xy_z_tight_locus_periodic", function="npairs_xy_z",
         reference="PC/test_pair_counters/test_npairs_xy_z.py:38-53",
         sample1=dict(kind="locus", npts=100, xc=0.1, yc=0.05, zc=0.05, seed=43),
         sample2=dict(kind="locus", npts=300, xc=0.1, yc=0.95, zc=0.95, seed=43),
         args=dict(rp_bins=[0.05, 0.25, 0.3], pi_bins=[0.05, 0.15]), kwargs=dict(period=1.0),
         expected=[[0, 0], [0, 100 * 300], [0, 100 * 300]]),
    dict(name="npairs_xy_z_tight_locus_unperiodic", function="npairs_xy_z",
         reference="PC/test_pair_counters/test_npairs_xy_z.py:55-57",
         sample1=dict(kind="locus", npts=100, xc=0.1, yc=0.05, zc=0.05, seed=43),
         sample2=dict(kind="locus", npts=300, xc=0.1, yc=0.95, zc=0.95, seed=43),
         args=dict(rp_bins=[0.05, 0.25, 0.3], pi_bins=[0.05, 0.15]), kwargs=dict(),
         expected=[[0, 0], [0, 0], [0, 0]]),
    dict(name="npairs_xy_z_grid", function="npairs_xy_z",
         reference="PC/test_pair_counters/test_npairs_xy_z.py:175-198",
         sample1=dict(kind="grid", npts_per_dim=10), sample2=dict(kind="grid", npts_per_dim=10),
         args=dict(rp_bins=GRID_BINS, pi_bins=[0.0, 1.1 * SPACING]), kwargs=dict(period=1.0, approx_cell1_size=0.1),
         expected=[[N3, 3 * N3], [5 * N3, 15 * N3], [9 * N3, 27 * N3], [9 * N3, 27 * N3]]),
]

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.json")
    with open(out, "w") as f:
        json.dump({"cases": CASES}, f, indent=1)
    print("wrote", out, len(CASES), "cases")
