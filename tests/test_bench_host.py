"""Host-side pieces of bench.py (no GPU): the module imports, the work
counts behind its roofline figures match their tabulated values, brute force
agrees with the closed counts, and more than one GPU is refused with a
message."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    sys.path.insert(0, ROOT)
    import bench
    return bench


def test_csteps_matches_tabulated_counts():
    """T_con(even) and T_con(odd) with l3 <= lmax (LW = lmax + 1) for the
    configurations' lmax values."""
    b = _bench()
    assert b.csteps(0, 767, 768) == (56_844_480, 56_623_104)
    assert b.csteps(0, 3071, 3072) == (3_627_418_368, 3_623_878_656)


def test_csteps_and_qp_terms_brute_force():
    b = _bench()
    lmax, LW = 23, 20
    ev = od = 0
    for l1 in range(lmax + 1):
        for l2 in range(l1, lmax + 1):
            for c in range(min(l1, l2) + 1):
                ev += (l1 + l2 - 2 * c) <= LW - 1
            for c in range(min(l1, l2)):
                od += (l1 + l2 - 2 * c - 1) <= LW - 1
    assert b.csteps(0, lmax, LW) == (ev, od)
    spins = (2, -2, 2, 2)
    bl, bu, LW = 3, 2, 30
    fam = con = 0
    for lpp in range(2, lmax + 1):
        for l in range(max(2, lpp - bl), min(lmax, lpp + bu) + 1):
            na = max(abs(l - lpp), abs(spins[2] + spins[0]))
            nb = max(abs(l - lpp), abs(spins[3] + spins[1]))
            fam += (l + lpp - na + 1) + (l + lpp - nb + 1)
            con += max(0, min(l + lpp, LW - 1) - max(na, nb) + 1)
    assert b.qp_terms(lmax, bl, bu, LW, spins) == (fam, con)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
