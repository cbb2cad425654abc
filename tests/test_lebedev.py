"""Lebedev-Laikov angular rules: the project table and both expansions
(library and oracle) against what the mathematics fixes, and against the
reference's own rule code where it has been built here (oracle/_ref)."""
import os
import subprocess

import numpy as np
import pytest

import chinium_amd as ca
import oracle_lib as ol

# rule size -> algebraic degree (Lebedev & Laikov 1999)
DEGREE = {6: 3, 14: 5, 26: 7, 38: 9, 50: 11, 74: 13, 86: 15, 110: 17, 146: 19, 170: 21, 194: 23,
          230: 25, 266: 27, 302: 29, 350: 31, 434: 35, 590: 41}
REF_BIN = os.path.join(ol.ROOT, "oracle", "_ref", "lebedev_ref")


def dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def sphere_mean(a, b, c):
    """(1/4pi) * integral of x^a y^b z^c over the unit sphere."""
    if a % 2 or b % 2 or c % 2:
        return 0.0
    return dfact(a - 1) * dfact(b - 1) * dfact(c - 1) / dfact(a + b + c + 1)


def exactness_error(rule, deg):
    x, y, z, w = rule.T
    worst = 0.0
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                q = np.sum(w * x ** a * y ** b * z ** c)
                worst = max(worst, abs(q - sphere_mean(a, b, c)))
    return worst


@pytest.mark.parametrize("n", sorted(DEGREE))
def test_library_rule_exact_to_degree(n):
    rule = ca.lebedev_rule(n)
    assert rule.shape == (n, 4)
    assert np.abs(np.linalg.norm(rule[:, :3], axis=1) - 1).max() < 1e-15
    assert abs(rule[:, 3].sum() - 1) < 1e-14
    # exact for every monomial x^a y^b z^c with a+b+c up to the rule's degree:
    # a wrong digit in an orbit generator or weight, or a missing orbit, breaks it
    assert exactness_error(rule, DEGREE[n]) < 2e-15


def test_exactness_check_is_sensitive():
    rule = ca.lebedev_rule(86).copy()
    rule[0, 3] *= 1 + 1e-9
    assert exactness_error(rule, DEGREE[86]) > 1e-12


@pytest.mark.parametrize("n", sorted(DEGREE))
def test_oracle_rule_equals_library_rule(n):
    a = ca.lebedev_rule(n)
    b = ol.lebedev(n)
    ka = np.lexsort((a[:, 2], a[:, 1], a[:, 0]))
    kb = np.lexsort((b[:, 2], b[:, 1], b[:, 0]))
    assert np.abs(a[ka] - b[kb]).max() < 1e-15
    assert len(np.unique(np.round(a[:, :3], 12), axis=0)) == n


@pytest.mark.skipif(not os.path.exists(REF_BIN), reason="reference Lebedev rules not built (oracle/_ref)")
def test_rules_match_reference_build():
    orders = sorted(DEGREE)
    txt = subprocess.run([REF_BIN] + [str(o) for o in orders], capture_output=True, text=True, check=True).stdout
    blocks = txt.split("rule ")[1:]
    for blk in blocks:
        lines = blk.strip().split("\n")
        n = int(lines[0])
        ref = np.array([[float(t) for t in l.split()] for l in lines[1:]])
        mine = ca.lebedev_rule(n)
        kr = np.lexsort((ref[:, 2], ref[:, 1], ref[:, 0]))
        km = np.lexsort((mine[:, 2], mine[:, 1], mine[:, 0]))
        assert np.abs(ref[kr] - mine[km]).max() < 1e-15, n
