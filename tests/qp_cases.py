"""Shared QP fixtures: the reference's example, quadprog test and SMP files
(tests/golden: quadprog_1.json = tests/1.json, *.yml and out_correct.csv =
data/ of the reference)."""
import csv
import json
import os

import numpy as np
import scipy.sparse as sp

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def nasoq_main_problem():
    # examples/nasoq_main.cpp:17-46
    H = sp.csc_matrix(np.diag([2.0, 2.0]))
    q = np.array([-4.0, -4.0])
    C = sp.csc_matrix(np.array([[2.0, 1.0], [1.0, -1.0], [-1.0, -1.0], [-2.0, 1.0]]))
    d = np.array([2.0, 1.0, 1.0, 2.0])
    return H, q, C, d


def _json_csc(m):
    if m["rows"] == 0:
        return sp.csc_matrix((0, m["cols"]))
    return sp.csc_matrix((np.array(m["values"], float), np.array(m["inner_indices"]), np.array(m["outer_starts"])),
                         shape=(m["rows"], m["cols"]))


def quadprog_1():
    with open(os.path.join(GOLDEN, "quadprog_1.json")) as f:
        d = json.load(f)
    out = {k: _json_csc(d[k]) for k in ("H", "A", "C")}
    out.update({k: np.array(d[k], float) for k in ("q", "b", "d", "x", "y", "z")})
    return out


def out_correct():
    rows = {}
    with open(os.path.join(GOLDEN, "out_correct.csv")) as f:
        for r in csv.DictReader(f):
            rows[(r["Tool Name"], r["Problem Name"])] = r
    return rows
