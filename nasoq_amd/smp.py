"""Reader for the SMP (sparse mathematical programming) QP format of the
reference's smp-format library (smp-format/smp_format.h:78-335,
mp_format_converter.h:580-860).

An SMP file is YAML whose sections hold MatrixMarket text:

    "Quadratic"            H (symmetric, coordinate)       -> lower CSC
    "Linear"               q (array)
    "Equality"             A (coordinate, me x n)
    "Equality bounds"      a (array)
    "Inequality"           C (coordinate, mi x n)
    "Inequality u-bounds"  u (array)       Cx <= u
    "Inequality l-bounds"  l (array)       l <= Cx   (optional)
    "Fixed"                constant term of the objective

load_smp() returns the inequality form (IEForm) that NASOQ consumes:
min 1/2 x'Hx + q'x + r  s.t.  Ax = a,  Bx <= b.  Two-sided rows are split
as in find_inequalities_by_bounds (l == u becomes an equality, a finite l
adds the row -C x <= -l).
"""
import numpy as np
import scipy.sparse as sp
import yaml


def _mm_lines(text):
    lines = [ln.strip() for ln in text.strip().splitlines()]
    header = lines[0].lower()
    body = [ln for ln in lines[1:] if ln and not ln.startswith("%")]
    return header, body


def read_mm(text, symmetric_lower=False):
    """MatrixMarket text -> scipy CSC (coordinate) or dense column (array)."""
    header, body = _mm_lines(text)
    dims = body[0].split()
    if "array" in header:
        m, ncol = int(dims[0]), int(dims[1])
        vals = np.array([float(v) for v in body[1:1 + m * ncol]])
        return vals.reshape(ncol, m).T[:, 0] if ncol == 1 else vals.reshape(ncol, m).T
    m, n, nnz = int(dims[0]), int(dims[1]), int(dims[2])
    rows, cols, vals = [], [], []
    for ln in body[1:1 + nnz]:
        parts = ln.split()
        r, c, v = int(parts[0]) - 1, int(parts[1]) - 1, float(parts[2])
        if symmetric_lower and r < c:
            r, c = c, r
        rows.append(r)
        cols.append(c)
        vals.append(v)
    M = sp.csc_matrix((vals, (rows, cols)), shape=(m, n))
    M.sum_duplicates()
    M.sort_indices()
    return M


def load_smp(path):
    with open(path) as f:
        doc = yaml.safe_load(f)
    sec = {k.lower(): v for k, v in doc.items()}
    H = read_mm(sec["quadratic"], symmetric_lower=True)
    n = H.shape[0]
    q = read_mm(sec["linear"]) if "linear" in sec else np.zeros(n)
    r = float(str(sec.get("fixed", "0")).split()[0]) if "fixed" in sec else 0.0
    A = read_mm(sec["equality"]) if "equality" in sec else sp.csc_matrix((0, n))
    a = read_mm(sec["equality bounds"]) if "equality bounds" in sec else np.zeros(0)
    C = read_mm(sec["inequality"]) if "inequality" in sec else sp.csc_matrix((0, n))
    mi = C.shape[0]
    u = read_mm(sec["inequality u-bounds"]) if "inequality u-bounds" in sec else None
    l = read_mm(sec["inequality l-bounds"]) if "inequality l-bounds" in sec else None
    if l is None:
        B, b = C, (u if u is not None else np.zeros(mi))
    else:
        big = 1e20
        u = np.full(mi, big) if u is None else u
        l = np.where(np.abs(l) < 1e-14, 0.0, l)
        u = np.where(np.abs(u) < 1e-14, 0.0, u)
        Cr = sp.csr_matrix(C)
        eq_rows, eq_b, in_rows, in_b = [], [], [], []
        for i in range(mi):
            row = Cr[i]
            if l[i] == u[i]:
                eq_rows.append(row)
                eq_b.append(u[i])
                continue
            if u[i] < big:
                in_rows.append(row)
                in_b.append(u[i])
            if l[i] > -big:
                in_rows.append(-row)
                in_b.append(-l[i])
        if eq_rows:
            A = sp.vstack([A] + eq_rows, format="csc")
            a = np.concatenate([a, eq_b])
        B = sp.vstack(in_rows, format="csc") if in_rows else sp.csc_matrix((0, n))
        b = np.array(in_b)
    return {"H": sp.tril(H, format="csc"), "q": np.asarray(q, float), "r": r,
            "A": sp.csc_matrix(A), "a": np.asarray(a, float), "B": sp.csc_matrix(B), "b": np.asarray(b, float)}
