"""Synthetic KKT / QP inputs of the benchmark configurations.

All generators draw from one seeded numpy Generator and return matrices as
symmetric lower-triangular CSC arrays (diagonal first in every column, rows
sorted), the layout the reference's CSC type expects
(include/nasoq/common/def.h:73-122; find_perturbation reads x[p[j]] as the
diagonal, src/QP/linear_solver_wrapper.cpp:352-363).
"""
import numpy as np
import scipy.sparse as sp


def lower_csc(K):
    """Symmetric scipy matrix -> (n, Ap, Ai, Ax) lower CSC, explicit diagonal first."""
    K = sp.csc_matrix(K)
    n = K.shape[0]
    L = sp.tril(K, format="csc")
    # make every diagonal entry structurally present (explicit zeros allowed)
    d = L.diagonal()
    L = (L + sp.csc_matrix((np.full(n, 1.0), (np.arange(n), np.arange(n))), shape=(n, n))).tocsc()
    L.sort_indices()
    Ap = L.indptr.astype(np.int32)
    Ai = L.indices.astype(np.int32)
    Ax = L.data.astype(np.float64)
    # restore the diagonal values (may be exact zeros)
    assert np.array_equal(Ai[Ap[:-1]], np.arange(n))
    Ax[Ap[:-1]] = d
    return n, Ap, Ai, Ax


def lower_to_full(n, Ap, Ai, Ax):
    L = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    return (L + sp.tril(L, -1).T).tocsc()


def laplacian_2d(r, c, shift):
    e = lambda m: sp.diags([-np.ones(m - 1), 2 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
    H = sp.kron(sp.identity(c), e(r)) + sp.kron(e(c), sp.identity(r))
    return (H + shift * sp.identity(r * c)).tocsc()


def laplacian_3d(g, shift):
    e = sp.diags([-np.ones(g - 1), 2 * np.ones(g), -np.ones(g - 1)], [-1, 0, 1])
    I = sp.identity(g)
    H = (sp.kron(sp.kron(I, I), e) + sp.kron(sp.kron(I, e), I) + sp.kron(sp.kron(e, I), I))
    return (H + shift * sp.identity(g ** 3)).tocsc()


def random_rows(rng, m, n, k):
    """m x n sparse matrix with k distinct uniform random columns per row, N(0,1) values."""
    cols = np.empty((m, k), dtype=np.int64)
    for r in range(m):
        cols[r] = rng.choice(n, size=k, replace=False)
    vals = rng.standard_normal((m, k))
    rows = np.repeat(np.arange(m), k)
    return sp.csr_matrix((vals.ravel(), (rows, cols.ravel())), shape=(m, n))


def kkt(H, A):
    """K = [H A'; A 0] with an explicitly stored zero (2,2) diagonal."""
    m = A.shape[0]
    Z = sp.csc_matrix((m, m))
    return sp.bmat([[H, A.T], [A, Z]], format="csc")


def kkt_config(grid, m, nnz_per_row, shift=1e-3, seed=0, nrhs=4, ill=False):
    """KKT of a 3-D 7-point Laplacian on grid^3 plus m random constraint rows
    (the plain-LBL and mixed-precision configurations).  ill=True scales H's
    diagonal by 10^U[-3,3] (the ill-conditioned variant)."""
    rng = np.random.default_rng(seed)
    H = laplacian_3d(grid, shift)
    n = H.shape[0]
    if ill:
        s = 10.0 ** rng.uniform(-3, 3, size=n)
        H = (H + sp.diags(H.diagonal() * (s - 1.0))).tocsc()
    A = random_rows(rng, m, n, nnz_per_row)
    K = kkt(H, A)
    B = rng.uniform(-1.0, 1.0, size=(K.shape[0], nrhs))
    return lower_csc(K) + (B,)


def kkt_small(n_h, m, seed=0, density=0.05, nrhs=1, shift=1.0):
    """Small random KKT [H A'; A 0] with H SPD (test sizes)."""
    rng = np.random.default_rng(seed)
    R = sp.random(n_h, n_h, density=density, random_state=rng, format="csc")
    H = (R @ R.T + shift * sp.identity(n_h)).tocsc()
    A = sp.random(m, n_h, density=min(1.0, 4.0 / max(n_h, 1)), random_state=rng, format="csr")
    # make sure each constraint row has at least one entry
    for r in range(m):
        if A.indptr[r + 1] == A.indptr[r]:
            A = A.tolil()
            A[r, rng.integers(n_h)] = rng.standard_normal()
            A = A.tocsr()
    K = kkt(H, A)
    B = rng.uniform(-1.0, 1.0, size=(K.shape[0], nrhs))
    return lower_csc(K) + (B,)


def tall_border(n1, m2, chains=2, seed=0, nrhs=2):
    """Quasi-definite [T B'; B -E] in natural order: T = `chains` independent
    dense SPD blocks of c = n1 / chains columns (T_ij = 1/2 off the diagonal,
    c/2 + 1 on it), each one fundamental supernode c wide with c + m2 rows
    (relaxed amalgamation off); B a dense m2 x n1 N(0,1) block; E =
    diag(U[1, 2]).  The m2 rows form one dense m2 x m2 clique, the blocks'
    common parent.  Inertia (n1, m2, 0).  Built directly as lower CSC."""
    rng = np.random.default_rng(seed)
    n = n1 + m2
    c = n1 // chains
    assert c * chains == n1
    Bd = rng.standard_normal((m2, n1))
    jj = np.arange(n1)
    per = np.r_[(c - jj % c) + m2, np.ones(m2, dtype=np.int64)]
    cnt = np.r_[0, np.cumsum(per)]
    nnz = int(cnt[-1])
    Ai = np.empty(nnz, dtype=np.int32)
    Ax = np.empty(nnz, dtype=np.float64)
    rows_b = np.arange(n1, n, dtype=np.int32)
    for j in range(n1):
        q = int(cnt[j])
        e = (j // c + 1) * c  # end of j's block
        k = e - j
        Ai[q:q + k] = np.arange(j, e)
        Ax[q:q + k] = 0.5
        Ax[q] = c / 2.0 + 1.0
        q += k
        Ai[q:q + m2] = rows_b
        Ax[q:q + m2] = Bd[:, j]
    q0 = int(cnt[n1])
    Ai[q0:] = rows_b
    Ax[q0:] = -rng.uniform(1.0, 2.0, m2)
    B = rng.uniform(-1.0, 1.0, size=(n, nrhs))
    return n, cnt.astype(np.int32), Ai, Ax, B


def checksum(*arrays):
    """Order-sensitive checksum of generated arrays (for recording inputs)."""
    import hashlib
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


def block_indefinite(nblk, bs, seed=0, coupling=0.3, nrhs=1):
    """Block-tridiagonal symmetric indefinite matrix whose diagonal blocks are
    dense with an all-zero diagonal (so Bunch-Kaufman must take 2x2 pivots
    inside supernodes), coupled by sparse random off-diagonal blocks."""
    rng = np.random.default_rng(seed)
    n = nblk * bs
    blocks = []
    for b in range(nblk):
        D = rng.standard_normal((bs, bs))
        D = D + D.T
        np.fill_diagonal(D, 0.0)
        blocks.append(sp.coo_matrix(D))
    K = sp.block_diag(blocks, format="lil")
    for b in range(nblk - 1):
        C = sp.random(bs, bs, density=coupling, random_state=rng).toarray()
        K[(b + 1) * bs:(b + 2) * bs, b * bs:(b + 1) * bs] = C
        K[b * bs:(b + 1) * bs, (b + 1) * bs:(b + 2) * bs] = C.T
    B = rng.uniform(-1.0, 1.0, size=(n, nrhs))
    return lower_csc(K.tocsc()) + (B,)


def batch_qp_kkts(count=1024, seed=0, nmin=16, nmax=160, meq_max=64, nnz_row=4):
    """The KKT systems [H A'; A 0] of a batch of QPs of uneven size: H the
    5-point Laplacian of an r x c grid (r, c ~ U{nmin..nmax}) + 1e-2 I,
    A ~ U{1..meq_max} rows of nnz_row N(0,1) entries, right-hand side
    [-q; A x0] with q ~ N(0,1) and x0 ~ U[-1,1].  Returns a list of
    (n, Ap, Ai, Ax, b) with the lower CSC of each KKT (explicit zero
    diagonal on the constraint block)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        rr, cc = int(rng.integers(nmin, nmax + 1)), int(rng.integers(nmin, nmax + 1))
        H = laplacian_2d(rr, cc, 1e-2)
        nh = H.shape[0]
        m = int(rng.integers(1, meq_max + 1))
        A = random_rows(rng, m, nh, nnz_row)
        q = rng.standard_normal(nh)
        x0 = rng.uniform(-1.0, 1.0, nh)
        n, Ap, Ai, Ax = lower_csc(kkt(H, A))
        out.append((n, Ap, Ai, Ax, np.concatenate([-q, A @ x0])))
    return out


def box_qp(rng, r, c, meq, nnz_row=4):
    """One member of the batched QP configuration: H the 5-point Laplacian of
    an r x c mesh + 1e-2 I, q ~ N(0,1), meq equality rows of nnz_row N(0,1)
    entries with b = A x0, x0 ~ U[-1,1], and the box |x| <= 1 as the 2n
    inequality rows [I; -I] x <= 1.  Returns (H (full), q, A, b, C, d)."""
    H = laplacian_2d(r, c, 1e-2)
    n = H.shape[0]
    q = rng.standard_normal(n)
    A = random_rows(rng, meq, n, nnz_row).tocsc()
    x0 = rng.uniform(-1.0, 1.0, n)
    b = A @ x0
    C = sp.vstack([sp.identity(n), -sp.identity(n)], format="csc")
    d = np.ones(2 * n)
    return H, q, A, b, C, d


def box_qp_batch(count, seed=0, lo=16, hi=160, meq_max=64):
    """The batched QP configuration: `count` independent box-constrained QPs
    (box_qp) with r, c ~ U{lo..hi} and meq ~ U{1..meq_max}, all from one
    seeded generator (the first k members do not depend on count)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        r, c = (int(v) for v in rng.integers(lo, hi + 1, size=2))
        meq = int(rng.integers(1, meq_max + 1))
        out.append(box_qp(rng, r, c, meq))
    return out


def qp_config(grid=40, meq=100, nnz_row=5, seed=0):
    """The convex QP min 1/2 x'Hx + q'x s.t. Ax = b, Cx <= d of the plain
    NASOQ configuration: H the 5-point Laplacian of a grid x grid mesh
    + 1e-2 I, q ~ N(0,1), A meq x n with nnz_row N(0,1) entries at uniform
    random columns, b = A x0 with x0 ~ U[-1,1], C = -I and d = 0.5 (x >= -0.5);
    all from one seeded generator.  Returns (H (full), q, A, b, C, d)."""
    rng = np.random.default_rng(seed)
    H = laplacian_2d(grid, grid, 1e-2)
    n = H.shape[0]
    q = rng.standard_normal(n)
    A = random_rows(rng, meq, n, nnz_row).tocsc()
    x0 = rng.uniform(-1.0, 1.0, n)
    b = A @ x0
    C = -sp.identity(n, format="csc")
    d = np.full(n, 0.5)
    return H, q, A, b, C, d
