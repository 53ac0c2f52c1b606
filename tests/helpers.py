"""Shared test helpers (global assembly from subdomain matrices, etc.)."""
import numpy as np
import scipy.sparse as sp


def global_system(model, om):
    """Global matrix/load assembled from the oracle's subdomain matrices,
    with the PEC edges removed: (A_free, f_free, free_mask, n_edges)."""
    N = model.n_edges
    rows, cols, vals = [], [], []
    F = np.zeros(N, complex)
    pec = np.zeros(N, bool)
    off = 0
    for s in range(model.nsub):
        g = om.get(s)
        n = g["n"]
        gl = model.ledge_global[off:off + n]
        off += n
        r = np.repeat(np.arange(n), np.diff(g["rowptr"]))
        rows.append(gl[r])
        cols.append(gl[g["col"]])
        vals.append(g["val"])
        F[gl] += g["f"]
        pec[gl[model.ibc[model.bc_off[s]:model.bc_off[s + 1]]]] = True
    L = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N)).tocsr()
    A = L + L.T - sp.diags(L.diagonal())
    free = ~pec
    return A, F, free


def to_global(model, xs):
    x = np.zeros(model.n_edges, complex)
    off = 0
    for s in range(model.nsub):
        n = model.nmp[s]
        x[model.ledge_global[off:off + n]] = xs[s]
        off += n
    return x


def dense_lower(rowptr, col, val, n):
    A = np.zeros((n, n), complex)
    for i in range(n):
        for k in range(rowptr[i], rowptr[i + 1]):
            A[i, col[k]] += val[k]
            if col[k] != i:
                A[col[k], i] += val[k]
    return A
