"""The sparse Cholesky analysis (nested-dissection order, supernodes, front maps) on the host, no GPU.

A numpy restatement of the multifrontal factorization driven only by the structure the analysis
exports (the CSR-to-front map, the child-to-parent maps, the front index lists) must reproduce the
dense FP64 solution of the system: a wrong front row, a missing fill row or a bad map changes it.
Systems: the 125-point stencil pattern on sparse voxel sets (a sphere shell and two disjoint blobs,
so the dissection also meets disconnected parts), with padded rows repeating their diagonal and an
upper triangle that differs from the lower one by rounding-sized perturbations."""
import ctypes

import numpy as np
import pytest

from oracle import direct


def _stencil_csr(coords, seed, pad=True):
    """Symmetric positive definite matrix with the 5^3 neighbour pattern of the voxel set, CSR with
    columns ascending and, when pad, rows padded to multiples of 4 with zero diagonal entries."""
    rng = np.random.default_rng(seed)
    key = {tuple(c): i for i, c in enumerate(coords.tolist())}
    n = coords.shape[0]
    rows = [[] for _ in range(n)]
    offs = [(a, b, c) for a in range(-2, 3) for b in range(-2, 3) for c in range(-2, 3)]
    for i, c in enumerate(coords.tolist()):
        for o in offs:
            j = key.get((c[0] + o[0], c[1] + o[1], c[2] + o[2]))
            if j is not None:
                rows[i].append(j)
    vals = {}
    for i in range(n):
        for j in rows[i]:
            if j > i:
                v = rng.uniform(-1.0, 0.2)
                vals[(i, j)] = vals[(j, i)] = v
    A = np.zeros((n, n))
    for (i, j), v in vals.items():
        A[i, j] = v
    np.fill_diagonal(A, np.abs(A).sum(1) + rng.uniform(0.5, 1.5, n))
    rp, cl, vv = [0], [], []
    for i in range(n):
        cs = sorted(rows[i])
        cl += cs
        vv += [A[i, j] for j in cs]
        if pad:
            while (len(cl) - rp[-1]) % 4:
                cl.append(i)
                vv.append(0.0)
        rp.append(len(cl))
    return A, np.array(rp, np.int32), np.array(cl, np.int32), np.array(vv, np.float32)


def _shell(r=7.0, w=1.6):
    g = np.arange(-9, 10)
    c = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    d = np.linalg.norm(c, axis=1)
    return c[np.abs(d - r) < w]


def _blobs():
    g = np.arange(0, 6)
    c = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    return np.concatenate([c, c + np.array([20, 0, 0])], 0)


def _analyze(rp, cl):
    from ngspsr import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    n = rp.shape[0] - 1
    assert L.ngs_chol_analyze_host(rp.ctypes.data_as(ctypes.c_void_p), cl.ctypes.data_as(ctypes.c_void_p), n,
                                   ctypes.byref(h)) == 0, L.ngs_last_error()
    st = (ctypes.c_int64 * 8)()
    assert L.ngs_chol_info(h, st) == 0
    st = [int(v) for v in st]
    nodes = np.zeros(8 * st[0], np.int64)
    fidx = np.zeros(max(st[5], 1), np.int32)
    cmap = np.zeros(max(st[6], 1), np.int32)
    amap = np.zeros(max(st[7], 1), np.int64)
    perm = np.zeros(max(n, 1), np.int32)
    assert L.ngs_chol_structure(h, nodes.ctypes.data_as(ctypes.c_void_p), fidx.ctypes.data_as(ctypes.c_void_p),
                                cmap.ctypes.data_as(ctypes.c_void_p), amap.ctypes.data_as(ctypes.c_void_p)) == 0
    assert L.ngs_chol_permutation(h, perm.ctypes.data_as(ctypes.c_void_p)) == 0
    assert L.ngs_chol_workspace_bytes(h) > 8 * st[4]
    L.ngs_chol_free(h)
    return st, nodes.reshape(-1, 8), fidx, cmap, amap, perm[:n]


def _multifrontal_solve(st, nodes, fidx, cmap, amap, vals, b_new):
    """Multifrontal Cholesky and solve from the exported structure (numpy, FP64)."""
    pool = np.zeros(st[4])
    for k, pos in enumerate(amap[:st[7]]):
        if pos >= 0:
            pool[pos] += float(vals[k])

    def front(i):
        start, ns, f, parent, off = nodes[i, :5]
        return pool[off:off + f * f].reshape(f, f).T  # column-major storage: element (r, c) at c f + r

    heights = nodes[:, 7]
    order = np.argsort(heights, kind="stable")
    for i in order:
        start, ns, f, parent, off, idx_off, cm_off, ht = nodes[i]
        F = front(i)
        for c in np.nonzero(nodes[:, 3] == i)[0]:  # children: extend-add their update matrices
            cs, cns, cf = nodes[c, 0], nodes[c, 1], nodes[c, 2]
            nu = cf - cns
            if nu == 0:
                continue
            U = front(c)[cns:, cns:]
            m = cmap[nodes[c, 6]:nodes[c, 6] + nu]
            for a in range(nu):
                for bb in range(a + 1):
                    F[m[a], m[bb]] += U[a, bb]
        Fl = np.tril(F)
        S = Fl + np.tril(Fl, -1).T
        L11 = np.linalg.cholesky(S[:ns, :ns])
        L21 = np.linalg.solve(L11, S[ns:, :ns].T).T
        F[:ns, :ns] = L11
        F[ns:, :ns] = L21
        F[ns:, ns:] = np.tril(S[ns:, ns:] - L21 @ L21.T)
    y = b_new.copy()
    for i in order:
        start, ns, f, parent, off, idx_off = nodes[i, :6]
        F = front(i)
        idx = fidx[idx_off:idx_off + f]
        z = np.linalg.solve(F[:ns, :ns], y[start:start + ns])
        y[start:start + ns] = z
        y[idx[ns:]] -= F[ns:, :ns] @ z
    for i in order[::-1]:
        start, ns, f, parent, off, idx_off = nodes[i, :6]
        F = front(i)
        idx = fidx[idx_off:idx_off + f]
        t = y[start:start + ns] - F[ns:, :ns].T @ y[idx[ns:]]
        y[start:start + ns] = np.linalg.solve(F[:ns, :ns].T, t)
    return y


@pytest.mark.parametrize("kind,pad", [("shell", True), ("blobs", True), ("shell", False)])
def test_multifrontal_structure_reproduces_dense_solve(kind, pad):
    coords = _shell() if kind == "shell" else _blobs()
    A, rp, cl, vv = _stencil_csr(coords, seed=1, pad=pad)
    n = A.shape[0]
    st, nodes, fidx, cmap, amap, perm = _analyze(rp, cl)
    assert np.array_equal(np.sort(perm), np.arange(n))
    # supernodes tile the new numbering in postorder; parents come after their children
    assert nodes[0, 0] == 0 and np.all(nodes[1:, 0] == nodes[:-1, 0] + nodes[:-1, 1]) and nodes[-1, 0] + nodes[-1, 1] == n
    for i in range(nodes.shape[0]):
        p = nodes[i, 3]
        assert p == -1 or p > i
    if kind == "blobs":
        assert int((nodes[:, 3] == -1).sum()) >= 2  # one tree per connected part
    assert st[2] < n  # the dissection split the system
    # the values as the matrix's entries in the CSR order: A's own values (FP32), padding 0; the
    # upper triangle is perturbed (an assembled matrix is symmetric only to rounding): the factor
    # is of the lower triangle in the caller's numbering, as cuSOLVER reads it
    vv = vv.copy()
    upper = cl > np.repeat(np.arange(n), np.diff(rp))
    vv[upper] *= np.float32(1.0) + np.float32(1e-3) * np.random.default_rng(3).uniform(-1, 1, int(upper.sum())).astype(np.float32)
    A32 = np.zeros_like(A)
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            A32[i, cl[k]] += float(vv[k])
    b = np.random.default_rng(2).standard_normal(n)
    x_ref = np.linalg.solve(direct.lower_symmetric(A32), b)
    assert np.linalg.norm(np.linalg.solve(A32, b) - x_ref) > 1e-8 * np.linalg.norm(x_ref)  # the triangles differ
    y = _multifrontal_solve(st, nodes, fidx, cmap, amap, vv, b[perm])
    x = np.empty(n)
    x[perm] = y
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-11
    # stored factor entries stay well below the dense triangle
    assert st[3] < n * (n + 1) // 2
