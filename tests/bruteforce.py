"""Independent numpy brute force for pinning the oracle (and, through it, the GPU).

Everything here is written from the mathematics, not from the reference's
code: the M'4 kernel in its factored textbook form (Monaghan 1985), periodic
images by modular arithmetic, sums by np.add.at.  Sizes are kept tiny.
"""
import numpy as np


def m4(r):
    """M'4: 1 - 5/2 r^2 + 3/2 r^3 (r<1), 1/2 (2-r)^2 (1-r) (1<=r<2), 0 beyond."""
    r = np.abs(r)
    inner = 1.0 - 2.5 * r ** 2 + 1.5 * r ** 3
    outer = 0.5 * (2.0 - r) ** 2 * (1.0 - r)
    return np.where(r < 1.0, inner, np.where(r < 2.0, outer, 0.0))


def dm4(s):
    """d/ds M'4(|s|)."""
    r = np.abs(s)
    sg = np.sign(s)
    inner = -5.0 * r + 4.5 * r ** 2
    outer = -(2.0 - r) * (1.0 - r) - 0.5 * (2.0 - r) ** 2   # d/dr [0.5 (2-r)^2 (1-r)]
    return sg * np.where(r < 1.0, inner, np.where(r < 2.0, outer, 0.0))


def bsp2(r):
    r = np.abs(r)
    return np.where(r < 1.0, 1.0 - r, 0.0)


def dbsp2(s):
    return np.where(np.abs(s) < 1.0, -np.sign(s), 0.0)


def kernel_fn(kernel):
    return (m4, 2) if kernel == 2 else (bsp2, 1)


def p2m_periodic(xp, up, N, kernel=2, L=1.0, xmin=0.0):
    """Folded periodic field on the N^dim unique nodes, shape (N_z, N_y, N_x, lda) / (N_y, N_x, lda).

    m_n = sum_p u_p prod_d W(x0_pd - n_d) over all periodic images of n."""
    W, reach = kernel_fn(kernel)
    xp = np.asarray(xp, np.float64)
    up = np.asarray(up, np.float64)
    if up.ndim == 1:
        up = up[:, None]
    dim = xp.shape[1]
    N = [N] * dim if np.isscalar(N) else list(N)
    lda = up.shape[1]
    x0 = (xp - xmin) / (L / np.array(N, dtype=np.float64))
    out = np.zeros(tuple(reversed(N)) + (lda,))
    base = np.floor(x0).astype(np.int64)
    offs = np.arange(-reach, reach + 2)
    # 1D weights per particle and axis on nodes base+offs (wider than the support; zeros are harmless)
    w1 = [W(x0[:, d:d + 1] - (base[:, d:d + 1] + offs[None, :])) for d in range(dim)]
    nodes = [np.mod(base[:, d:d + 1] + offs[None, :], N[d]) for d in range(dim)]
    K = len(offs)
    if dim == 3:
        for a in range(K):
            for b in range(K):
                for c in range(K):
                    w = w1[0][:, a] * w1[1][:, b] * w1[2][:, c]
                    np.add.at(out, (nodes[2][:, c], nodes[1][:, b], nodes[0][:, a]), w[:, None] * up)
    else:
        for a in range(K):
            for b in range(K):
                w = w1[0][:, a] * w1[1][:, b]
                np.add.at(out, (nodes[1][:, b], nodes[0][:, a]), w[:, None] * up)
    return out


def m2p_periodic(xp, field, N, kernel=2, L=1.0, xmin=0.0):
    """u_p = sum_n W(x0_p - n) f_{n mod N}; field on unique nodes (N_z,N_y,N_x,lda)."""
    W, reach = kernel_fn(kernel)
    xp = np.asarray(xp, np.float64)
    dim = xp.shape[1]
    N = [N] * dim if np.isscalar(N) else list(N)
    x0 = (xp - xmin) / (L / np.array(N, dtype=np.float64))
    base = np.floor(x0).astype(np.int64)
    offs = np.arange(-reach, reach + 2)
    w1 = [W(x0[:, d:d + 1] - (base[:, d:d + 1] + offs[None, :])) for d in range(dim)]
    nodes = [np.mod(base[:, d:d + 1] + offs[None, :], N[d]) for d in range(dim)]
    K = len(offs)
    up = np.zeros((xp.shape[0], field.shape[-1]))
    if dim == 3:
        for a in range(K):
            for b in range(K):
                for c in range(K):
                    w = w1[0][:, a] * w1[1][:, b] * w1[2][:, c]
                    up += w[:, None] * field[nodes[2][:, c], nodes[1][:, b], nodes[0][:, a]]
    else:
        for a in range(K):
            for b in range(K):
                w = w1[0][:, a] * w1[1][:, b]
                up += w[:, None] * field[nodes[1][:, b], nodes[0][:, a]]
    return up


def rank_stable(xp, xmin, xmax, Ncells):
    """Cell keys x-fastest and the stable order: numpy's stable argsort (a library routine)."""
    xp = np.asarray(xp)
    dim = xp.shape[1]
    idx = [np.floor((xp[:, d] - xmin[d]) * (Ncells[d] / (xmax[d] - xmin[d]))).astype(np.int64)
           for d in range(dim)]
    key = idx[0] + Ncells[0] * idx[1]
    if dim == 3:
        key = key + Ncells[0] * Ncells[1] * idx[2]
    order = np.argsort(key, kind="stable")
    nbox = int(np.prod(Ncells))
    counts = np.bincount(key, minlength=nbox)
    heads = np.concatenate([[0], np.cumsum(counts)])
    return key, order, heads
