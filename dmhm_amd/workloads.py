"""Synthetic operators for the H-matrix path, generated on the host in FP64.

The same bytes feed the GPU library and the CPU oracle. The stencil is the
flux form of the reference's inversion test program
(tests/HMat2d/Invert.cpp:12-67): the
coefficient lives on the (n+2)^d node grid including the Dirichlet boundary
nodes, each face coefficient is the mean of the two node values it joins,
off-diagonals are -face/h^2 and the diagonal is the sum of the face
coefficients / h^2. Two deliberate differences from that program:
h = 1/(n+1) (it uses 1/n) and a +u mass term (+1 on the diagonal; its
DomainV is zero). Rows and columns are in the hierarchical ordering of
HMat2d::BuildNaturalToHierarchicalMap (src/HMat2d/HMat2d.cpp:27-87) /
HMat3d (src/HMat3d/HMat3d.cpp:28-90).
"""
import numpy as np


def natural_to_hierarchical_map(sizes, num_levels):
    """map[natural] = hierarchical index; natural = x + nx*y (+ nx*ny*z).

    Recursion over boxes: split each axis into size//2 (low half) and the
    rest, visit children in the order c = 0..2^d-1 with axis bit a = (c>>a)&1
    (lower-left, lower-right, upper-left, upper-right, then the back octants),
    and number the points of each leaf box x-fastest.
    """
    sizes = list(sizes)
    dim = len(sizes)
    full = sizes + [1] * (3 - dim)
    grid = np.zeros(full[::-1], dtype=np.int64)      # indexed [z, y, x]
    counter = [0]

    def rec(level, off, box):
        if level == num_levels - 1:
            cnt = box[0] * box[1] * box[2]
            block = np.arange(counter[0], counter[0] + cnt, dtype=np.int64).reshape(box[2], box[1], box[0])
            grid[off[2]:off[2] + box[2], off[1]:off[1] + box[1], off[0]:off[0] + box[0]] = block
            counter[0] += cnt
            return
        for c in range(1 << dim):
            sub, o = [1, 1, 1], list(off)
            for a in range(dim):
                half = box[a] // 2
                hi = (c >> a) & 1
                sub[a] = box[a] - half if hi else half
                o[a] = off[a] + (half if hi else 0)
            rec(level + 1, o, sub)

    rec(0, [0, 0, 0], full)
    return grid.reshape(-1).astype(np.int32)


def coefficient_smooth_2d(n):
    """a(x,y) = 1 + 0.5 sin(2 pi x) cos(2 pi y) on the (n+2)^2 node grid."""
    h = 1.0 / (n + 1)
    x = np.arange(n + 2) * h
    X, Y = np.meshgrid(x, x, indexing="xy")          # [y, x]
    return 1.0 + 0.5 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)


def coefficient_lognormal(shape_nodes, sigma, corr_len, seed, features=2048):
    """log a = sigma sqrt(2/M) sum_j cos(w_j . x + b_j), w_j ~ N(0, corr^-2 I),
    b_j ~ U[0, 2 pi): random Fourier features of a squared-exponential field,
    evaluated on the node grid (closed form, platform independent in FP64)."""
    dim = len(shape_nodes)
    rng = np.random.Generator(np.random.PCG64(seed))
    w = rng.normal(0.0, 1.0 / corr_len, size=(features, dim))
    b = rng.uniform(0.0, 2 * np.pi, size=features)
    axes = [np.arange(s) / (s - 1) for s in shape_nodes]     # node coordinates in [0,1]
    mesh = np.meshgrid(*axes[::-1], indexing="ij")            # [z,] y, x order
    pts = np.stack([m.reshape(-1) for m in mesh[::-1]], axis=1)   # columns x, y(, z)
    out = np.zeros(pts.shape[0])
    chunk = 4096
    for i in range(0, pts.shape[0], chunk):
        out[i:i + chunk] = np.cos(pts[i:i + chunk] @ w.T + b).sum(axis=1)
    loga = sigma * np.sqrt(2.0 / features) * out
    return np.exp(loga).reshape([s for s in shape_nodes[::-1]])


def coefficient_uniform_3d(n, seed):
    """a = 1 + 0.5 U[0,1) i.i.d. per node of the (n+2)^3 grid."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return 1.0 + 0.5 * rng.random((n + 2, n + 2, n + 2))


def flux_operator(a_nodes, num_levels, mass=1.0):
    """CSR (hierarchical ordering) of -div(a grad u) + mass*u with Dirichlet
    boundary; a_nodes is indexed [y, x] (2D) or [z, y, x] (3D) with shape
    (n+2, ...). Returns (rowptr, colind, vals, sizes, map)."""
    dim = a_nodes.ndim
    sizes = [s - 2 for s in a_nodes.shape[::-1]]          # nx, ny(, nz)
    n_pts = int(np.prod(sizes))
    h = 1.0 / (sizes[0] + 1)
    hh = h * h
    hmap = natural_to_hierarchical_map(sizes, num_levels)
    inv = np.empty(n_pts, dtype=np.int64)
    inv[hmap] = np.arange(n_pts)
    nat = inv                                              # natural index of hierarchical row i
    coords = []
    rem = nat.copy()
    for a in range(dim):
        coords.append(rem % sizes[a])
        rem //= sizes[a]

    def a_at(c):                                           # node value at unknown coords + 1
        idx = tuple((c[a] + 1) for a in reversed(range(dim)))
        return a_nodes[idx]

    self_a = a_at(coords)
    diag = np.full(n_pts, mass, dtype=np.float64)
    cols, vals = [], []
    for a in range(dim):
        for d in (-1, 1):
            nb = [c.copy() for c in coords]
            nb[a] = nb[a] + d
            face = 0.5 * (self_a + a_at(nb)) / hh          # boundary nodes exist in a_nodes
            diag += face
            inside = (nb[a] >= 0) & (nb[a] < sizes[a])
            nat_nb = np.zeros(n_pts, dtype=np.int64)
            stride = 1
            for b in range(dim):
                nat_nb += np.clip(nb[b], 0, sizes[b] - 1) * stride
                stride *= sizes[b]
            col = np.where(inside, hmap[nat_nb], -1)
            cols.append(col)
            vals.append(np.where(inside, -face, 0.0))
    cols.insert(0, np.arange(n_pts))
    vals.insert(0, diag)
    C = np.stack(cols, axis=1)
    V = np.stack(vals, axis=1)
    mask = C >= 0
    counts = mask.sum(axis=1)
    rowptr = np.zeros(n_pts + 1, dtype=np.int32)
    rowptr[1:] = np.cumsum(counts)
    colind = C[mask].astype(np.int32)
    values = V[mask].astype(np.float64)
    return rowptr, colind, values, sizes, hmap


def config_operator(name, seed=0):
    """The named synthetic operators (dim, sizes, numLevels, maxRank, CSR)."""
    if name == "c1":        # 2D 64^2, smooth coefficient, L=4, maxRank 8
        a = coefficient_smooth_2d(64)
        L, r = 4, 8
    elif name == "c2":      # 2D 256^2, lognormal sigma=1, l=0.1, L=6, maxRank 16
        a = coefficient_lognormal((258, 258), 1.0, 0.1, seed)
        L, r = 6, 16
    elif name == "c4":      # 2D 512^2, lognormal sigma=1.5, l=0.05, L=7, maxRank 32
        a = coefficient_lognormal((514, 514), 1.5, 0.05, seed)
        L, r = 7, 32
    elif name == "c3":      # 3D 32^3, a = 1 + 0.5 U, L=4, maxRank 32
        a = coefficient_uniform_3d(32, seed)
        L, r = 4, 32
    elif name == "c5":      # 3D 64^3, lognormal sigma=1, l=0.1, L=5, maxRank 32
        a = coefficient_lognormal((66, 66, 66), 1.0, 0.1, seed)
        L, r = 5, 32
    else:
        raise ValueError("unknown operator " + name)
    rowptr, colind, vals, sizes, hmap = flux_operator(a, L)
    return dict(dim=a.ndim, sizes=sizes, num_levels=L, max_rank=r, rowptr=rowptr,
                colind=colind, vals=vals, map=hmap)


def c3_batch(count=48, seed=0, sides=(16, 20, 24, 28, 32), num_levels=4, max_rank=32):
    """The batch of independent 3D operators: -div(a grad u) + u on n^3 grids,
    n drawn uniformly from `sides`, a = 1 + 0.5 U[0,1) i.i.d. per cell, each
    with its own seed; HMat3d numLevels 4, maxRank 32."""
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.asarray(sides), size=count)
    out = []
    for i, n in enumerate(ns):
        a = coefficient_uniform_3d(int(n), seed * 1000 + i)
        rowptr, colind, vals, sizes, hmap = flux_operator(a, num_levels)
        out.append(dict(dim=3, sizes=sizes, num_levels=num_levels, max_rank=max_rank, rowptr=rowptr,
                        colind=colind, vals=vals, map=hmap))
    return out


def small_operator(n, num_levels, dim=2, smooth=True, seed=0):
    """A small operator of the same family for tests (n per axis)."""
    if dim == 2:
        a = coefficient_smooth_2d(n) if smooth else coefficient_lognormal((n + 2, n + 2), 1.0, 0.1, seed)
    else:
        a = coefficient_uniform_3d(n, seed)
    rowptr, colind, vals, sizes, hmap = flux_operator(a, num_levels)
    return dict(dim=dim, sizes=sizes, num_levels=num_levels, rowptr=rowptr, colind=colind,
                vals=vals, map=hmap)


def rect_operator(sizes, num_levels, seed=0):
    """Operator on a rectangular grid (per-axis sizes) with a = 1 + 0.5 U[0,1)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a = 1.0 + 0.5 * rng.random([s + 2 for s in sizes[::-1]])
    rowptr, colind, vals, sz, hmap = flux_operator(a, num_levels)
    return dict(dim=len(sizes), sizes=sz, num_levels=num_levels, rowptr=rowptr, colind=colind, vals=vals,
                map=hmap)


def csr_to_dense(rowptr, colind, vals, n):
    D = np.zeros((n, n))
    for i in range(n):
        D[i, colind[rowptr[i]:rowptr[i + 1]]] += vals[rowptr[i]:rowptr[i + 1]]
    return D
