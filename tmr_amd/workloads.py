"""Synthetic workloads of the cantilever SIMP problems (seeded, no data files).

* cube_l4     : unit cube, 1 block, uniform level 4 (4,096 hex), rho = 0.5,
                3 multigrid levels;
* box_256     : 2 x 1 x 1 box, 2 blocks at level 7 (256 x 128 x 128 =
                4,194,304 hex, 12,830,211 dofs), seeded smoothed densities,
                6 multigrid levels (coarsest 8 x 4 x 4);
* adaptive    : 2 x 1 x 1 box of 2 x 2 x 2 blocks, per-octant levels from a
                seeded Gaussian-sum density field, 2:1 balanced.
All clamp the face x = 0 and load the face x = Lx with a unit traction -z.
"""
import numpy as np

from . import TMR

FACES = {"x-": "fixed", "x+": "load"}


def hierarchy(forest, nlevels, device=None):
    """Finest-first list of forests: coarsen + balance(1) per level
    (TopOptUtils.createTopoProblem, tmr/TopOptUtils.py:69-99); with `device`
    (a GPU index) the balance and node passes run on that GPU."""
    forest.balance(1, device=device)
    forests = [forest]
    for _ in range(nlevels - 1):
        c = forests[-1].coarsen()
        c.balance(1, device=device)
        forests.append(c)
    for f in forests:
        f.createNodes(device=device)
    return forests


def smoothed_density(shape, seed, passes=3, lo=1e-3, hi=1.0):
    """Uniform [lo, hi] densities, then `passes` passes of a 27-point box mean
    (boundary-truncated), clipped to [lo, hi].  shape = (nz, ny, nx)."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(lo, hi, size=shape)
    for _ in range(passes):
        s = np.zeros_like(r)
        c = np.zeros_like(r)
        ones = np.ones_like(r)
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    sl_dst, sl_src = [], []
                    for d, n in zip((dz, dy, dx), shape):
                        if d < 0:
                            sl_dst.append(slice(0, n + d)); sl_src.append(slice(-d, n))
                        elif d > 0:
                            sl_dst.append(slice(d, n)); sl_src.append(slice(0, n - d))
                        else:
                            sl_dst.append(slice(0, n)); sl_src.append(slice(0, n))
                    s[tuple(sl_dst)] += r[tuple(sl_src)]
                    c[tuple(sl_dst)] += ones[tuple(sl_src)]
        r = s / c
    return np.clip(r, lo, hi)


def element_density_from_grid(forest, grid, L):
    """Sample a (nz, ny, nx) cell grid over the box [0,L] at element centres."""
    X = forest._Xv
    octs = forest.getOctants()
    bc = forest.getConnectivity()["block_conn"]
    x0 = X[bc[octs["block"], 0]]
    ext = X[bc[:, 7]] - X[bc[:, 0]]
    h = ext[octs["block"]] * np.ldexp(1.0, -octs["level"].astype(np.int64))[:, None]
    u = np.stack([octs["x"], octs["y"], octs["z"]], 1) / float(1 << 30)
    c = x0 + u * ext[octs["block"]] + 0.5 * h
    nz, ny, nx = grid.shape
    i = np.minimum((c[:, 0] / L[0] * nx).astype(np.int64), nx - 1)
    j = np.minimum((c[:, 1] / L[1] * ny).astype(np.int64), ny - 1)
    k = np.minimum((c[:, 2] / L[2] * nz).astype(np.int64), nz - 1)
    return grid[k, j, i]


def cube_l4(nlevels=3):
    f = TMR.createBoxForest(1, 1, 1, (1.0, 1.0, 1.0), face_names=FACES)
    f.createTrees(4)
    forests = hierarchy(f, nlevels)
    rho = np.full(len(forests[0].getOctants()), 0.5)
    return forests, rho, (1.0, 1.0, 1.0)


def box(level=7, nlevels=6, seed=0, passes=3, device=None):
    L = (2.0, 1.0, 1.0)
    f = TMR.createBoxForest(2, 1, 1, L, face_names=FACES)
    f.createTrees(level)
    forests = hierarchy(f, nlevels, device=device)
    n = 1 << level
    grid = smoothed_density((n, n, 2 * n), seed, passes)
    rho = element_density_from_grid(forests[0], grid, L)
    return forests, rho, L


def gaussian_field(pts, seed, L=(2.0, 1.0, 1.0), ngauss=32):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 1, size=(ngauss, 3)) * np.array(L)
    sig = rng.uniform(0.05, 0.2, size=ngauss)
    amp = rng.uniform(0.0, 1.0, size=ngauss)
    d2 = ((pts[:, None, :] - c[None, :, :]) ** 2).sum(-1)
    return np.clip((amp[None, :] * np.exp(-0.5 * d2 / sig[None, :] ** 2)).sum(1), 1e-3, 1.0)


def _centres(f, L):
    octs = f.getOctants()
    X = f._Xv
    bc = f.getConnectivity()["block_conn"]
    ext = X[bc[:, 7]] - X[bc[:, 0]]
    h = ext[octs["block"]] * np.ldexp(1.0, -octs["level"].astype(np.int64))[:, None]
    u = np.stack([octs["x"], octs["y"], octs["z"]], 1) / float(1 << 30)
    return X[bc[octs["block"], 0]] + u * ext[octs["block"]] + 0.5 * h, h, octs


def _field_gradient(c, h, seed, L):
    """|grad rho| at the points c by central differences of step h/4."""
    eps = 0.25 * h
    g = np.zeros(len(c))
    for a in range(3):
        dp = c.copy(); dm = c.copy()
        dp[:, a] += eps[:, a]; dm[:, a] -= eps[:, a]
        g += ((gaussian_field(dp, seed, L) - gaussian_field(dm, seed, L)) / (2 * eps[:, a])) ** 2
    return np.sqrt(g)


def adaptive(seed=0, base_level=3, max_level=7, nlevels=5, L=(2.0, 1.0, 1.0), max_sweeps=None,
             base_median=False, quantile=0.5, face_names=FACES, device=None):
    """Adaptive forest on 2x2x2 blocks: octants where the seeded density has a
    gradient above its median are refined towards a seeded level in
    [base_level, max_level]; balance(1) after each sweep.  The median is the
    one of the gradient over the current octants (base_median False) or over
    the centres of the uniform base-level octants, fixed for every sweep
    (base_median True: a median of the field itself, the C2 rule); with
    base_median, `quantile` replaces the median by another quantile."""
    rng = np.random.default_rng(seed + 1000)
    f = TMR.createBoxForest(2, 2, 2, L, face_names=face_names)
    f.createTrees(base_level)
    sweeps = max_level - base_level if max_sweeps is None else max_sweeps
    c, h, _ = _centres(f, L)
    med0 = np.quantile(_field_gradient(c, h, seed, L), quantile)
    for _ in range(sweeps):
        c, h, octs = _centres(f, L)
        g = _field_gradient(c, h, seed, L)
        target = rng.integers(base_level, max_level + 1, size=len(octs))
        med = med0 if base_median else np.median(g)
        ref = ((g > med) & (octs["level"] < target)).astype(np.intc)
        if not ref.any():
            break
        f.refine(ref, 0, max_level)
        f.balance(1, device=device)
    forests = hierarchy(f, nlevels, device=device)
    cen, _, _ = _centres(forests[0], L)
    rho = gaussian_field(cen, seed, L)
    return forests, rho, L


def c2_adaptive(seed=0, nlevels=5, device=None):
    """The adaptive cantilever of about a million hex (1,227,458 at seed 0):
    2 x 2 x 2 blocks on a 2 x 1 x 1 box, base level 3, per-octant target
    levels drawn from the seed in 3..7, refined where the gradient of the
    32-Gaussian density exceeds its median, balance(1) after each sweep;
    `nlevels` multigrid forests."""
    return adaptive(seed=seed, base_level=3, max_level=7, nlevels=nlevels, base_median=True,
                    device=device)


ALL_FACES = {"x-": "fixed", "x+": "load", "y-": "y-", "y+": "y+", "z-": "z-", "z+": "z+"}


def c5a_adaptive(seed=1, nlevels=5, device=None):
    """The multi-load-case forest: the C2 generator at the 0.1 quantile of the
    field gradient (2,160,152 hex at seed 1), every outer face named
    (x- clamped)."""
    return adaptive(seed=seed, base_level=3, max_level=7, nlevels=nlevels, base_median=True,
                    quantile=0.1, face_names=ALL_FACES, device=device)


def face_load(forest, faces, tr):
    """Consistent nodal load of a constant traction tr on element faces
    [(element, local face), ...]: the exact bilinear face integral t * area / 4
    per face corner, hanging corners passing their share to their parents
    (the rule of TopOptUtils.compute3DTractionLoad on box faces)."""
    conn = forest.getMeshConn()
    ext = forest.getElementExtents()
    dp, dc, dw = forest.getDepNodeConn()
    n = int(forest.getNodeRange()[1])
    f = np.zeros((n, 3))
    tr = np.asarray(tr, dtype=np.double)
    for e, face in faces:
        ax = face // 2
        area = ext[e][(ax + 1) % 3] * ext[e][(ax + 2) % 3]
        for k in range(8):
            if [(k & 1), (k >> 1) & 1, (k >> 2)][ax] != face % 2:
                continue
            g = conn[e][k]
            if g >= 0:
                f[g] += 0.25 * area * tr
            else:
                d = -g - 1
                for j in range(dp[d], dp[d + 1]):
                    f[dc[j]] += dw[j] * 0.25 * area * tr
    return f.reshape(-1)


def patch_loads(forest, count, seed, L=(2.0, 1.0, 1.0), names=("x+", "y-", "y+", "z-", "z+")):
    """`count` load cases, each a unit traction in a seeded direction on a
    seeded disc (radius 0.1 .. 0.3) of one seeded outer face (never the
    clamped one).  Returns F [count, ndof] and the patch parameters."""
    rng = np.random.default_rng(seed)
    octs = forest.getOctants()
    cen, h, _ = _centres(forest, L)
    side = {"x-": (0, 0), "x+": (0, 1), "y-": (1, 0), "y+": (1, 1), "z-": (2, 0), "z+": (2, 1)}
    faces_of = {nm: [(int(o["tag"]), int(o["info"])) for o in forest.getOctsWithName(nm)]
                for nm in names}
    F, params = [], []
    for _ in range(count):
        while True:
            nm = names[rng.integers(len(names))]
            ax, hi = side[nm]
            c = rng.uniform(0.0, 1.0, 3) * np.array(L)
            c[ax] = L[ax] if hi else 0.0
            r = rng.uniform(0.1, 0.3)
            t = rng.standard_normal(3)
            t /= np.linalg.norm(t)
            sel = []
            for e, fi in faces_of[nm]:
                p = cen[e].copy()
                p[ax] = c[ax]
                if np.linalg.norm(p - c) <= r:
                    sel.append((e, fi))
            if sel:
                break
        F.append(face_load(forest, sel, t))
        params.append({"face": nm, "centre": c.tolist(), "radius": r, "traction": t.tolist(),
                       "element_faces": len(sel)})
    return np.array(F), params


def c5b_edges(count=96, seed=0, lo=8, hi=48):
    """Edge lengths (elements) of the uneven batch: every integer in
    [lo, hi] equally likely, drawn from the seed."""
    return np.random.default_rng(seed + 77).integers(lo, hi + 1, size=count)


def level_edges(n, nlevels):
    """Elements per edge on each level of a cube of n elements: n, then
    halved and rounded up (13, 7, 4, 2)."""
    ms = [int(n)]
    for _ in range(nlevels - 1):
        ms.append(max(1, (ms[-1] + 1) // 2))
    return ms


def odd_split(m):
    """(b, L) with m = b 2^L and b odd: a lattice of b^3 blocks at depth L
    meshes a cube with m elements per edge."""
    m, L = int(m), 0
    while m % 2 == 0:
        m //= 2
        L += 1
    return m, L


def uneven_cube_batch(edges, nlevels=4, h=0.125, gap=1.0, seed=0, passes=2, device=None):
    """Independent cantilever cubes of `edges` elements per edge (elements of
    size h on the finest level), side by side along x, each clamped on its
    x- face ("fixed") and loaded on its x+ face ("load").  Level l meshes a
    cube of n with level_edges(n)[l] elements per edge as a lattice of b^3
    blocks at depth L (odd_split), a forest of its own per level: an edge
    such as 13 has no coarsen() level (13 level-0 blocks), and its coarser
    levels (7, 4, 2 elements) are not nested in it; the problem builds their
    transfers from physical positions.  Returns (forests, rho) with seeded
    smoothed densities (seed + k for cube k)."""
    edges = [int(n) for n in edges]
    forests = []
    for l in range(nlevels):
        conns, Xs, names, blev = [], [], [], []
        off_v, off_b, x0 = 0, 0, 0.0
        for n in edges:
            b, L = odd_split(level_edges(n, nlevels)[l])
            E = n * h
            conn, X = TMR.box_connectivity(b, b, b, L=(E, E, E))
            X[:, 0] += x0
            conns.append(conn + off_v)
            Xs.append(X)
            q = np.arange(len(conn))
            names += [(off_b + int(v), 0, "fixed") for v in q[q % b == 0]]
            names += [(off_b + int(v), 1, "load") for v in q[q % b == b - 1]]
            blev.append(np.full(len(conn), L, dtype=np.intc))
            off_v += len(X)
            off_b += len(conn)
            x0 += E + gap
        f = TMR.OctForest()
        f.setConnectivity(np.vstack(conns).astype(np.intc))
        f.setNodePoints(np.vstack(Xs))
        bfc = f.getConnectivity()["block_face_conn"]
        for q, fi, name in names:
            f.setName("face", int(bfc[q, fi]), name)
        f.createTrees(block_levels=np.concatenate(blev))
        forests.append(f)
    for f in forests:
        f.createNodes(device=device)
    octs = forests[0].getOctants()
    rho = np.empty(len(octs))
    off_b = 0
    for k, n in enumerate(edges):
        b, L = odd_split(n)
        nb = b ** 3
        sel = (octs["block"] >= off_b) & (octs["block"] < off_b + nb)
        grid = smoothed_density((n, n, n), seed + k, passes=passes)
        q = octs["block"][sel] - off_b
        hh = 1 << (30 - L)
        ix = (q % b) * (1 << L) + octs["x"][sel] // hh
        iy = ((q // b) % b) * (1 << L) + octs["y"][sel] // hh
        iz = (q // (b * b)) * (1 << L) + octs["z"][sel] // hh
        rho[sel] = grid[iz, iy, ix]
        off_b += nb
    return forests, rho


def box_batch(dims_list, level, nlevels, h=1.0, gap=1.0, seed=0, passes=2):
    """Independent cantilevers solved side by side: one forest holding a
    disjoint box lattice per problem (dims in blocks of edge h, each block at
    `level`), placed along x with a gap.  Every problem clamps its own x-
    face ("fixed") and is loaded on its x+ face ("load").  Returns (forests,
    rho) with seeded smoothed densities per problem."""
    conns, Xs, names = [], [], []
    off_v, off_b, x0 = 0, 0, 0.0
    for (nx, ny, nz) in dims_list:
        conn, X = TMR.box_connectivity(nx, ny, nz, L=(nx * h, ny * h, nz * h))
        X = X.copy()
        X[:, 0] += x0
        conns.append(conn + off_v)
        Xs.append(X)
        for b in range(len(conn)):
            i = b % nx
            if i == 0:
                names.append((off_b + b, 0, "fixed"))
            if i == nx - 1:
                names.append((off_b + b, 1, "load"))
        off_v += len(X)
        off_b += len(conn)
        x0 += nx * h + gap
    f = TMR.OctForest()
    f.setConnectivity(np.vstack(conns).astype(np.intc))
    f.setNodePoints(np.vstack(Xs))
    bfc = f.getConnectivity()["block_face_conn"]
    for b, fi, name in names:
        f.setName("face", int(bfc[b, fi]), name)
    f.createTrees(level)
    forests = hierarchy(f, nlevels)
    octs = forests[0].getOctants()
    n = 1 << level
    rho = np.empty(len(octs))
    hh = 1 << (30 - level)
    off_b = 0
    for k, (nx, ny, nz) in enumerate(dims_list):
        nb = nx * ny * nz
        sel = (octs["block"] >= off_b) & (octs["block"] < off_b + nb)
        grid = smoothed_density((nz * n, ny * n, nx * n), seed + k, passes=passes)
        b = octs["block"][sel] - off_b
        ix = (b % nx) * n + octs["x"][sel] // hh
        iy = ((b // nx) % ny) * n + octs["y"][sel] // hh
        iz = (b // (nx * ny)) * n + octs["z"][sel] // hh
        rho[sel] = grid[iz, iy, ix]
        off_b += nb
    return forests, rho


def cube_batch(count=96, nlevels=4, seed=0, device=None):
    """The uneven-size batch of independent problems: `count` cubes with
    edges drawn from the seed in 8..48 (c5b_edges), each with a hierarchy
    of its own (uneven_cube_batch); returns (forests, rho, edges)."""
    edges = [int(n) for n in c5b_edges(count, seed)]
    forests, rho = uneven_cube_batch(edges, nlevels, seed=seed, device=device)
    return forests, rho, edges
