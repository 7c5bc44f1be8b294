"""Pins of the geometric forest oracle (oracle/forest_oracle.c) that do not
involve the product: the reference's own Morton sort (built from
src/TMROctant.cpp into oracle/_ref), closed-form node counts, 2:1 balance and
minimality checked leaf by leaf, hanging-node weight rules, and exact
reproduction of trilinear fields by the interpolation."""
import itertools

import numpy as np
import pytest

from oracle import pyoracle as po

H = 1 << 30


def rand_octants(rng, n, nblocks, maxlev):
    lev = rng.integers(0, maxlev + 1, n)
    out = np.zeros(n, dtype=po.OCT_DT)
    out["block"] = rng.integers(0, nblocks, n)
    out["level"] = lev
    for a in "xyz":
        h = (1 << (30 - lev)).astype(np.int64)
        out[a] = (rng.integers(0, 1 << 20, n) % (1 << lev)) * h
    return out


@pytest.mark.parametrize("use_node", [0, 1])
def test_sort_matches_reference_octant_array(use_node):
    """TMROctantArray::sort (src/TMROctant.cpp:357-399) from the reference."""
    R = po.ref_octant_lib()
    if R is None:
        pytest.skip("reference octant library not built (oracle/_ref)")
    assert R.ref_sizeof_octant() == 24
    rng = np.random.default_rng(11 + use_node)
    for trial in range(5):
        a = rand_octants(rng, 3000, 3, 6)
        # duplicates of positions at several levels
        a = np.concatenate([a, a[:500]])
        if use_node:
            a["info"] = rng.integers(0, 3, len(a))
            a["x"][:200] = H - 1  # the node truncation value
        ref = np.ascontiguousarray(a.copy())
        n = R.ref_sort_octants(ref.ctypes.data, len(ref), use_node)
        mine = po.sort_octants(a, use_node=bool(use_node))
        key = ["block", "x", "y", "z", "level"] + (["info"] if use_node else [])
        assert n == len(mine)
        assert np.array_equal(ref[:n][key], mine[key])


def test_child_order_matches_reference():
    R = po.ref_octant_lib()
    if R is None:
        pytest.skip("reference octant library not built (oracle/_ref)")
    parent_level = 3
    h = 1 << (30 - parent_level - 1)
    ch = np.zeros(8, dtype=po.OCT_DT)
    for c in range(8):
        ch[c] = (0, (c & 1) * h, ((c >> 1) & 1) * h, (c >> 2) * h, 0, parent_level + 1, 0)
        assert R.ref_child_id(np.ascontiguousarray(ch[c:c + 1]).ctypes.data) == c
    s = po.sort_octants(ch)
    order = [((o["x"] > 0) + 2 * (o["y"] > 0) + 4 * (o["z"] > 0)) for o in s]
    assert order == [0, 4, 2, 6, 1, 5, 3, 7]


def uniform_octants(dims, level):
    n = 1 << level
    h = 1 << (30 - level)
    out = []
    for b in range(dims[0] * dims[1] * dims[2]):
        for x, y, z in itertools.product(range(n), repeat=3):
            out.append((b, x * h, y * h, z * h, 0, level, 0))
    return np.array(out, dtype=po.OCT_DT)


@pytest.mark.parametrize("dims,level", [((1, 1, 1), 3), ((2, 1, 1), 2), ((2, 2, 2), 2)])
def test_uniform_node_counts(dims, level):
    m = po.OracleMesh(dims, uniform_octants(dims, level)).create_nodes()
    n = 1 << level
    assert len(m.leaves) == n ** 3 * np.prod(dims)
    assert m.ndep == 0
    assert m.nindep == (dims[0] * n + 1) * (dims[1] * n + 1) * (dims[2] * n + 1)


def _leaf_boxes(m):
    L = m.leaves
    dims = m.dims
    b = L["block"].astype(np.int64)
    bi, bj, bk = b % dims[0], (b // dims[0]) % dims[1], b // (dims[0] * dims[1])
    lo = np.stack([(bi << 30) + L["x"], (bj << 30) + L["y"], (bk << 30) + L["z"]], 1)
    h = (np.int64(1) << (30 - L["level"].astype(np.int64)))
    return lo, h


def _touching(lo, h):
    """pairs of leaves whose closed boxes intersect (face/edge/corner)"""
    n = len(h)
    pairs = []
    for i in range(n):
        hi_i = lo[i] + h[i]
        hi = lo + h[:, None]
        ok = np.all((lo <= hi_i) & (hi >= lo[i]), axis=1)
        ok[i] = False
        for j in np.nonzero(ok)[0]:
            if j > i:
                # classify the contact dimension: 3 - (# axes with a shared face)
                touch = np.sum((lo[j] == hi_i) | (hi[j] == lo[i]))
                pairs.append((i, j, touch))
    return pairs


@pytest.mark.parametrize("seed,corner", [(0, 1), (1, 1), (2, 0)])
def test_balance_is_complete_balanced_and_minimal(seed, corner):
    rng = np.random.default_rng(seed)
    dims = (2, 1, 1)
    base = uniform_octants(dims, 1)
    # child-0 insertion as TMROctForest::refine does
    pick = rng.choice(len(base), 4, replace=False)
    extra = []
    for p in pick:
        o = base[p]
        lev = int(rng.integers(3, 5))
        hn = 1 << (30 - lev)
        extra.append((o["block"], o["x"], o["y"], o["z"], 0, lev, 0))
    inp = np.concatenate([base, np.array(extra, dtype=po.OCT_DT)])
    m = po.OracleMesh(dims, inp, corner)
    lo, h = _leaf_boxes(m)
    # complete, non-overlapping: volumes add up to the blocks
    assert np.sum(h.astype(float) ** 3) == pytest.approx(np.prod(dims) * float(H) ** 3, rel=0)
    lev = m.leaves["level"].astype(int)
    pairs = _touching(lo, h)
    for i, j, touch in pairs:
        if touch == 3 and not corner:
            continue  # corner contact is not balanced by balance(0)
        assert abs(lev[i] - lev[j]) <= 1
    # minimality: every refined octant that the input does not force would,
    # if its 8 children were merged back, touch a leaf two levels finer
    forced = set()
    for o in inp:
        l = int(o["level"])
        x, y, z = int(o["x"]), int(o["y"]), int(o["z"])
        while l > 0:
            l -= 1
            hh = 1 << (30 - l)
            x, y, z = x & ~(hh - 1), y & ~(hh - 1), z & ~(hh - 1)
            forced.add((int(o["block"]), x, y, z, l))
    fam = {}
    for k, o in enumerate(m.leaves):
        l = int(o["level"])
        if l == 0:
            continue
        hh = 1 << (30 - l + 1)
        key = (int(o["block"]), int(o["x"]) & ~(hh - 1), int(o["y"]) & ~(hh - 1), int(o["z"]) & ~(hh - 1), l - 1)
        fam.setdefault(key, []).append(k)
    for key, kids in fam.items():
        if len(kids) != 8 or key in forced:
            continue
        b, x, y, z, l = key
        bi = b % dims[0]
        plo = np.array([(bi << 30) + x, y, z], dtype=np.int64)
        ph = np.int64(1) << (30 - l)
        need = False
        for j in range(len(h)):
            if j in kids:
                continue
            if np.all((lo[j] <= plo + ph) & (lo[j] + h[j] >= plo)):
                touch = np.sum((lo[j] == plo + ph) | (lo[j] + h[j] == plo))
                if touch == 3 and not corner:
                    continue
                if lev[j] >= l + 2:
                    need = True
                    break
        assert need, "refinement of %s is not required" % (key,)


def test_hanging_weights_and_parents():
    rng = np.random.default_rng(5)
    dims = (2, 2, 2)
    base = uniform_octants(dims, 1)
    extra = [(o["block"], o["x"], o["y"], o["z"], 0, 3, 0) for o in base[rng.choice(len(base), 6, replace=False)]]
    m = po.OracleMesh(dims, np.concatenate([base, np.array(extra, dtype=po.OCT_DT)])).create_nodes()
    assert m.ndep > 0
    lens = np.diff(m.dep_ptr)
    assert set(lens) <= {2, 4}
    for d in range(m.ndep):
        w = m.dep_w[m.dep_ptr[d]:m.dep_ptr[d + 1]]
        assert np.all(w == 1.0 / len(w))
        assert np.all(m.dep_conn[m.dep_ptr[d]:m.dep_ptr[d + 1]] >= 0)


def test_interpolation_reproduces_trilinear_fields():
    rng = np.random.default_rng(9)
    dims = (2, 2, 2)
    base = uniform_octants(dims, 1)
    extra = [(o["block"], o["x"], o["y"], o["z"], 0, 3, 0) for o in base[rng.choice(len(base), 5, replace=False)]]
    fine_in = np.concatenate([base, np.array(extra, dtype=po.OCT_DT)])
    fm0 = po.OracleMesh(dims, fine_in)
    # coarse = parents of the 0-children (TMROctForest::coarsen)
    coarse = []
    for o in fm0.leaves:
        l = int(o["level"])
        if l == 0:
            coarse.append(tuple(o))
            continue
        h = 1 << (30 - l)
        if (o["x"] & h) or (o["y"] & h) or (o["z"] & h):
            continue
        coarse.append((o["block"], o["x"], o["y"], o["z"], 0, l - 1, 0))
    fm, cm, (rowp, cols, w) = po.interpolation(dims, fine_in, np.array(coarse, dtype=po.OCT_DT))
    sums = np.add.reduceat(w, rowp[:-1])
    assert np.all(sums == 1.0)
    a = np.array([0.3, -1.25, 2.0, 0.5, 0.75, -0.4, 1.1, -0.6])

    # a trilinear field (cross terms xy, yz, xz and xyz included) at the
    # independent nodes via the mesh conn/leaves: trilinear in every box
    # element, so the interpolation must reproduce it exactly
    def field(p):
        x, y, z = p / float(H)
        return (a[0] + a[1] * x + a[2] * y + a[3] * z + a[4] * x * y + a[5] * y * z +
                a[6] * x * z + a[7] * x * y * z)

    def field_at_indep(mesh):
        lo, h = _leaf_boxes(mesh)
        vals = {}
        for e in range(len(mesh.leaves)):
            for c in range(8):
                n = mesh.conn[e, c]
                if n < 0:
                    continue
                p = lo[e] + np.array([c & 1, (c >> 1) & 1, c >> 2]) * h[e]
                vals[n] = field(p)
        return np.array([vals[i] for i in range(mesh.nindep)])
    uc = field_at_indep(cm)
    uf = field_at_indep(fm)
    Pu = np.add.reduceat(w * uc[cols], rowp[:-1])
    assert np.max(np.abs(Pu - uf)) < 1e-13
