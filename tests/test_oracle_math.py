"""Checks that pin the oracle by mathematics rather than by the reference:
closed-form Morton id arithmetic, brute-force trees on tiny inputs, the
M2M-equals-S2M identity, the single-blob velocity in closed form, and the
treecode truncation bound.  Each is chosen so that a plausible mistake (a
dropped term, a wrong sign or index, a transposed operand) fails it."""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import rel_err, workloads


# ---- Morton ids (morton_order.c), known answers from closed forms ---------
def test_morton_known_answers(oracle):
    L = oracle.lib()
    c = (C.c_int64 * 4)()
    L.orc_get_child(2, c)  # node 2 = level 1, local 1 -> level-2 locals 4..7, offset 5
    assert list(c) == [9, 10, 11, 12]
    assert L.orc_get_parent(19) == 4      # level 2 local 14 -> level 1 local 3 -> 4
    assert L.orc_get_level(13) == 2
    assert L.orc_global2local(13, 2) == 8
    assert L.orc_local2global(8, 2) == 13
    assert L.orc_quadrant_pos(np.array([1.0, -1.0]), np.array([0.0, 0.0])) == 3
    assert L.orc_quadrant_pos(np.array([0.0, 0.0]), np.array([0.0, 0.0])) == 3  # ties: x>=, y<=
    assert L.orc_quadrant_pos(np.array([-1.0, 1.0]), np.array([0.0, 0.0])) == 0


def test_morton_closed_forms(oracle):
    L = oracle.lib()
    c = (C.c_int64 * 4)()
    for level in range(0, 9):
        off = (4 ** level - 1) // 3
        for m in list(range(min(4 ** level, 64))) + [4 ** level - 1]:
            g = off + m
            assert L.orc_get_level(g) == level
            L.orc_get_child(g, c)
            assert list(c) == [(4 ** (level + 1) - 1) // 3 + 4 * m + q for q in range(4)]
            for q in range(4):
                assert L.orc_get_parent(c[q]) == g
    # ids of deep levels stay exact in 64 bits
    g = (4 ** 30 - 1) // 3 + 4 ** 30 - 1
    assert L.orc_get_level(g) == 30 and L.orc_get_parent(L.orc_local2global(4 ** 31 - 1, 31)) == g


def test_morton_matches_reference_helpers(oracle):
    if not oracle.ref_available():
        pytest.skip("reference build not available here")
    R = oracle.reflib()
    L = oracle.lib()
    c1 = (C.c_int * 4)()
    c2 = (C.c_int64 * 4)()
    for g in list(range(0, 2000)) + [21845, 87380, 349524]:
        assert R.ref_getLevel(g) == L.orc_get_level(g)
        R.ref_getChild(g, c1)
        L.orc_get_child(g, c2)
        assert list(c1) == list(c2)
        if g > 0:
            assert R.ref_getParent(g) == L.orc_get_parent(g)


# ---- tree: brute force on tiny inputs -------------------------------------
def brute_tree(xy, num_p):
    """Independent construction: every blob walks down from the root; a prefix
    is split iff it holds >= num_p blobs (the root always)."""
    x, y = xy[:, 0], xy[:, 1]
    rx = abs(max(x) - min(x)) + 0.0001
    ry = abs(max(y) - min(y)) + 0.0001
    region0 = rx if rx > ry else ry
    c0 = (0.5 * (max(x) + min(x)), 0.5 * (max(y) + min(y)))
    paths = []
    for i in range(len(x)):
        cx, cy, reg, path = c0[0], c0[1], region0, []
        for _ in range(20):
            q = (1 if x[i] >= cx else 0) + (2 if y[i] <= cy else 0)
            path.append(q)
            cx = cx + 0.25 * reg if q % 2 else cx - 0.25 * reg
            cy = cy + 0.25 * reg if q < 2 else cy - 0.25 * reg
            reg = 0.5 * reg
        paths.append(tuple(path))
    leaf_of = {}
    for i, p in enumerate(paths):
        for level in range(1, 21):
            pre = p[:level]
            cnt = sum(1 for pp in paths if pp[:level] == pre)
            if cnt < num_p:
                m = 0
                for q in pre:
                    m = 4 * m + q
                leaf_of[i] = m + (4 ** level - 1) // 3
                break
    return leaf_of


@pytest.mark.parametrize("n,num_p,seed", [(1, 4, 1), (5, 2, 2), (40, 3, 3), (64, 8, 4), (50, 1 + 1, 5)])
def test_tree_brute_force(oracle, n, num_p, seed):
    xy, cir, cor = workloads().uniform_square(n, seed=seed)
    T = oracle.Tree(xy, cir, cor, 4, num_p, 0.5)
    nd = T.nodes()
    got = {}
    for i in np.nonzero(nd["is_leaf"])[0]:
        for b in nd["members"][nd["lo"][i]:nd["lo"][i] + nd["count"][i]]:
            assert int(b) not in got
            got[int(b)] = int(nd["gid"][i])
    assert got == brute_tree(xy, num_p)


@pytest.mark.parametrize("kind,n,num_p", [("uniform", 5000, 16), ("mixture", 20000, 32),
                                          ("sheet", 8000, 10)])
def test_tree_invariants(oracle, kind, n, num_p):
    wl = workloads()
    xy, cir, cor = {"uniform": wl.uniform_square, "sheet": wl.sheet,
                    "mixture": wl.gaussian_mixture}[kind](n, seed=9)
    T = oracle.Tree(xy, cir, cor, 6, num_p, 0.5)
    nd = T.nodes()
    leaf = nd["is_leaf"] == 1
    assert (nd["count"][leaf] < num_p).all()
    assert (nd["count"][~leaf][1:] >= num_p).all()
    members = np.concatenate([nd["members"][nd["lo"][i]:nd["lo"][i] + nd["count"][i]]
                              for i in np.nonzero(leaf)[0]])
    assert np.array_equal(np.sort(members), np.arange(n))
    pos = {int(g): i for i, g in enumerate(nd["gid"])}
    for i in np.nonzero(~leaf)[0]:
        lv = int(nd["level"][i])
        m = int(nd["gid"][i]) - (4 ** lv - 1) // 3
        kids = [pos[(4 ** (lv + 1) - 1) // 3 + 4 * m + q] for q in range(4)]
        assert sum(int(nd["count"][k]) for k in kids) == int(nd["count"][i])
        h = 0.25 * nd["region"][i]
        for q, k in enumerate(kids):
            assert nd["region"][k] == 0.5 * nd["region"][i]
            assert nd["cx"][k] == (nd["cx"][i] + h if q % 2 else nd["cx"][i] - h)
            assert nd["cy"][k] == (nd["cy"][i] + h if q < 2 else nd["cy"][i] - h)
    # every leaf member lies in its leaf's square, ties going right / down
    for i in np.nonzero(leaf)[0][:500]:
        b = nd["members"][nd["lo"][i]:nd["lo"][i] + nd["count"][i]]
        h = 0.5 * nd["region"][i]
        assert (xy[b, 0] >= nd["cx"][i] - h).all() and (xy[b, 0] < nd["cx"][i] + h).all()
        assert (xy[b, 1] > nd["cy"][i] - h).all() and (xy[b, 1] <= nd["cy"][i] + h).all()


# ---- expansions: M2M(children) == S2M(all blobs of the subtree) ------------
def test_m2m_equals_s2m_of_subtree(oracle):
    xy, cir, cor = workloads().gaussian_mixture(4000, seed=12)
    p = 12
    T = oracle.Tree(xy, cir, cor, p, 20, 0.5)
    nd = T.nodes()
    co = T.coeffs()
    z = xy[:, 0] + 1j * xy[:, 1]
    leaf = nd["is_leaf"] == 1
    # blobs under each node: union of its descendant leaves, by gid prefix
    gl = {int(g): i for i, g in enumerate(nd["gid"])}
    def blobs(i):
        if leaf[i]:
            return nd["members"][nd["lo"][i]:nd["lo"][i] + nd["count"][i]]
        lv = int(nd["level"][i])
        m = int(nd["gid"][i]) - (4 ** lv - 1) // 3
        return np.concatenate([blobs(gl[(4 ** (lv + 1) - 1) // 3 + 4 * m + q]) for q in range(4)])
    checked = 0
    for i in range(1, len(nd["gid"])):
        if nd["count"][i] == 0:
            continue
        b = blobs(i)
        d = z[b] - (nd["cx"][i] + 1j * nd["cy"][i])
        k = np.arange(p)
        want = (cir[b][:, None] * d[:, None] ** k[None, :]).sum(axis=0)
        # rounding in the shifted sums scales with the binomial terms, bounded by
        # (|z - z_c| + the largest centre shift, region / sqrt 2)^k
        reach = np.abs(d) + nd["region"][i] / math.sqrt(2.0)
        scale = (np.abs(cir[b])[:, None] * reach[:, None] ** k[None, :]).sum(axis=0)
        assert (np.abs(co[i] - want) <= 1e-12 * scale + 1e-300).all(), i
        checked += 1
    assert checked > 50
    assert np.all(co[0] == 0)  # nothing is shifted into the root (fmm.c:41)


# ---- velocities: closed forms and truncation bound ------------------------
def test_single_blob_closed_form(oracle):
    """One blob: the root square is 1e-4 wide, so targets further out see the
    far field (M2S, the point-vortex kernel, fmm.c:65) and targets inside the
    opened leaf see the regularised blob kernel (fmm.c:121-142)."""
    g, s = 0.37, 0.05
    xy = np.array([[0.3, 0.6]])
    T = oracle.Tree(xy, np.array([g]), np.array([s]), 8, 4, 0.5)
    for tx, ty in [(0.31, 0.62), (0.35, 0.55), (0.3, 0.7), (0.2, 0.5)]:
        rx, ry = tx - 0.3, ty - 0.6
        f = g / (2 * math.pi * (rx * rx + ry * ry))  # u = -f ry, v = f rx
        u, v = T.eval(np.array([[tx, ty]]))[0]
        assert abs(u + f * ry) <= 1e-13 * abs(f) and abs(v - f * rx) <= 1e-13 * abs(f)
        gid, kind = T.list(tx, ty)
        assert set(kind.tolist()) == {1}
    tx, ty = 0.3 + 2e-5, 0.6 - 1e-5
    rx, ry = tx - 0.3, ty - 0.6
    r2 = rx * rx + ry * ry
    f = g * (1 - math.exp(-r2 / (s * s))) / (2 * math.pi * r2)
    u, v = T.eval(np.array([[tx, ty]]))[0]
    assert abs(u + f * ry) <= 1e-12 * abs(f * ry) and abs(v - f * rx) <= 1e-12 * abs(f * rx)
    assert 0 in T.list(tx, ty)[1].tolist()
    # direct sum agrees (both cut-off rules)
    for style in (0, 1):
        d = oracle.direct(xy, np.array([g]), np.array([s]), np.array([[0.35, 0.55]]), style=style)
        rx, ry = 0.05, -0.05
        f = g * (1 - math.exp(-(rx * rx + ry * ry) / (s * s))) / (2 * math.pi * (rx * rx + ry * ry))
        assert rel_err(d[0], [-f * ry, f * rx])[1] < 1e-14


@pytest.mark.parametrize("theta", [0.5, 0.3])
def test_truncation_bound(oracle, theta):
    """With point-like cores the treecode converges to the direct sum at the
    rate (theta / sqrt 2)^(p - 1) of its M2S series."""
    wl = workloads()
    xy, cir, _ = wl.uniform_square(6000, seed=14)
    cor = np.full(6000, 1e-7)
    t = xy[::37]
    d = oracle.direct(xy, cir, cor, t, style=1)
    rho = theta / math.sqrt(2)
    prev = None
    for p in (3, 6, 9, 12):
        T = oracle.Tree(xy, cir, cor, p, 16, theta)
        e = rel_err(T.eval(t), d)[0]
        assert e <= 50 * rho ** (p - 1), (p, e)
        if prev is not None:
            assert e < prev
        prev = e


def test_direct_styles_agree(oracle):
    xy, cir, cor = workloads().uniform_square(3000, seed=15)
    t = xy[:200] + 1e-3
    a = oracle.direct(xy, cir, cor, t, style=0)
    b = oracle.direct(xy, cir, cor, t, style=1)
    assert rel_err(a, b)[1] < 1e-14


def test_param_reader(oracle, tmp_path):
    f = tmp_path / "p.dat"
    f.write_text("15    nterms\n250   num_p\n0.5d0 theta\n0.01  r_min ! for Tree\n")
    got, nt, npp, th, _ = oracle.read_params(str(f))
    assert (got, nt, npp, th) == (3, 15, 250, 0.5)  # "0.5d0" stops the r_min read
    f.write_text("10    nterms\n32   num_p\n0.4 theta\n0.01  r_min ! for Tree\n")
    assert oracle.read_params(str(f)) == (4, 10, 32, 0.4, 0.01)
    assert oracle.read_params(str(tmp_path / "missing.dat"))[0] == -1
