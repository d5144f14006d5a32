"""The oracle against the reference's own outputs (golden fixtures written by
tests/golden/make_golden.py from the reference build) and, where the
reference sources are present, against a live run of the reference build on
fresh seeded inputs.  The bar is bit-exactness: same tree, same leaf lists,
same coefficients, same velocities, same interaction lists."""
import numpy as np
import pytest

from conftest import golden_cases, load_golden, workloads


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def _check_tree(T, ref):
    nd = T.nodes()
    for k in ("gid", "level", "count", "is_leaf"):
        assert np.array_equal(nd[k], ref[k]), k
    for k in ("cx", "cy", "region"):
        assert np.array_equal(_bits(nd[k]), _bits(ref[k])), k
    rl = ref["leaf_lo"]
    for i in np.nonzero(ref["is_leaf"])[0]:
        a = nd["members"][nd["lo"][i]:nd["lo"][i] + nd["count"][i]]
        b = ref["leaf_members"][rl[i]:rl[i] + ref["count"][i]]
        assert np.array_equal(a, b), i
    assert np.array_equal(_bits(T.coeffs().view(np.float64)), _bits(ref["coeffs"].view(np.float64)))


@pytest.mark.parametrize("case", golden_cases())
def test_oracle_matches_reference_golden(oracle, case):
    g = load_golden(case)
    nterms, num_p, theta = int(g["params"][0]), int(g["params"][1]), float(g["params"][2])
    T = oracle.Tree(g["xy"], g["cir"], g["cor"], nterms, num_p, theta)
    _check_tree(T, g)
    txy = g["txy"] if "txy" in g else g["xy"]
    sel = np.arange(txy.shape[0]) if txy.shape[0] <= 12000 else np.arange(0, txy.shape[0], 7)
    uv = T.eval(txy[sel])
    assert np.array_equal(_bits(uv), _bits(g["uv"][sel]))
    off = g["list_off"]
    for j, (x, y) in enumerate(g["list_xy"]):
        gid, kind = T.list(x, y)
        assert np.array_equal(gid, g["list_gid"][off[j]:off[j + 1]])
        assert np.array_equal(kind, g["list_kind"][off[j]:off[j + 1]])


REF_CASES = [("uniform", 3000, 41, 8, 16, 0.5), ("sheet", 5000, 42, 14, 40, 0.45),
             ("mixture", 6000, 43, 12, 24, 0.6), ("uniform", 800, 44, 3, 2, 0.9)]


@pytest.mark.parametrize("kind,n,seed,nterms,num_p,theta", REF_CASES)
def test_oracle_matches_live_reference(oracle, kind, n, seed, nterms, num_p, theta):
    if not oracle.ref_available():
        pytest.skip("reference build (oracle/_ref) not available here")
    wl = workloads()
    xy, cir, cor = {"uniform": wl.uniform_square, "sheet": wl.sheet,
                    "mixture": wl.gaussian_mixture}[kind](n, seed=seed)
    R = oracle.RefRun(xy, cir, cor, nterms, num_p, theta)
    ref = R.nodes()
    T = oracle.Tree(xy, cir, cor, nterms, num_p, theta)
    _check_tree(T, ref)
    idx = np.arange(0, n, max(1, n // 300))
    assert np.array_equal(_bits(T.eval(xy[idx])), _bits(R.eval(xy[idx])))
    for i in idx[:10]:
        g1, k1 = T.list(*xy[i])
        g2, k2, uv = R.list(*xy[i])
        assert np.array_equal(g1, g2) and np.array_equal(k1, k2)
        # the restated walk reproduces biot_tree bit for bit, so it is biot_tree's list
        assert np.array_equal(_bits(np.array(uv)), _bits(R.eval(xy[i:i + 1])[0]))
