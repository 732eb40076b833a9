"""Box partition on the CPU: the library's sort-free partition_boxes (host code, C ABI) against the
oracle's sort-based restatement of src/Box.jl:159-374, on the geometries of the reference's own
tests (test_geometry_windows.jl:29-52, test_cost.jl:66-73) and a few more: the same boxes, in the
same order, with the same points."""
import numpy as np
import pytest


def _lib_groups(th, ph, L, B=1, tol=1e-8):
    import pfsht
    return pfsht.partition_boxes(th, ph, L, B, tol)


def _geometry(name, rng):
    if name == "band":            # an equatorial band stays whole
        half = np.deg2rad(8)
        th = np.pi / 2 + half * (2 * rng.uniform(size=3000) - 1)
        ph = 2 * np.pi * rng.uniform(size=3000)
        return th, ph, 256, 1
    if name == "antipodal":       # two antipodal caps separate
        th = np.concatenate([0.2 * np.sqrt(rng.uniform(size=1500)), np.pi - 0.2 * np.sqrt(rng.uniform(size=1500))])
        return th, 2 * np.pi * rng.uniform(size=3000), 256, 2
    if name == "patch":           # a compact patch is not split
        return 0.2 * np.sqrt(rng.uniform(size=2000)), 2 * np.pi * rng.uniform(size=2000), 256, 1
    if name == "cap_band":
        n = 4000
        th = np.where(np.arange(n) % 2 == 0, np.arccos(1 - rng.uniform(size=n) * (1 - np.cos(0.2))),
                      rng.uniform(1.0, 1.3, n))
        return th, rng.uniform(0, 2 * np.pi, n), 256, 1
    if name == "three_patches":   # three patches apart in colatitude and longitude
        n = 1500
        th = np.concatenate([rng.uniform(0.5, 0.7, n), rng.uniform(1.5, 1.7, n), rng.uniform(2.5, 2.7, n)])
        ph = np.concatenate([rng.uniform(0.5, 0.8, n), rng.uniform(3.0, 3.3, n), rng.uniform(5.0, 5.4, n)])
        return th, ph, 512, 1
    raise ValueError(name)


@pytest.mark.parametrize("name,expect", [("band", 1), ("antipodal", 2), ("patch", 1), ("cap_band", 2),
                                         ("three_patches", 3)])
def test_partition_matches_oracle(O, name, expect):
    rng = np.random.default_rng(22)
    th, ph, L, B = _geometry(name, rng)
    got = _lib_groups(th, ph, L, B)
    ref, costs = O.partition_boxes(th, ph, L, B, 1e-8)
    if expect is not None:
        assert len(got) == expect
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)
    assert np.array_equal(np.sort(np.concatenate(got)), np.arange(len(th)))  # a partition of the points
    assert all(np.isfinite(costs))


def test_partition_single_point_and_errors():
    import pfsht
    g = pfsht.partition_boxes([0.3], [1.0], 64)
    assert len(g) == 1 and np.array_equal(g[0], [0])
    with pytest.raises(pfsht.ArgumentError):
        pfsht.partition_boxes([0.3], [1.0], 64, maxboxes=0)


def test_select_method_oracle(O):
    # test_cost.jl:57-73: a tiny problem takes Direct, two antipodal caps take the multi-box leaf
    rng = np.random.default_rng(73)
    th, ph = 0.3 * rng.uniform(size=12), 2 * np.pi * rng.uniform(size=12)
    assert O.select_method(th, ph, 4, 1e-8) == "direct"
    rng = np.random.default_rng(74)
    n1 = 2000
    th = np.concatenate([0.2 * np.sqrt(rng.uniform(size=n1)), np.pi - 0.2 * np.sqrt(rng.uniform(size=n1))])
    ph = 2 * np.pi * rng.uniform(size=2 * n1)
    assert O.select_method(th, ph, 512, 1e-8) == "multibox"
    # one compact cap stays on the stretched leaf
    assert O.select_method(th[:n1], ph[:n1], 512, 1e-8) == "stretch"


@pytest.mark.parametrize("name", ["antipodal", "three_patches", "cap_band"])
def test_partition_threaded_scans_match_oracle(O, name):
    # groups of >= 65536 points take the threaded bucket scans; the result must not change
    rng = np.random.default_rng(23)
    n = 240_000
    if name == "antipodal":
        th = np.concatenate([0.2 * np.sqrt(rng.uniform(size=n // 2)), np.pi - 0.2 * np.sqrt(rng.uniform(size=n // 2))])
        ph = 2 * np.pi * rng.uniform(size=n)
        L = 1024
    elif name == "three_patches":
        k = n // 3
        th = np.concatenate([rng.uniform(0.5, 0.7, k), rng.uniform(1.5, 1.7, k), rng.uniform(2.5, 2.7, k)])
        ph = np.concatenate([rng.uniform(0.5, 0.8, k), rng.uniform(3.0, 3.3, k), rng.uniform(5.0, 5.4, k)])
        L = 1024
    else:
        th = np.where(np.arange(n) % 2 == 0, np.arccos(1 - rng.uniform(size=n) * (1 - np.cos(0.2))),
                      rng.uniform(1.0, 1.3, n))
        ph = rng.uniform(-np.pi, 3 * np.pi, n)  # longitudes outside [0, 2 pi) too
        L = 1024
    got = _lib_groups(th, ph, L)
    ref, _ = O.partition_boxes(th, ph, L, 1, 1e-8)
    assert len(got) == len(ref) >= 2
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)
