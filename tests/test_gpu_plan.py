"""Plan construction on the device (src/Stretch.jl:129-248): the coefficient tables, the points'
map onto the fine grid, the stable bin sort and the reached rows are built by kernels. The checks:
the plan is a pure function of its points (two builds give the same bits), the synthesis at a point
does not depend on where the point sits in the input, bad coordinates are refused, and the build
is fast enough for uncached serving geometries."""
import os
import time

import numpy as np
import pytest

from conftest import cap_points, relmax

pytestmark = pytest.mark.gpu
NT = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def torch():
    import torch as T
    assert T.cuda.is_available(), "a GPU is required for the gpu-marked tests"
    return T


@pytest.fixture(scope="module")
def P():
    import pfsht
    pfsht.lib()
    return pfsht


def test_rebuild_and_input_order(torch, P, O, rng):
    L, tol, M = 384, 1e-8, 60000
    th, ph = cap_points(rng, 20, M)
    a = O.rand_alm_fast(L, rng)
    f = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    p1 = P.Plan(th, ph, lmax=L, tol=tol, method="stretch")
    p2 = P.Plan(th, ph, lmax=L, tol=tol, method="stretch")
    assert np.array_equal(p1 @ a, p2 @ a) and np.array_equal(p1.H @ f, p2.H @ f)
    # a shuffled input: the same value at every point (interpolation is per point), the adjoint
    # sums the same terms in another order
    sh = rng.permutation(M)
    p3 = P.Plan(th[sh], ph[sh], lmax=L, tol=tol, method="stretch")
    assert np.array_equal((p3 @ a), (p1 @ a)[sh])
    m = O.valid_mask(L)
    assert relmax((p3.H @ f[sh])[m], (p1.H @ f)[m]) < 1e-13
    # and against the oracle's host build of the same plan
    po = O.Plan(th, ph, L, tol=tol, method="stretch", nufft="es", nthreads=NT)
    assert relmax(p1 @ a, po.synth(a)) < 1e-11


@pytest.mark.parametrize("bad", ["phi_nan", "theta_nan", "phi_inf"])
def test_bad_coordinates_refused(torch, P, rng, bad):
    th, ph = cap_points(rng, 20, 5000)
    if bad == "phi_nan":
        ph[17] = np.nan
    elif bad == "theta_nan":
        th[17] = np.nan
    else:
        ph[17] = np.inf
    with pytest.raises(P.ArgumentError):
        P.Plan(th, ph, lmax=256, tol=1e-8, method="stretch")
    # the library is usable afterwards
    th2, ph2 = cap_points(rng, 20, 5000)
    assert P.Plan(th2, ph2, lmax=256, tol=1e-8, method="stretch").info["method"] == 2


def test_build_time_lmax2048_M2e6(torch, P, rng):
    # uncached geometry of a seen size: the reference quotes 0.24-0.48 s (README.md:68-70)
    th, ph = cap_points(rng, 20, 2_000_000)
    P.Plan(th, ph, lmax=2048, tol=1e-8, method="stretch").close()  # rocFFT kernels for these sizes
    th, ph = cap_points(rng, 20, 2_000_000)
    t0 = time.perf_counter()
    p = P.Plan(th, ph, lmax=2048, tol=1e-8, method="stretch")
    dt = time.perf_counter() - t0
    print(f"\nplan build, new points: {dt * 1e3:.1f} ms", flush=True)
    assert dt < 0.24
    p.close()
