"""dg_masks_noalloc: the integer wet/dry masks through the library's static staging
area (no device allocation) equal dg_masks's, over two calls in a row on one GPU,
the second on a smaller mesh, and a mesh whose masks exceed the area is refused."""
import numpy as np
import pytest

from dgcom_amd import cases, swe2d


def flat_case(n, order, depth=10.0):
    mesh = cases.structured_mesh(0.0, 1000.0 * n, 0.0, 1000.0 * n, n, n)
    nn = (order + 1) ** 2
    bed = np.full(mesh.n_elements * nn, -depth)
    h, hu, hv = cases.still_water(bed)
    return cases.Case("flat", order, mesh, bed, np.full(mesh.n_elements, 0.025), 0.0, cases.m2_tide(0.0), h, hu, hv)


@pytest.mark.gpu
def test_noalloc_masks_equal_masks_on_two_calls():
    big = cases.drying_beach(nx=48, ny=12, order=3, jitter=0.2)
    small = cases.drying_beach(nx=10, ny=4, order=2, jitter=0.1, seed=3)
    runs = []
    for case, steps in ((big, 30), (small, 20)):
        phys = swe2d.SWEPhysics2D.from_case(case)
        state = phys.state(case.h, case.hu, case.hv)
        phys.run(state, 0.0, steps, case.cfl)
        runs.append((phys, state))
    outs = []
    for phys, state in runs:  # the second call's input is smaller than the first's
        nd, es = phys.masks_noalloc(state)
        nd_ref, es_ref = phys.masks(state)
        assert nd.sum() > 0 and es.sum() > 0
        assert np.array_equal(nd, nd_ref) and np.array_equal(es, es_ref)
        outs.append((nd, es))
    # the small mesh again after the large one, then the large one again: same bytes
    for (phys, state), (nd, es) in zip(reversed(runs), reversed(outs)):
        nd2, es2 = phys.masks_noalloc(state)
        assert np.array_equal(nd2, nd) and np.array_equal(es2, es)


@pytest.mark.gpu
def test_noalloc_masks_refuse_beyond_the_area():
    # K (NN + 1) bytes: 1410^2 elements at N = 3 take 33,797,700 > 32 MiB
    case = flat_case(1410, 3)
    phys = swe2d.SWEPhysics2D.from_case(case)
    state = phys.state(case.h, case.hu, case.hv)
    with pytest.raises(RuntimeError, match="static staging area"):
        phys.masks_noalloc(state)
    nd, es = phys.masks(state)  # the allocating call still serves it
    assert nd.sum() == 0 and es.sum() == 0
