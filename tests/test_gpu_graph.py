"""HIP-graph replay of the BiCGSTAB iteration: the same kernels in the same
order as the eager launch, so every result must be bit-identical to graph = 0.

Workload: 64x64 (level 4) single solves, forward and adjoint, FP64; a
16-transmitter batched mixed-precision DBIM (3 seeded blobs) for 2 iterations
with the first-order start (FP32-MFMA kernels, and the split-FP16 near field
and P2M), which replays each captured iteration graph several times per
solve, with the driver's own eager work between replays.
"""
import numpy as np
import pytest

from conftest import oracle_for, rand_c

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def isg():
    import isg as m
    return m


@pytest.mark.parametrize("herm", [False, True])
def test_graph_solve_bit_identical(isg, herm):
    print("graph vs eager bicgs, herm", herm, flush=True)
    o = oracle_for(4)
    g = 0.3 * rand_c(np.random.default_rng(1), o.N) * o.k0 ** 2
    b = rand_c(np.random.default_rng(2), o.N)
    out = []
    for graph in (0, 1):
        s = isg.Solver(level=4, toli=1e-10, iti=300, graph=graph)
        x, info = s.bicgs(g, b, herm=herm)
        out.append((x, info["iters"]))
    assert out[0][1] == out[1][1]
    np.testing.assert_array_equal(out[0][0], out[1][0])


@pytest.mark.parametrize("level,split", [(4, 0), (4, 128), (5, 256)])
def test_graph_dbim_bit_identical(isg, level, split):
    """split: the FP32 operator's near field, L2P and P2M on the split-FP16
    kernels (near tuning value 128: pairs of right-hand sides on 64-box
    tiles; 256: one right-hand side on 128-box tiles, which needs level 5),
    captured and replayed like the others."""
    print("graph vs eager batched DBIM, level", level, "split kernels", split, flush=True)
    o = oracle_for(level, numrx=16)
    otrue = isg.raster_to_tree(isg.phantom_blobs(o.side, 3, 1), o.unkmap())
    res = []
    for graph in (0, 1):
        s = isg.Solver(level=level, numrx=16, numtx=16, toli=1e-6, iti=300, regular=float(np.sqrt(1e-3)),
                       max_rhs=16, precision=1, graph=graph)
        if split:
            for cls in (0, 1):
                s.tuning_set(1, cls, "near", 0, split)
                s.tuning_set(1, cls, "p2m", 0, 128)
        d = isg.Dbim(s, otrue, field_start=2)
        its = [d.iterate()["solve_iters"] for _ in range(2)]
        res.append((d.get("o"), its))
    assert res[0][1] == res[1][1]
    np.testing.assert_array_equal(res[0][0], res[1][0])


def test_kernel_timing_with_graph_config(isg):
    """Live kernel timing needs eager launches (events recorded by graph
    nodes carry no timestamps): with timing on, a graph-enabled context
    launches eagerly and every near-field launch is timed."""
    o = oracle_for(4, numrx=16)
    otrue = isg.raster_to_tree(isg.phantom_blobs(o.side, 3, 1), o.unkmap())
    s = isg.Solver(level=4, numrx=16, numtx=16, toli=1e-6, iti=300, regular=float(np.sqrt(1e-3)), max_rhs=16,
                   precision=1, graph=1)
    d = isg.Dbim(s, otrue, field_start=2)
    s.profile(True)
    d.iterate()
    ms, n = s.profile_read("near_fp32")
    s.profile(False)
    assert n > 0 and ms > 0
