"""Orderings on the GPU: edge colouring and RCM, bit-exact against the host
implementation and the oracle's restatement of the same rules, up to the
128^3 benchmark box (2,146,689 nodes, 14,827,904 edges)."""
import numpy as np
import pytest

import oracle as O
from aerof_dev_amd import mesh as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(20, 20, 20), (24, 8, 8), (40, 40, 40), (128, 128, 128)])
def test_gpu_colouring_bit_exact(gpu, shape):
    m = M.kuhn_box(*shape, jitter=0.15, seed=4)
    ptr = M.edges(m)
    c_gpu, n_gpu = M.color_edges_gpu(m.num_nodes, ptr)
    c_host, n_host = M.color_edges(m.num_nodes, ptr)
    assert n_gpu == n_host
    assert np.array_equal(c_gpu, c_host)
    if shape[0] <= 24:
        c_or, n_or = O.color_edges(m.num_nodes, ptr)
        assert np.array_equal(c_gpu, c_or) and n_or == n_gpu
    # proper colouring: no two edges of one colour share a node
    for col in range(n_gpu):
        e = ptr[c_gpu == col]
        nodes = e.ravel()
        assert np.unique(nodes).size == nodes.size


@pytest.mark.parametrize("shape", [(20, 20, 20), (24, 8, 8), (40, 40, 40), (128, 128, 128)])
def test_gpu_rcm_bit_exact(gpu, shape):
    m = M.kuhn_box(*shape, jitter=0.15, seed=4)
    r_gpu = M.rcm_gpu(m)
    r_host = M.rcm(m)
    assert np.array_equal(r_gpu, r_host)
    assert np.array_equal(np.sort(r_gpu), np.arange(m.num_nodes))
    if shape[0] <= 24:
        r_or = O.rcm(m.num_nodes, M.edges(m))
        assert np.array_equal(r_gpu, r_or)
