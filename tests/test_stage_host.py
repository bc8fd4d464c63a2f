"""Host-side argument checks of ``LocalBatch`` (no device calls).

Every entry point checks the full shape of its arrays against the layout
before any device work is enqueued: a smaller or mis-laid-out tensor would
otherwise be read past its end by the kernels.  The batch below is a host-only
stand-in (no library handle), so reaching the C ABI would fail the test.
"""

from types import SimpleNamespace

import numpy as np
import pytest
import torch

from pymhm_amd.stage import LocalBatch


@pytest.fixture
def batch():
    b = LocalBatch.__new__(LocalBatch)
    b.layout = SimpleNamespace(ne=4, nq=16, dim=2, nnz=30, n=10, m=4, k=1)
    b.macros, b.n, b.m, b.k, b.N, b.R, b.width = 3, 10, 4, 1, 11, 6, 5
    b.max_columns = 2
    b.precision = "fp64"
    b.device = torch.device("cpu")
    b.handle = None
    b._keep = []
    return b


def test_sources_shapes(batch):
    with pytest.raises(ValueError, match="fq"):
        batch.sources(np.zeros((2, 3, 4, 15)))      # wrong quadrature count
    with pytest.raises(ValueError, match="fq"):
        batch.sources(np.zeros((2, 3, 5, 16)))      # wrong element count
    with pytest.raises(ValueError, match="fq"):
        batch.sources(np.zeros((3, 4, 16)))         # missing field axis
    with pytest.raises(ValueError, match="max_columns"):
        batch.sources(np.zeros((3, 3, 4, 16)))      # more fields than max_columns
    with pytest.raises(ValueError, match="max_columns"):
        batch.sources(None, ncols=0)


def test_geometry_and_coupling_shapes(batch):
    with pytest.raises(ValueError, match="grad_lambda"):
        batch.set_geometry(np.zeros((3, 4, 3, 3)), np.zeros((3, 4)))
    with pytest.raises(ValueError, match="measure"):
        batch.set_geometry(np.zeros((3, 4, 3, 2)), np.zeros((3, 5)))
    with pytest.raises(ValueError, match="couplings"):
        batch.set_coupling(couplings=np.zeros((3, 10, 4)))   # transposed
    with pytest.raises(ValueError, match="constraints"):
        batch.set_coupling(constraints=np.zeros((3, 11)))
    with pytest.raises(ValueError, match="colmap"):
        batch.set_coupling(colmap=np.full((3, 4), 4))
    with pytest.raises(ValueError, match="kq"):
        batch.assemble(np.ones((3, 4, 15)))
    with pytest.raises(ValueError, match="f_const"):
        batch.assemble(np.ones((3, 4, 16)), f_const=np.nan)


def test_solve_and_reconstruct_shapes(batch):
    with pytest.raises(ValueError, match="rhs"):
        batch.solve(np.zeros((3, 2, 10)))
    with pytest.raises(ValueError, match="max_columns"):
        batch.solve(np.zeros((3, 3, 11)))
    with pytest.raises(ValueError, match="trace"):
        batch.reconstruct(np.zeros((3, 1, 5)), np.zeros((3, 1)))
    with pytest.raises(ValueError, match="coarse"):
        batch.reconstruct(np.zeros((3, 2, 4)), np.zeros((3, 3)))
    with pytest.raises(ValueError, match="coarse"):
        batch.reconstruct(np.zeros((3, 2, 4)), None)
