"""Host-side logic: mesh-only maps against the oracle, and the C ABI exports.

Integer maps (local meshes, DOF numbering, trace DOFs) must match the oracle
bit for bit; template couplings with per-macro signed column maps must equal
the oracle's per-macro trace coupling.  No GPU is used here.
"""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import workloads as ow
from pymhm_amd import _lib, hybrid, layout
from pymhm_amd import workloads as pw

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "mhm_hip.h").read_text()
    declared = set(re.findall(r"^int (mhm_\w+)\(", header, flags=re.M))
    assert declared == set(_lib.EXPORTS)
    lib = ctypes.CDLL(str(_lib.LIBRARY))
    for name in declared:
        assert hasattr(lib, name), name
    major, minor = ctypes.c_int32(), ctypes.c_int32()
    assert lib.mhm_version(ctypes.byref(major), ctypes.byref(minor)) == 0
    assert major.value == 1


def test_no_device_means_clean_error():
    count = ctypes.c_int32(-1)
    assert _lib.load().mhm_device_count(ctypes.byref(count)) == 0
    assert count.value >= 0


@pytest.mark.parametrize("make,macros", [(lambda: pw.config1(nx=4), None),
                                         (lambda: pw.config2(nx=3, refinement=4), [0, 4, 8])])
def test_template_matches_oracle_per_macro(make, macros):
    case = make()
    prob = case.problem
    problems, skel = ow.build_problems(case, macros)
    assert skel.size == prob.trace_size
    for c, p, fine in problems:
        assert np.array_equal(fine.cells, prob.mesh.cells)
        assert np.allclose(fine.points - prob.offsets[c], prob.mesh.points, atol=1e-15, rtol=0)
        assert np.array_equal(p.trace_dofs, prob.trace_dofs[c])
        coupling = prob.coupling[:, prob.colmap[c]] * prob.colscale[c]
        assert np.abs(coupling - p.coupling).max() <= 4e-15 * np.abs(p.coupling).max()
        assert np.abs(prob.constraint - p.constraints[:, 0]).max() <= 4e-15 * prob.constraint.max()


def test_band_layout_is_a_permutation_with_small_bandwidth():
    prob = pw.config2(nx=2).problem
    lay = prob.layout
    assert sorted(lay.perm.tolist()) == list(range(lay.n))
    assert lay.bandwidth == 66            # two lattice rows of a 33-node P2 row
    rows = np.repeat(np.arange(lay.n), np.diff(lay.rowptr))
    assert np.all(np.abs(rows - lay.cols) <= lay.bandwidth)
    # contribution lists cover every (element, i, j) exactly once
    code = lay._host["slot_code"]
    assert np.array_equal(np.sort(code), np.arange(lay.ne * lay.nb * lay.nb))


def test_scatter_plan_matches_reference_coo_stream():
    prob = pw.config1(nx=3).problem
    idx = prob.global_indices()
    plan = hybrid.HybridScatter(idx, prob.trace_size, prob.kernel_offsets(), device=0)
    w = idx.shape[1]
    rows = np.repeat(idx, w, axis=1).ravel()
    cols = np.tile(idx, (1, w)).ravel()
    from scipy import sparse
    ref = sparse.coo_matrix((np.arange(1, len(rows) + 1, dtype=float), (rows, cols)),
                            shape=(plan.size, plan.size)).tocsc()
    assert np.array_equal(ref.indptr, plan.csc_indptr)
    assert np.array_equal(ref.indices, plan.csc_indices)
    host = plan._host
    vals = np.array([np.sum(np.arange(1, len(rows) + 1)[host["vcode"][host["vptr"][g]:host["vptr"][g + 1]]])
                     for g in range(len(plan.csc_indices))], dtype=float)
    assert np.array_equal(vals, ref.data)


def test_scatter_plan_rejects_bad_partitions():
    prob = pw.config1(nx=2).problem
    idx = prob.global_indices()
    bad = idx.copy()
    bad[0, -1] = bad[1, -1]
    with pytest.raises(ValueError):
        hybrid.HybridScatter(bad, prob.trace_size, prob.kernel_offsets(), device=0)


def test_level_order_handles_rcm_fallback():
    indptr = np.array([0, 2, 4, 6])
    indices = np.array([0, 1, 0, 1, 2, 2])
    with pytest.raises(ValueError):
        layout.level_order(indptr, indices, [0])   # node 2 unreachable


def test_ragged_scatter_plan_in_bucket_order_matches_reference():
    """Uneven macros: per-cell widths differ and the blocks sit bucket by
    bucket in the device buffer; the gathered values equal the reference's
    assembly of the same contributions (core/contributions.py:77-195)."""
    from oracle.local import assemble_hybrid_contributions

    prob = pw.config4(npts=20, h=1 / 64, dof_range=(30, 900)).problem
    idx = prob.global_indices()
    rng = np.random.default_rng(4)
    blocks = [rng.standard_normal((len(i), len(i))) for i in idx]
    loads = [rng.standard_normal(len(i)) for i in idx]
    # device buffers in bucket order
    order = np.concatenate([bk.cells for bk in prob.buckets])
    wid = np.array([len(i) for i in idx])
    bpos = np.zeros(prob.M, dtype=np.int64)
    lpos = np.zeros(prob.M, dtype=np.int64)
    bpos[order] = np.concatenate(([0], np.cumsum(wid[order] ** 2)[:-1]))
    lpos[order] = np.concatenate(([0], np.cumsum(wid[order])[:-1]))
    flat_b = np.concatenate([blocks[c].ravel() for c in order])
    flat_l = np.concatenate([loads[c] for c in order])
    plan = hybrid.HybridScatter(idx, prob.trace_size, prob.kernel_offsets(), device=0, positions=bpos,
                                load_positions=lpos)
    host = plan._host
    vals = np.array([np.sum(flat_b[host["vcode"][host["vptr"][g]:host["vptr"][g + 1]]])
                     for g in range(len(plan.csc_indices))])
    rhs = np.array([np.sum(flat_l[host["lcode"][host["lptr"][g]:host["lptr"][g + 1]]])
                    for g in range(plan.size)])
    mat, orhs, _ = assemble_hybrid_contributions([(idx[c], blocks[c], loads[c]) for c in range(prob.M)],
                                                 prob.trace_size, prob.kernel_offsets())
    from scipy import sparse
    got = sparse.csc_matrix((vals, plan.csc_indices, plan.csc_indptr), shape=(plan.size, plan.size))
    assert np.array_equal(got.toarray(), mat.toarray())
    assert np.array_equal(rhs, orhs)
