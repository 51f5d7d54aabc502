"""Node model host side: the family's index table and target map, the assembly map,
and the merit-order inputs (CPU only).

The gather semantics of ``pdlp_node_gather`` are restated here in NumPy
(:func:`gather_ref`) so the GPU tests compare the kernel with the same function.
"""

import numpy as np
import pytest

from sovereign_solver_amd.learned.nodes import NodeSpec, assembly_map
from sovereign_solver_amd.workloads import DispatchFamily


def gather_ref(src, shared, table, scale=None, mean=None, inv_std=None, src2=None, rows=None, ld_out=None):
    """Rows of ``pdlp_node_gather``: row r holds node q = rows[r] (q = b * nodes + j)."""
    src = np.asarray(src, np.float64)
    B, Ls = src.shape
    nodes, width = table.shape
    ld_out = width if ld_out is None else ld_out
    sh = np.asarray(shared, np.float64)
    s2 = np.zeros((B, 0)) if src2 is None else np.asarray(src2, np.float64)
    rows = np.arange(B * nodes) if rows is None else np.asarray(rows)
    b, j = rows // nodes, rows % nodes
    t = table[j]                                                     # [R, width]
    full = np.concatenate([src, np.broadcast_to(sh, (B, sh.size)), s2], axis=1)
    g = np.where(t >= 0, full[b[:, None], np.maximum(t, 0)], 0.0)
    if scale is not None:
        g = g * scale[j]
    if mean is not None:
        g = g - mean[None, :width]
    if inv_std is not None:
        g = g * inv_std[None, :width]
    out = np.zeros((len(rows), ld_out))
    out[:, :width] = g
    return out


def assemble_ref(net, nodes, indptr, node, chan, weight, bias, length):
    B = net.shape[0] // nodes
    out = np.zeros((B, length))
    for i in range(length):
        e = slice(indptr[i], indptr[i + 1])
        out[:, i] = bias[i] + (weight[e][None, :] * net.reshape(B, nodes, -1)[:, node[e], chan[e]]).sum(axis=1)
    return out


def _toy_spec():
    # 3 nodes, 2 channels; solution z of length 5: z0..z2 one node each, z3 predicted by
    # nodes 0 and 2 (averaged), z4 by node 1 (channel 1)
    table = np.array([[0, 3], [1, -1], [2, 3]], np.int32)
    tidx = np.array([[0, 3], [1, 4], [2, 3]], np.int32)
    tscale = np.array([[2.0, 0.5], [1.0, 4.0], [0.25, 0.5]], np.float32)
    return NodeSpec(nodes=3, table=table, shared=np.array([7.0], np.float32), src_len=3, target_index=tidx,
                    target_scale=tscale)


def test_assembly_inverts_the_target_map():
    sp = _toy_spec()
    z = np.array([[1.0, -2.0, 3.0, 0.5, 8.0], [0.0, 4.0, -1.0, 2.0, -3.0]])
    # targets per node, normalised by (mean, std) per channel
    T = gather_ref(z, np.zeros(0), sp.target_index, scale=sp.target_scale)
    mean, std = np.array([0.3, -0.2]), np.array([2.0, 0.5])
    Tn = (T - mean) / std
    ip, nd, ch, w, b = assembly_map(sp, 5, mean, std)
    assert list(np.diff(ip)) == [1, 1, 1, 2, 1]
    back = assemble_ref(Tn, 3, ip, nd, ch, w, b, 5)
    np.testing.assert_allclose(back, z, rtol=1e-6, atol=1e-6)


def test_assembly_rejects_bad_maps():
    sp = _toy_spec()
    with pytest.raises(ValueError):
        assembly_map(sp, 4, np.zeros(2), np.ones(2))                  # index 4 beyond the start
    bad = NodeSpec(sp.nodes, sp.table, sp.shared, sp.src_len, sp.target_index, sp.target_scale.copy())
    bad.target_scale[1, 1] = 0.0
    with pytest.raises(ValueError):
        assembly_map(bad, 5, np.zeros(2), np.ones(2))


def test_gather_ref_reads_src_shared_and_src2():
    sp = _toy_spec()
    src = np.arange(6.0).reshape(2, 3)
    out = gather_ref(src, sp.shared, sp.table, ld_out=4)
    # node 0 of instance 1: src[1, 0] = 3 and shared[0] = 7; node 1: src[1, 1], then zero
    np.testing.assert_array_equal(out[3], [3.0, 7.0, 0.0, 0.0])
    np.testing.assert_array_equal(out[4], [4.0, 0.0, 0.0, 0.0])
    t2 = np.array([[4, 0]], np.int32)                                 # index 4 = src2[b][0]
    out2 = gather_ref(src, sp.shared, t2, src2=np.array([[10.0], [20.0]]))
    np.testing.assert_array_equal(out2[:, 0], [10.0, 20.0])


@pytest.fixture(scope="module")
def fam():
    return DispatchFamily()


def test_dispatch_node_spec_covers_the_solution(fam):
    sp = fam.node_spec()
    n, m, G, T = fam.n, fam.m, fam.G, fam.T
    assert sp.nodes == n and sp.table.shape == (n, sp.width)
    assert sp.table.min() >= 0 and sp.table.max() < sp.src_len + sp.shared.size
    cnt = np.bincount(sp.target_index[sp.target_index >= 0], minlength=n + m)
    assert cnt.size == n + m
    assert np.all(cnt[:n] == 1)                                       # every p_gt by its own node
    assert np.all(cnt[n:n + T] == G)                                  # a demand dual by the period's G nodes
    assert np.all(cnt[n + T:] == 1)                                   # every ramp dual once
    # ramp rows of (t, g): the up row r0 + 2((t-1)G + g) has +1 at p_gt and -1 at p_g,t-1
    A = fam.A.tocsr()
    for j in (G + 5, 7 * G + 199, (T - 1) * G):
        up = sp.target_index[j, 1] - n
        row = A.getrow(up)
        assert dict(zip(row.indices, row.data)) == {j: 1.0, j - G: -1.0}
        dn = sp.target_index[j, 2] - n
        row = A.getrow(dn)
        assert dict(zip(row.indices, row.data)) == {j: -1.0, j - G: 1.0}
    assert np.all(sp.target_index[:G, 1:3] == -1)                     # no ramp rows at t = 0


def test_dispatch_merit_order_inputs(fam):
    # a small family checked against a brute-force merit order
    f = DispatchFamily(G=7, T=4)
    d, c = f.params(3)
    v = f.node_inputs_params(d[None], c[None])[0]
    o, L = f._node_offsets()
    assert v.shape == (L,)
    S = f.pmax.sum()
    cheaper = np.array([f.pmax[c < c[g]].sum() for g in range(f.G)])
    np.testing.assert_allclose(v[o["CC"]:o["CC"] + f.G], 2 * cheaper / S, rtol=1e-6)
    xm = v[o["X"]:].reshape(f.T, f.G) * f.pmax
    for t in range(f.T):
        # the merit-order dispatch meets the demand (capacity exceeds it) and fills cheaper units first
        assert abs(xm[t].sum() - min(d[t], S)) <= 1e-3 * S
        marg = np.argmin(np.where(xm[t] < f.pmax - 1e-6, c, np.inf))
        assert np.all(xm[t][c < c[marg]] >= f.pmax[c < c[marg]] - 1e-3)
        assert np.all(xm[t][c > c[marg]] <= 1e-6)
        assert v[o["P"] + t] == pytest.approx(c[marg] / 60.0, rel=1e-6)


def test_dispatch_neighbours(fam):
    nb = fam.node_neighbours()
    G, T = fam.G, fam.T
    assert nb.shape == (fam.n, 5)
    j = 5 * G + 17
    assert list(nb[j]) == [3 * G + 17, 4 * G + 17, 5 * G + 17, 6 * G + 17, 7 * G + 17]
    assert list(nb[17]) == [17, 17, 17, G + 17, 2 * G + 17]            # clamped at the horizon
    assert list(nb[(T - 1) * G]) == [(T - 3) * G, (T - 2) * G, (T - 1) * G, (T - 1) * G, (T - 1) * G]


def test_node_inputs_from_problems_match_ids(fam):
    ids = [3, 11]
    a = fam.node_inputs(ids)
    b = fam.node_inputs_from([fam.problem(k) for k in ids])
    np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-7)
