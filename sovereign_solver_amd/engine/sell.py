"""SELL-64 row slices of a CSR matrix for the dual step.

A slice holds 64 consecutive rows, one per lane of a wavefront, with its
entries stored column-major (entry k of lane l at ``ptr[s] + 64 k + l``), so
a wavefront reads entry k of all 64 rows in one coalesced access.  A slice
pays off when, for most k, the 64 rows' k-th columns are neighbours — their
gathers of xbar then share cache lines (the demand rows of a transportation
LP: row j's k-th entry is column k·S + j).  Rows whose slice does not pay off
stay with the CSR kernel.  Format: Monakov et al. 2010 (sliced ELLPACK),
Kreutzer et al. 2014 (SELL-C-σ), here with C = 64 and σ = 1 (no sorting:
consecutive rows keep the neighbour property).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

SLICE = 64


@dataclass
class SellSlices:
    ptr: np.ndarray        # int32, slices + 1: first entry of each slice
    rows: np.ndarray       # int32, 64 * slices: row of each lane, -1 = padding
    col: np.ndarray        # int32, entries (padding 0)
    val: np.ndarray        # float64, entries (padding 0)
    csr_rows: np.ndarray   # int32: rows outside the slices
    padded: bool = False   # some slice pads a row (rows of unequal length)
    stride: np.ndarray = None   # int32 per slice: the column step of every row (col_k = col_0 + k stride), or 0

    @property
    def slices(self) -> int:
        return len(self.ptr) - 1


def _windows(rl: np.ndarray):
    """Candidate slices: 64-row windows inside each run of rows of equal length
    (at least 32 rows; a shorter last window is padded), then 64-row windows
    over the remaining rows in order."""
    m = rl.size
    change = np.flatnonzero(np.diff(rl)) + 1
    starts = np.concatenate([[0], change])
    ends = np.concatenate([change, [m]])
    used = np.zeros(m, bool)
    out = []
    for s0, e0 in zip(starts, ends):
        if e0 - s0 < SLICE:
            continue
        for w in range(s0, e0, SLICE):
            r = np.arange(w, min(w + SLICE, e0))
            if r.size >= SLICE // 2:
                out.append(r)
                used[r] = True
    rest = np.flatnonzero(~used)
    for w in range(0, rest.size - SLICE // 2 + 1, SLICE):
        out.append(rest[w: w + SLICE])
    return out


def build(K: sp.csr_matrix, line_elems: int = 32, max_lines: float = 4.0, max_pad: float = 1.25) -> SellSlices:
    """Slices of up to 64 rows whose k-th columns span at most ``max_lines``
    cache lines of ``line_elems`` elements on average, with at most ``max_pad``
    padding (width x 64 over the slice's entries); the other rows stay CSR."""
    m = K.shape[0]
    indptr, indices, data = K.indptr, K.indices, K.data
    rl = np.diff(indptr)
    ptr = [0]
    rows_out, col_out, val_out = [], [], []
    in_slice = np.zeros(m, bool)
    padded = False
    strides = []
    for r in _windows(rl):
        lens = rl[r]
        if lens.min() == 0:
            continue
        W = int(lens.max())
        if W * SLICE > max_pad * lens.sum():
            continue
        if lens.min() == W and r[-1] - r[0] == r.size - 1:
            # consecutive rows of equal length: their entries are one block, [rows, W]
            e0 = indptr[r[0]]
            cols = indices[e0: e0 + r.size * W].reshape(r.size, W).T       # [W, rows]: entry-major
            vals = data[e0: e0 + r.size * W].reshape(r.size, W).T
            lines = np.sort(cols // line_elems, axis=1)
        else:
            k = np.arange(W)
            valid = k[:, None] < lens[None, :]
            pos = np.minimum(indptr[r][None, :] + k[:, None], len(indices) - 1)
            cols = np.where(valid, indices[pos], 0)
            vals = np.where(valid, data[pos], 0.0)
            lines = np.sort(np.where(valid, cols // line_elems, -1), axis=1)
        distinct = 1 + (np.diff(lines, axis=1) != 0).sum(axis=1)           # per k (padding counts once)
        if distinct.mean() > max_lines:
            continue
        in_slice[r] = True
        padded |= bool(lens.min() != W)
        st = 0                                                            # index run: rows of full width only
        if lens.min() == W:
            st = 1 if W == 1 else int(cols[1, 0] - cols[0, 0])
            if W > 1 and not np.all(np.diff(cols, axis=0) == st):
                st = 0
        strides.append(st)
        rows_out.append(np.concatenate([r, np.full(SLICE - r.size, -1)]).astype(np.int32))  # -1: no row
        c = np.zeros((W, SLICE), np.int32)                                  # column-major: [W, 64]
        v = np.zeros((W, SLICE))
        c[:, : r.size] = cols
        v[:, : r.size] = vals
        col_out.append(c.ravel())
        val_out.append(v.ravel())
        ptr.append(ptr[-1] + W * SLICE)
    if len(ptr) == 1:
        return SellSlices(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0),
                          np.arange(m, dtype=np.int32), False, np.zeros(0, np.int32))
    return SellSlices(np.asarray(ptr, np.int32), np.concatenate(rows_out), np.concatenate(col_out),
                      np.concatenate(val_out), np.flatnonzero(~in_slice).astype(np.int32), padded,
                      np.asarray(strides, np.int32))


def row_strides(K: sp.csr_matrix, rows: np.ndarray) -> np.ndarray:
    """For each row of ``rows``: the step s with col_k = col_0 + k s for all its entries
    (1 for a single entry), or 0 (no entries, or no constant step): the smallest and
    largest step between neighbouring entries of each row agree."""
    rows = np.asarray(rows, np.int64)
    out = np.zeros(rows.size, np.int32)
    if rows.size == 0:
        return out
    ip = K.indptr.astype(np.int64)
    lens = ip[rows + 1] - ip[rows]
    out[lens == 1] = 1
    two = lens >= 2
    if not two.any() or K.nnz < 2:
        return out
    d = np.diff(K.indices.astype(np.int64))                  # d[e] = col[e + 1] - col[e]
    cross = ip[1:-1] - 1                                      # steps from one row into the next
    cross = cross[(cross >= 0) & (cross < d.size)]
    lo, hi = d.copy(), d.copy()
    lo[cross], hi[cross] = np.iinfo(np.int64).max, np.iinfo(np.int64).min
    st = ip[rows[two]]                                        # a row of >= 2 entries: steps [st, end - 1)
    smin = np.minimum.reduceat(lo, st)
    smax = np.maximum.reduceat(hi, st)
    # reduceat runs each segment to the next start: clip it at the row's own end
    # (later rows' steps are masked as crossings only at row boundaries, so recompute
    # rows whose next listed start is not their own end)
    nxt = np.append(st[1:], d.size)
    own_end = ip[rows[two] + 1] - 1
    fix = np.flatnonzero(nxt != own_end)
    for k in fix:
        seg = d[st[k]: own_end[k]]
        smin[k], smax[k] = seg.min(), seg.max()
    out[np.flatnonzero(two)] = np.where(smin == smax, smin, 0).astype(np.int32)
    return out
    r2, l2 = rows[two], lens[two]
    step = ind[ip[r2] + 1] - ind[ip[r2]]
    # every entry of those rows against col_0 + k step
    seg = np.repeat(np.arange(r2.size), l2)
    k = np.arange(int(l2.sum())) - np.repeat(np.cumsum(l2) - l2, l2)
    e = np.repeat(ip[r2], l2) + k
    bad = np.bincount(seg[ind[e] != ind[ip[r2]][seg] + k * step[seg]], minlength=r2.size)
    out[np.flatnonzero(two)] = np.where(bad == 0, step, 0).astype(np.int32)
    return out


# Largest padding (entries of the SELL copy beyond nnz, relative to nnz) for which
# the primal step's SELL copy of K^T is built.
COLUMN_PADDING_MAX = 0.125


def build_columns(K: sp.csr_matrix, padding_max=None):
    """Slice pointers of a SELL-64 copy of K^T (K's columns, natural order): slice s holds
    columns 64 s .. 64 s + 63 and is as wide as the longest of them; the copy's entries
    are filled on the device (``pdlp_csr_to_sell64``).  None when the padding would exceed
    ``padding_max`` (default :data:`COLUMN_PADDING_MAX`) of K's entries, when K has no
    entries, or beyond int32 indexing."""
    n = K.shape[1]
    if K.nnz == 0 or n == 0:
        return None
    lens = np.bincount(K.indices, minlength=n)
    slices = -(-n // SLICE)
    padded = np.zeros(slices * SLICE, np.int64)
    padded[:n] = lens
    return _column_pointers(padded.reshape(slices, SLICE).max(axis=1), K.nnz, padding_max)


def _column_pointers(width: np.ndarray, nnz: int, padding_max=None):
    if padding_max is None:
        padding_max = COLUMN_PADDING_MAX
    ptr = np.zeros(width.size + 1, np.int64)
    ptr[1:] = np.cumsum(SLICE * np.asarray(width, np.int64))
    if ptr[-1] >= 2**31 or ptr[-1] > (1.0 + padding_max) * nnz:
        return None
    return ptr.astype(np.int32)


def build_columns_device(t_row_ptr, nnz: int, stream, padding_max=None):
    """:func:`build_columns` from K^T's row pointers on the device (``t_row_ptr``, a
    torch int32 tensor of n + 1 entries): the slice widths come from
    ``pdlp_sell64_widths`` (one wavefront per slice), the pointers from their running
    sum on the host.  Same result as :func:`build_columns` on K."""
    import torch
    from . import _lib as L
    from .device import stream_handle
    n = t_row_ptr.numel() - 1
    if nnz == 0 or n == 0:
        return None
    slices = -(-n // SLICE)
    with torch.cuda.stream(stream):
        width = torch.empty(slices, dtype=torch.int32, device=t_row_ptr.device)
        L.check(L.lib().pdlp_sell64_widths(n, t_row_ptr.data_ptr(), width.data_ptr(), stream_handle(stream)),
                "pdlp_sell64_widths")
        host = width.cpu()
    return _column_pointers(host.numpy(), nnz, padding_max)
