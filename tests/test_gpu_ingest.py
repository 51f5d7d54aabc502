"""GPU: the device CSR transpose of problem ingestion (needs an MI355X).

``pdlp_csr_transpose`` (through the C ABI) against SciPy's canonical
transpose, which the reference builds on the host (``pdhg.py:179``,
``pdhg_gpu.py:137,161,186``): row pointers, indices and values must agree bit
for bit, in FP32 and FP64, including empty rows and columns, no entries, a
dense column and the transport LP's 3,000 x 2,000,000 matrix.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from sovereign_solver_amd.engine.device import csr_transpose, dev_tensor
from sovereign_solver_amd.engine.forms import box_form
from sovereign_solver_amd.workloads import transportation_lp

pytestmark = pytest.mark.gpu


def _check(A, dtype):
    A = sp.csr_matrix(A)
    A.sum_duplicates()
    A.sort_indices()
    ref = A.T.tocsr()
    ref.sort_indices()
    dv = torch.device("cuda", 0)
    s = torch.cuda.Stream(device=dv)
    rp = dev_tensor(A.indptr, torch.int32, dv)
    ci = dev_tensor(A.indices, torch.int32, dv)
    v = dev_tensor(A.data, dtype, dv)
    trp, tci, tv = csr_transpose(rp, ci, v, A.shape[1], s)
    s.synchronize()
    np.testing.assert_array_equal(trp.cpu().numpy(), ref.indptr)
    np.testing.assert_array_equal(tci.cpu().numpy(), ref.indices)
    np.testing.assert_array_equal(tv.cpu().numpy(), ref.data.astype(np.float32 if dtype == torch.float32 else np.float64))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("shape,density,seed", [((1, 1), 1.0, 0), ((7, 13), 0.3, 1), ((300, 500), 0.02, 2),
                                                ((2000, 90), 0.2, 3), ((50, 70000), 0.001, 4)])
def test_random_matrices(gpu, dtype, shape, density, seed):
    A = sp.random(*shape, density, random_state=seed, format="csr")
    A.data = np.random.default_rng(seed).standard_normal(A.nnz)
    _check(A, dtype)


def test_empty_lines_and_no_entries(gpu):
    A = sp.csr_matrix(np.array([[0.0, 2.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0], [1.0, 3.0, 0.0, 0.0]]))
    _check(A, torch.float64)                         # empty row 1, empty columns 2 and 3
    _check(sp.csr_matrix((5, 9)), torch.float32)     # no entries: every pointer 0


def test_dense_column(gpu):
    rng = np.random.default_rng(5)
    A = sp.random(4000, 300, 0.01, random_state=5, format="lil")
    A[:, 17] = rng.standard_normal((4000, 1))
    _check(A.tocsr(), torch.float32)


def test_transport_matrix(gpu):
    K = box_form(transportation_lp(1000, 2000)).K
    _check(K, torch.float32)
