"""Cross-check of the row-wise product against rocSPARSE's SpGEMM (the ROCm
library) on the GPU: same structure, values within 1e-12 of each other
relative to the sum of the terms (the oracle gives that sum)."""
import numpy as np
import pytest

import oracle
from helpers import powerlaw_rows, uniform_rows
from sparsews_amd import rocsparse_cmp
from sparsews_amd.device import DeviceCSR, spgemm_rowwise

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not rocsparse_cmp.available(), reason="librocsparse.so not found")]
TOL = 1e-12


@pytest.mark.parametrize("kind", ["uniform", "powerlaw"])
def test_rowwise_matches_rocsparse(ctx, kind):
    print(f"case: rocSPARSE cross-check {kind}", flush=True)
    if kind == "uniform":
        A, B = uniform_rows(3000, 2500, 2, 40, 41, 0), uniform_rows(2500, 5000, 2, 40, 41, 1)
    else:
        A, B = powerlaw_rows(3000, 2500, 1.8, 800, 42, 0), powerlaw_rows(2500, 5000, 1.8, 800, 42, 1)
    dA, dB = DeviceCSR.upload(ctx, A), DeviceCSR.upload(ctx, B)
    C = spgemm_rowwise(ctx, dA, dB).download()
    r = rocsparse_cmp.RocsparseSpgemm(ctx, dA, dB)
    try:
        r.run()
        pos, crd, vals = r.result()
    finally:
        r.close()
    np.testing.assert_array_equal(pos, np.asarray(C.pos))
    np.testing.assert_array_equal(crd, C.crd)
    _, absum = oracle.spgemm_refsum(A, B, C.pos, C.crd)
    assert oracle.max_value_error(vals, C.vals, absum) <= TOL
