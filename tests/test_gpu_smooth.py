"""smooth_photoion_rate on the GPU equals the oracle's smoothing of the
step's own rates bit for bit, on the device and in the host mesh."""
import numpy as np
import pytest

import oracle as O
from argot_amd import abi
from argot_amd import radiation as R
from helpers import small_workload

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [16, 12])
def test_smoothing_matches_oracle(require_gpu, n):
    w = small_workload(n=n, nsrc=5, seed=41)
    st = R.PointSourceStep(w)
    st.step()
    raw = st.rates()
    raw[~np.isfinite(raw)] = 0.0          # source cells: keep the sums finite
    v = abi.mesh_view(st.mesh)
    pc = v["prev_chem"]
    for k, f in enumerate(R.RATE_FIELDS):
        pc[f] = raw[:, k]
    v["prev_chem"] = pc
    import ctypes as C
    from argot_amd._lib import lib
    # merge adds the host rates to the device ones: start the device from zero
    lib().zero_out_photoion_rate_cuda(C.byref(st.cmem), C.byref(st.cpar), C.byref(st.run))
    lib().merge_photoion_data(st.mesh, C.byref(st.cmem), C.byref(st.cpar), C.byref(st.run))
    st.smooth()
    host = np.stack([abi.mesh_view(st.mesh)["prev_chem"][f] for f in R.RATE_FIELDS], axis=1)
    ref = O.smooth_rates(raw, (n, n, n))
    np.testing.assert_array_equal(host, ref)
    np.testing.assert_array_equal(st.rates(), ref)
