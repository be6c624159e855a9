"""A snapshot written and read back through the library's on-disk formats
(output_data / input_data) drives the step to the same per-cell rates, bit
for bit, as the workload it was written from."""
import numpy as np
import pytest

from argot_amd import radiation as R
from helpers import small_workload

pytestmark = pytest.mark.gpu


def test_snapshot_step_equals_direct_step(require_gpu, tmp_path):
    w = small_workload(n=16, nsrc=5, seed=51)
    a = R.PointSourceStep(w)
    a.step()
    ra = a.rates()

    freq, cs = R.freq_and_csect()
    run = R.run_param_for(w, freq, cs)
    src = R.blackbody_sources(w.src_pos, w.src_rate, w.src_temp, freq)
    prefix = str(tmp_path / "ic")
    R.write_snapshot(prefix, R.fill_mesh(w), src, run)
    mesh, src2, run2 = R.read_snapshot(prefix, w.n, max_sources=w.nsrc)
    assert run2.nsrc == w.nsrc
    b = R.PointSourceStep.from_state(mesh, src2, run2, w.helium_fact)
    b.step()
    np.testing.assert_array_equal(b.rates(), ra)
