"""nsem_dg_diagnostics_noalloc (scratch in the library's static device
array, no allocation) against nsem_dg_diagnostics (the handle's buffers):
the same kernels over the same state, so the six values are identical.
Two calls in a row on two problems, the second one smaller, each checked
against the existing entry point, then the first problem again (the
scratch keeps the previous call's contents, which every call rewrites)."""
import numpy as np
import pytest

from _cases import EULER_CASES
from test_gpu_dg_split import setup

pytestmark = pytest.mark.gpu


def test_noalloc_diagnostics_match_existing_entry_point(oracle, tmp_path):
    print("[gpu] diagnostics without allocation", flush=True)
    big = [c for c in EULER_CASES if c["name"] == "e3_n7"][0]
    small = [c for c in EULER_CASES if c["name"] == "e3_n3"][0]
    mesh1, om1, prm1, dg1, st1 = setup(oracle, tmp_path / "big", big)
    mesh2, om2, prm2, dg2, st2 = setup(oracle, tmp_path / "small", small)
    assert mesh2.info.nbcs_field < mesh1.info.nbcs_field
    dg1.split_steps(2)
    dg2.rk3_steps(2, 64)
    for dg in (dg1, dg2, dg1):
        want = dg.diagnostics()
        got = dg.diagnostics_noalloc()
        print(f"  nb {dg.nb}: {got}", flush=True)
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, want), (got, want)
    dg1.close()
    dg2.close()
