"""2:1 non-conforming (AMR) meshes on the HIP method-of-lines path, with
and without artificial viscosity: the stage kernels with the mortar pre-pass (scatter of the coarse side to each
refined face, Rusanov flux there, gather back to the coarse side) against
the oracle's SSP-RK3 composition, whose spatial operator on these decks is
bit-identical to the reference's euler (tests/test_oracle.py).

Tolerances are those of the conforming decks (test_gpu_dg_rk3.py, 1e-12
relative for every field): the device forms the same operator with
different rounding (sum-factorised volume term, normal fluxes interpolated
instead of flux vectors)."""
import numpy as np
import pytest

from _cases import AMR_CASES
from test_gpu_dg_rk3 import check_rk3_against_oracle
from test_gpu_dg_split import relerr, setup

pytestmark = pytest.mark.gpu

@pytest.mark.parametrize("case", AMR_CASES, ids=[c["name"] for c in AMR_CASES])
def test_amr_rk3_fp64_matches_oracle(oracle, tmp_path, case):
    print(f"[gpu] amr rk3 fp64 {case['name']}", flush=True)
    mesh, om, prm, dg, (p_ref, rho_ref, rho, U, T, p) = setup(oracle, tmp_path, case)
    nb = mesh.info.nbcs_field
    assert (mesh.array("FMC") > 0).sum() > 0
    check_rk3_against_oracle(oracle, om, prm, dg, nb, p_ref, rho_ref, rho, U, T, p, 3)
    dg.close()


@pytest.mark.parametrize("case", AMR_CASES, ids=[c["name"] for c in AMR_CASES])
def test_amr_rk3_fp32_against_fp64(oracle, tmp_path, case):
    """The FP32-perturbation form with mortars against FP64 (Theta' measured
    against Theta_ref, as for the conforming decks)."""
    print(f"[gpu] amr rk3 fp32 vs fp64 {case['name']}", flush=True)
    mesh, om, prm, dg, state = setup(oracle, tmp_path, case)
    dg.rk3_steps(10, 64)
    q64 = dg.get_conserved()
    p_ref = dg.background()[0]
    mesh2, om2, prm2, dg2, state2 = setup(oracle, tmp_path / "b", case)
    dg2.rk3_steps(10, 32)
    q32 = dg2.get_conserved()
    P0, R, gamma = 101325.0, 1004.67 - 715.5, 1004.67 / 715.5
    theta_ref = (P0 / R) * (p_ref / P0) ** (1.0 / gamma)
    for nm, a, b in zip(("rho'", "m", "Theta'"), q32, q64):
        l2, linf = relerr(a, b)
        print(f"  fp32 vs fp64 {nm}: rel L2 {l2:.2e} rel max {linf:.2e}", flush=True)
        if nm == "Theta'":
            assert np.linalg.norm(a - b) / np.linalg.norm(theta_ref) <= 1e-9
        else:
            assert l2 <= 1e-5, (nm, l2)
    dg.close()
    dg2.close()
