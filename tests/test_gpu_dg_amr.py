"""2:1 non-conforming (AMR) meshes on the HIP method-of-lines path, with
and without artificial viscosity: the stage kernels with the mortar pre-pass (scatter of the coarse side to each
refined face, Rusanov flux there, gather back to the coarse side) against
the oracle's SSP-RK3 composition, whose spatial operator on these decks is
bit-identical to the reference's euler (tests/test_oracle.py).

Tolerances are those of the conforming decks (test_gpu_dg_rk3.py): the
device forms the same operator with different rounding (conserved
perturbation variables, p' by the exact-difference formula, normal fluxes
interpolated instead of flux vectors)."""
import numpy as np
import pytest

from _cases import AMR_CASES
from test_gpu_dg_split import relerr, setup

pytestmark = pytest.mark.gpu

@pytest.mark.parametrize("case", AMR_CASES, ids=[c["name"] for c in AMR_CASES])
def test_amr_rk3_fp64_matches_oracle(oracle, tmp_path, case):
    print(f"[gpu] amr rk3 fp64 {case['name']}", flush=True)
    mesh, om, prm, dg, (p_ref, rho_ref, rho, U, T, p) = setup(oracle, tmp_path, case)
    nb = mesh.info.nbcs_field
    assert (mesh.array("FMC") > 0).sum() > 0
    oracle.euler_ssprk3_steps(om, prm, p_ref, rho_ref, rho, U, T, p, 3)
    dg.rk3_steps(3, 64)
    g_rho, g_U, g_T, g_p = dg.get_state()
    tol = {"rho": 1e-13, "U": 2e-9, "T": 1e-11}
    for name, a, b in (("rho", g_rho, rho[:nb]), ("U", g_U, U[:3 * nb]), ("T", g_T, T[:nb])):
        l2, linf = relerr(a, b)
        print(f"  {name}: rel L2 {l2:.2e} rel max {linf:.2e}", flush=True)
    for name, a, b in (("rho", g_rho, rho[:nb]), ("U", g_U, U[:3 * nb]), ("T", g_T, T[:nb])):
        l2, linf = relerr(a, b)
        assert l2 <= tol[name], (name, l2, linf)
    assert np.linalg.norm(g_p - p[:nb]) / np.linalg.norm(p_ref[:nb]) <= 1e-13
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
