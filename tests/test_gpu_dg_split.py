"""HIP split-step (reference-mode) Euler path vs the oracle: every field
within 1e-12 of the oracle (norm-relative L2 and max-norm), on the same decks
whose oracle results are bit-identical to the reference build."""
import sys

import numpy as np
import pytest

from _cases import EULER_CASES
from _euler_setup import bc_table, load_case
from nebulasem_amd import DGEuler, euler_params

pytestmark = pytest.mark.gpu

TOL = 1e-12


def relerr(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300), np.abs(a - b).max() / scale


def setup(oracle, tmp_path, case):
    mesh, f = load_case(tmp_path, case)
    nb = mesh.info.nbcs_field
    om = oracle.OracleMesh(mesh)
    prm = oracle.make_params(case["dt"], bc_table(f))
    T = f["T"][0].reshape(-1).copy()
    p = f["p"][0].reshape(-1).copy()
    U = f["U"][0].reshape(-1).copy()
    p_ref, rho_ref, rho = oracle.euler_hydrostatic(om, prm, T, p)
    dg = DGEuler.from_fields(mesh, euler_params(case["dt"]), f)
    pr, rr = dg.background()
    assert np.array_equal(pr, p_ref[:nb]) and np.array_equal(rr, rho_ref[:nb])
    rho_gpu = dg.initial_density(T[:nb], p[:nb])
    assert np.array_equal(rho_gpu, rho[:nb])
    dg.set_state(rho[:nb], U[:3 * nb], T[:nb], p[:nb])
    return mesh, om, prm, dg, (p_ref, rho_ref, rho, U, T, p)


@pytest.mark.parametrize("case", EULER_CASES, ids=[c["name"] for c in EULER_CASES])
def test_split_steps_match_oracle(oracle, tmp_path, case):
    print(f"[gpu] split steps {case['name']}", flush=True)
    mesh, om, prm, dg, (p_ref, rho_ref, rho, U, T, p) = setup(oracle, tmp_path, case)
    nb = mesh.info.nbcs_field
    done = 0
    for s in case["check_steps"][:2]:
        oracle.euler_split_steps(om, prm, p_ref, rho_ref, rho, U, T, p, s - done)
        dg.split_steps(s - done)
        done = s
        g_rho, g_U, g_T, g_p = dg.get_state()
        for name, a, b in (("rho", g_rho, rho[:nb]), ("U", g_U, U[:3 * nb]), ("T", g_T, T[:nb]),
                           ("p", g_p, p[:nb])):
            l2, linf = relerr(a, b)
            print(f"  step {s} {name}: rel L2 {l2:.2e} rel max {linf:.2e}", flush=True)
            assert l2 <= TOL and linf <= TOL, (s, name, l2, linf)
    dg.close()


def test_resting_hydrostatic_state_stays_at_rest(oracle, tmp_path):
    """theta' = 0, U = 0: the perturbation form keeps the discrete rest state
    to round-off (the flux carries p', the source (rho - rho_ref) g)."""
    case = dict(EULER_CASES[1], name="rest", T=["uniform 0"])
    mesh, om, prm, dg, state = setup(oracle, tmp_path, case)
    dg.split_steps(5)
    rho, U, T, p = dg.get_state()
    assert np.abs(U).max() < 1e-9
    assert np.abs(T).max() < 1e-9
