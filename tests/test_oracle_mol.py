"""The oracle's method-of-lines SSP-RK3 composition (no reference
counterpart) pinned by what the mathematics fixes: third-order temporal
convergence (a wrong Shu-Osher coefficient breaks it), discrete mass
conservation with no-flux walls, and the hydrostatic rest state."""
import numpy as np

from _cases import EULER_CASES
from test_oracle import run_oracle_euler


def case_n3():
    return dict([c for c in EULER_CASES if c["name"] == "e3_n3"][0])


def test_ssprk3_third_order_in_time(oracle, tmp_path):
    case = case_n3()
    finals = []
    T_end = 0.08
    for nsteps in (2, 4, 8, 16):
        case["dt"] = T_end / nsteps
        mesh, om, prm, p_ref, rho_ref, rho, U, T, p = run_oracle_euler(oracle, tmp_path / str(nsteps), case)
        assert oracle.euler_ssprk3_steps(om, prm, p_ref, rho_ref, rho, U, T, p, nsteps) == 0
        nb = mesh.info.nbcs_field
        finals.append(np.concatenate([rho[:nb], U[:3 * nb], T[:nb]]))
    e = [np.linalg.norm(finals[i] - finals[i + 1]) for i in range(3)]
    rates = [np.log2(e[i] / e[i + 1]) for i in range(2)]
    assert all(2.7 < r < 3.3 for r in rates), rates


def test_ssprk3_conserves_mass(oracle, tmp_path):
    case = case_n3()
    mesh, om, prm, p_ref, rho_ref, rho, U, T, p = run_oracle_euler(oracle, tmp_path, case)
    nb = mesh.info.nbcs_field
    cV = mesh.array("cV")[:nb]
    m0 = np.sum(cV * rho[:nb])
    oracle.euler_ssprk3_steps(om, prm, p_ref, rho_ref, rho, U, T, p, 10)
    m1 = np.sum(cV * rho[:nb])
    assert abs(m1 - m0) / m0 < 1e-14
    assert np.abs(U[:3 * nb]).max() > 1e-6  # the bubble did move


def test_ssprk3_rest_state(oracle, tmp_path):
    case = dict(case_n3(), T=["uniform 0"])
    mesh, om, prm, p_ref, rho_ref, rho, U, T, p = run_oracle_euler(oracle, tmp_path, case)
    oracle.euler_ssprk3_steps(om, prm, p_ref, rho_ref, rho, U, T, p, 5)
    nb = mesh.info.nbcs_field
    assert np.abs(U[:3 * nb]).max() < 1e-9 and np.abs(T[:nb]).max() < 1e-9


def test_ssprk3_viscous_limits(oracle, tmp_path):
    """The viscous method-of-lines operator: viscosity 0 reproduces the
    inviscid steps bit for bit (the viscous terms are exactly -0 mu grad),
    mass stays conserved (the density equation has no viscous flux), and the
    change the viscous terms make is linear in the viscosity at first order
    (doubling a small viscosity doubles the departure from the inviscid run)."""
    base = dict([c for c in EULER_CASES if c["name"] == "e3_visc_n3"][0])
    runs = {}
    for tag, visc in (("inviscid", None), ("zero", 0.0), ("v1", 0.5), ("v2", 1.0)):
        case = dict(base, viscosity=visc) if visc is not None else {k: v for k, v in base.items() if k != "viscosity"}
        mesh, om, prm, p_ref, rho_ref, rho, U, T, p = run_oracle_euler(oracle, tmp_path / tag, case)
        nb = mesh.info.nbcs_field
        cV = mesh.array("cV")[:nb]
        m0 = np.sum(cV * rho[:nb])
        assert oracle.euler_ssprk3_steps(om, prm, p_ref, rho_ref, rho, U, T, p, 6) == 0
        assert abs(np.sum(cV * rho[:nb]) - m0) / m0 < 1e-14, tag
        runs[tag] = np.concatenate([U[:3 * nb], T[:nb]])
    assert np.array_equal(runs["inviscid"], runs["zero"])
    d1 = np.linalg.norm(runs["v1"] - runs["inviscid"])
    d2 = np.linalg.norm(runs["v2"] - runs["inviscid"])
    assert d1 > 0 and abs(d2 / d1 - 2) < 0.02, (d1, d2)
