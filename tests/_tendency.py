"""Cases and states of the method-of-lines residual pin: the oracle's
residual against the reference's own divergence operators evaluated at the
same mid-RK state (tests/golden/make_tendency.py writes the fixtures)."""
import os

import numpy as np

from _cases import AMR_CASES, EULER_CASES
from _euler_setup import bc_table, load_case

_BY_NAME = {c["name"]: c for c in EULER_CASES + AMR_CASES}
TENDENCY_CASES = [dict(_BY_NAME[n], tend_steps=2) for n in
                  ("cfg0_2d_n4", "e3_n3", "e3_visc_n3", "e3_trilinear_n3", "e3_aniso_bubbles",
                   "amr3_n2", "amr3_n2_rev", "amr3_n3_visc")]


def prepare_case(oracle, tmp, case):
    """The deck in tmp/<name> and the oracle's initial state on every node."""
    mesh, f = load_case(tmp, case)
    d = os.path.join(str(tmp), case["name"])
    om = oracle.OracleMesh(mesh)
    prm = oracle.make_params(case["dt"], bc_table(f), viscosity=case.get("viscosity"))
    T = f["T"][0].reshape(-1).copy()
    p = f["p"][0].reshape(-1).copy()
    U = f["U"][0].reshape(-1).copy()
    p_ref, rho_ref, rho = oracle.euler_hydrostatic(om, prm, T, p)
    return d, mesh, om, prm, (p_ref, rho_ref, rho, U, T, p)


def mid_rk_state(oracle, om, prm, state0, case):
    """`tend_steps` SSP-RK3 steps of the oracle, then the first stage's
    forward-Euler update q + dt L(q): the state the second stage evaluates
    its residual at.  Returns (rho, U, theta') on every node, the ghost
    nodes set from the boundary conditions (the split step reads them as
    given)."""
    p_ref, rho_ref, rho, U, T, p = [a.copy() for a in state0]
    assert oracle.euler_ssprk3_steps(om, prm, p_ref, rho_ref, rho, U, T, p, case["tend_steps"]) == 0
    L, _ = oracle.euler_mol_rhs(om, prm, p_ref, rho_ref, rho, U, T)
    nb = om.info.nbcs_field
    T0, dt = prm.T0, prm.dt
    r0 = rho[:nb]
    m0 = U[:3 * nb].reshape(nb, 3) * r0[:, None]
    th0 = r0 * (T[:nb] + T0)
    r1 = r0 + dt * L[:nb, 0]
    m1 = m0 + dt * L[:nb, 1:4]
    th1 = th0 + dt * L[:nb, 4]
    rho2, U2, T2 = rho.copy(), U.copy(), T.copy()
    rho2[:nb] = r1
    U2[:3 * nb] = (m1 / r1[:, None]).reshape(-1)
    T2[:nb] = th1 / r1 - T0
    oracle.euler_mol_rhs(om, prm, p_ref, rho_ref, rho2, U2, T2)  # sets the ghosts in place
    return rho2, U2, T2
