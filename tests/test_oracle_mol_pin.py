"""The oracle's method-of-lines residual pinned against the reference.

1. At a mid-RK state of eight decks (2D and 3D, viscous, non-affine,
   anisotropic orders, 2:1 mortars in both orientations), the oracle's
   Rusanov divergences r of the density, momentum and rho theta fluxes are
   bit-identical to what the reference's own divf / rusanov / gradf<true>
   return at that state (oracle/ref_tendency.cpp, fixtures written by
   tests/golden/make_tendency.py).  L = -r/cV (+ (rho - rho_ref) g) is
   formed from r the same way in both.
2. The split step (bit-identical to the reference's euler) is a first-order
   consistent forward-Euler step of the same operator: (q_split(dt) - q)/dt
   tends to L(q) at first order as dt halves.  A wrong flux (theta' in place
   of theta, a missing pressure term) leaves an O(1) gap and fails this.
3. The cancellation-free forms (mol_exact: p' without subtracting two
   ~1e5 Pa numbers, the gradient of theta' rather than theta' + T0), which
   the device's method-of-lines kernels use, agree with the reference's
   expressions to the rounding those cancel."""
import json
import os
import subprocess

import numpy as np
import pytest

from _tendency import TENDENCY_CASES, mid_rk_state, prepare_case

HERE = os.path.dirname(os.path.abspath(__file__))
FX = json.load(open(os.path.join(HERE, "golden", "tendency_fixtures.json")))
ARR = np.load(os.path.join(HERE, "golden", "tendency_arrays.npz"))
REF_BIN = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "euler_tendency")


def sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def oracle_tendency(oracle, tmp_path, case):
    d, mesh, om, prm, state0 = prepare_case(oracle, tmp_path, case)
    nb = mesh.info.nbcs_field
    rho, U, T = mid_rk_state(oracle, om, prm, state0, case)
    p_ref, rho_ref = state0[0], state0[1]
    state_sha = sha(np.concatenate([rho[:nb], U[:3 * nb], T[:nb]]))
    L, r = oracle.euler_mol_rhs(om, prm, p_ref, rho_ref, rho.copy(), U.copy(), T.copy())
    out = {"r_rho": r[:nb, 0].copy(), "r_m": r[:nb, 1:4].reshape(-1).copy(), "r_T": r[:nb, 4].copy()}
    return d, mesh, (rho, U, T), state_sha, out, L


@pytest.mark.parametrize("case", TENDENCY_CASES, ids=[c["name"] for c in TENDENCY_CASES])
def test_mol_residual_bit_identical_to_reference(oracle, tmp_path, case):
    fx = FX[case["name"]]
    d, mesh, state, state_sha, out, L = oracle_tendency(oracle, tmp_path, case)
    assert state_sha == fx["state_sha256"], "the mid-RK state is not the one the fixture was made from"
    for k, v in out.items():
        if sha(v) != fx["sha256"][k]:
            key = f"{case['name']}_{k}"
            msg = f"{k}: not bit-identical to the reference"
            if key in ARR:
                ref = ARR[key]
                msg += f" (rel L2 {np.linalg.norm(v - ref) / np.linalg.norm(ref):.2e})"
            pytest.fail(msg)
        assert abs(np.linalg.norm(v) - fx["l2"][k]) <= 1e-15 * fx["l2"][k]


@pytest.mark.skipif(not os.path.exists(REF_BIN), reason="reference build (oracle/_ref) not present")
def test_mol_residual_live_reference(oracle, tmp_path):
    """The same comparison against the reference run now (this container),
    on the 2D deck and a viscous mortar deck."""
    import shutil
    from _euler_setup import amr_grid
    for case in (TENDENCY_CASES[0], TENDENCY_CASES[-1]):
        d, mesh, (rho, U, T), state_sha, out, L = oracle_tendency(oracle, tmp_path / case["name"], case)
        nb = mesh.info.nbcs_field
        if case.get("refine"):
            src = amr_grid(case)
            shutil.copy(src, os.path.join(d, "grid_0" + os.path.splitext(src)[1]))
        else:
            subprocess.run([os.path.join(os.path.dirname(REF_BIN), "mesh"), "spec", "-o", "grid_0.bin"],
                           cwd=d, check=True, capture_output=True)
        np.concatenate([rho[:nb], U[:3 * nb], T[:nb]]).tofile(os.path.join(d, "state.bin"))
        subprocess.run([REF_BIN, "controls", "state.bin", "tend.bin"], cwd=d, check=True, capture_output=True)
        ref = np.fromfile(os.path.join(d, "tend.bin"))
        assert np.array_equal(out["r_rho"], ref[:nb])
        assert np.array_equal(out["r_m"], ref[nb:4 * nb])
        assert np.array_equal(out["r_T"], ref[4 * nb:5 * nb])


@pytest.mark.parametrize("name", ["cfg0_2d_n4", "e3_visc_n3", "amr3_n2"])
def test_split_step_is_consistent_with_mol_residual(oracle, tmp_path, name):
    """(q_split(dt) - q)/dt -> L(q), first order in dt: the split step's
    density pass is exactly q + dt L_rho(q), its momentum and theta passes
    differ from q + dt L(q) by terms proportional to dt (they use rho^{n+1}
    in the pressure, the source and the Rusanov state)."""
    case = [c for c in TENDENCY_CASES if c["name"] == name][0]
    d, mesh, om, prm, state0 = prepare_case(oracle, tmp_path, case)
    nb = mesh.info.nbcs_field
    rho, U, T = mid_rk_state(oracle, om, prm, state0, case)
    p_ref, rho_ref = state0[0], state0[1]
    L, _ = oracle.euler_mol_rhs(om, prm, p_ref, rho_ref, rho.copy(), U.copy(), T.copy())
    T0 = prm.T0
    q = np.concatenate([rho[:nb], (U[:3 * nb].reshape(nb, 3) * rho[:nb, None]).T.reshape(-1),
                        rho[:nb] * (T[:nb] + T0)])
    Lq = np.concatenate([L[:nb, 0], L[:nb, 1:4].T.reshape(-1), L[:nb, 4]])
    errs = []
    dt0 = case["dt"] / 4
    for k in range(4):
        dt = dt0 / 2 ** k
        prm.dt = dt
        r2, U2, T2, p2 = rho.copy(), U.copy(), T.copy(), np.zeros_like(rho)
        assert oracle.euler_split_steps(om, prm, p_ref, rho_ref, r2, U2, T2, p2, 1) == 0
        q2 = np.concatenate([r2[:nb], (U2[:3 * nb].reshape(nb, 3) * r2[:nb, None]).T.reshape(-1),
                             r2[:nb] * (T2[:nb] + T0)])
        blocks = [slice(0, nb), slice(nb, 4 * nb), slice(4 * nb, 5 * nb)]
        e = [np.linalg.norm((q2[b] - q[b]) / dt - Lq[b]) / np.linalg.norm(Lq[b]) for b in blocks]
        errs.append(e)
    errs = np.array(errs)
    print(name, errs)
    # density: the split's density pass is the same forward-Euler step; only
    # the rounding of (q + dt L - q)/dt remains, which grows as 1/dt
    assert np.all(errs[:, 0] < 1e-5)
    assert np.all(errs[1:, 0] > errs[:-1, 0])
    # momentum, rho theta: first order, error halves with dt
    for v in (1, 2):
        ratios = errs[:-1, v] / errs[1:, v]
        assert np.all((ratios > 1.8) & (ratios < 2.2)), (v, errs[:, v])
        assert errs[-1, v] < 1e-3


@pytest.mark.parametrize("name", ["cfg0_2d_n4", "e3_n3", "amr3_n2"])
def test_exact_pressure_form_agrees_with_reference_form(oracle, tmp_path, name):
    """p_ref expm1(gamma log1p(Theta'/Theta_ref)) and P0 (rho theta R/P0)^gamma
    - p_ref are the same quantity; they differ by the rounding of the full
    pressure (a few ulp of ~1e5 Pa), so the residuals they give differ only
    through that rounding."""
    case = [c for c in TENDENCY_CASES if c["name"] == name][0]
    d, mesh, om, prm, state0 = prepare_case(oracle, tmp_path, case)
    nb = mesh.info.nbcs_field
    rho, U, T = mid_rk_state(oracle, om, prm, state0, case)
    p_ref, rho_ref = state0[0], state0[1]
    Tp = rho[:nb] * (T[:nb] + prm.T0) - rho_ref[:nb] * prm.T0
    L0, r0 = oracle.euler_mol_rhs(om, prm, p_ref, rho_ref, rho.copy(), U.copy(), T.copy())
    prm.mol_exact = 1
    L1, r1 = oracle.euler_mol_rhs(om, prm, p_ref, rho_ref, rho.copy(), U.copy(), T.copy(), Tp=Tp)
    # density and rho theta do not involve p: identical
    assert np.array_equal(L0[:nb, 0], L1[:nb, 0]) and np.array_equal(L0[:nb, 4], L1[:nb, 4])
    # momentum: the pressure gradient moved by the full pressure's rounding
    dm = np.linalg.norm(L1[:nb, 1:4] - L0[:nb, 1:4]) / np.linalg.norm(L0[:nb, 1:4])
    print(name, "momentum residual, exact vs reference pressure form:", dm)
    assert dm < 1e-6
