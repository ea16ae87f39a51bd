"""The oracle (oracle/nsem_oracle.c) is pinned to the reference: starting
from the same decks, its Euler split steps and its FV Poisson PCG reproduce
the reference build's output fields bit for bit (SHA-256 of each field, with
a norm check for diagnosis), and the same PCG iteration counts."""
import os

import numpy as np
import pytest

from _cases import AMR_CASES, EULER_CASES, POTENTIAL_CASES, write_potential_case
from _euler_setup import bc_table, load_case
from _golden import arrays, fixtures, sha
from nebulasem_amd import Mesh, read_field

FX = fixtures()


def run_oracle_euler(oracle, tmp_path, case):
    mesh, f = load_case(tmp_path, case)
    om = oracle.OracleMesh(mesh)
    prm = oracle.make_params(case["dt"], bc_table(f), viscosity=case.get("viscosity"))
    T = f["T"][0].reshape(-1).copy()
    p = f["p"][0].reshape(-1).copy()
    U = f["U"][0].reshape(-1).copy()
    p_ref, rho_ref, rho = oracle.euler_hydrostatic(om, prm, T, p)
    return mesh, om, prm, p_ref, rho_ref, rho, U, T, p


@pytest.mark.parametrize("case", EULER_CASES + AMR_CASES, ids=[c["name"] for c in EULER_CASES + AMR_CASES])
def test_oracle_euler_matches_reference(oracle, tmp_path, case):
    mesh, om, prm, p_ref, rho_ref, rho, U, T, p = run_oracle_euler(oracle, tmp_path, case)
    nb = mesh.info.nbcs_field
    done = 0
    fx = FX["euler"][case["name"]]["steps"]
    for s in case["check_steps"]:
        assert oracle.euler_split_steps(om, prm, p_ref, rho_ref, rho, U, T, p, s - done) == 0
        done = s
        for name, arr, nc in (("rho", rho, 1), ("U", U, 3), ("T", T, 1), ("p", p, 1)):
            v = arr[:nb * nc]
            ref = fx[str(s)][name]
            assert abs(np.linalg.norm(v) - ref["l2"]) <= 1e-12 * max(ref["l2"], 1e-300), (s, name)
            assert sha(v) == ref["sha256"], f"step {s} {name}: not bit-identical to the reference"


def test_oracle_euler_small_case_arrays(oracle, tmp_path):
    case = [c for c in EULER_CASES if c.get("keep_fields")][0]
    mesh, om, prm, p_ref, rho_ref, rho, U, T, p = run_oracle_euler(oracle, tmp_path, case)
    arr = arrays()
    nb = mesh.info.nbcs_field
    s = case["check_steps"][-1]
    oracle.euler_split_steps(om, prm, p_ref, rho_ref, rho, U, T, p, s)
    for name, a, nc in (("rho", rho, 1), ("U", U, 3), ("T", T, 1), ("p", p, 1)):
        assert np.array_equal(a[:nb * nc], arr[f"euler_{case['name']}_{name}{s}"]), name


def potential_inputs(tmp_path, case):
    d = os.path.join(str(tmp_path), case["name"])
    write_potential_case(d, case)
    mesh = Mesh.from_spec(os.path.join(d, "spec"), order=(0, 0, 0))
    U, kU, vU = read_field(mesh, os.path.join(d, "U0.txt"), 3)
    p, kp, vp = read_field(mesh, os.path.join(d, "p0.txt"), 1)
    # ghost values as applyExplicitBCs sets them after reading
    FO, FN = mesh.array("FO"), mesh.array("FN")
    boff, bf = mesh.array("boff"), mesh.array("bfaces")
    U, p = U.copy(), p.reshape(-1).copy()
    for b in range(len(mesh.boundaries)):
        fs = bf[boff[b]:boff[b + 1]]
        U[FN[fs]] = U[FO[fs]] if kU[b] == 0 else vU[b]
        p[FN[fs]] = vp[b, 0] if kp[b] == 2 else p[FO[fs]]
    return mesh, U, p, kp, vp


@pytest.mark.parametrize("case", POTENTIAL_CASES, ids=[c["name"] for c in POTENTIAL_CASES])
def test_oracle_poisson_matches_reference(oracle, tmp_path, case):
    mesh, U, p, kp, vp = potential_inputs(tmp_path, case)
    om = oracle.OracleMesh(mesh)
    st, it, r0, r = oracle.fv_poisson(om, kp, vp.reshape(-1), U.reshape(-1).copy(), p, 1e-10, 100000)
    ref = FX["potential"][case["name"]]
    assert st == 0
    assert it == ref["iterations"]
    assert abs(r - ref["final_residual"]) <= 1e-5 * ref["final_residual"]
    assert sha(p[:mesh.info.nbcs]) == ref["p_sha256"]
    # the velocity correction U -= gradf(p, true) the potential app closes with
    oracle.fv_correct_velocity(om, kp, vp.reshape(-1), p, U.reshape(-1))
    assert sha(U.reshape(-1)[:3 * mesh.info.nbcs]) == ref["U_sha256"]


def test_fv_gradient_closed_forms(oracle, tmp_path):
    """The FV Gauss gradient is exact for a linear p on cells whose faces are
    all interior (uniform grid: cds is the face-centre value), and it
    telescopes: sum cV grad p = sum over boundary faces of fN cds(p)."""
    case = dict(name="grad_lin", ndiv=(6, 5, 4), seed=5)
    mesh, U, p, kp, vp = potential_inputs(tmp_path, case)
    om = oracle.OracleMesh(mesh)
    n = mesh.info.nbcs
    cC = mesh.array("cC").reshape(-1, 3)
    a = np.array([0.7, -1.3, 2.1])
    p = (cC @ a).copy()
    kn = np.zeros_like(kp)   # all NEUMANN: ghosts copy their owners
    vn = np.zeros(len(kp))
    U0 = np.zeros((len(p), 3))
    U1 = U0.copy().reshape(-1)
    oracle.fv_correct_velocity(om, kn, vn, p.copy(), U1)
    grad = -U1.reshape(-1, 3)[:n]
    FO, FN, fI = mesh.array("FO"), mesh.array("FN"), mesh.array("fI")
    fi0, fi1, af = mesh.array("faceIndices0"), mesh.array("faceIndices1"), mesh.array("allFaces")
    inner = [c for c in range(n) if all(FN[af[f]] < n for f in range(fi0[c], fi1[c]))]
    assert len(inner) > 0
    assert np.abs(grad[inner] - a).max() <= 1e-12 * np.abs(a).max() * 10
    # telescoping with the ghost values the correction used (owner copies)
    pg = p.copy()
    pg[FN[FN >= n]] = pg[FO[FN >= n]]
    fN = mesh.array("fN").reshape(-1, 3)
    cV = mesh.array("cV")
    bnd = np.where(FN >= n)[0]
    lhs = (cV[:n, None] * grad).sum(axis=0)
    rhs = (fN[bnd] * (pg[FO[bnd]] * fI[bnd] + pg[FN[bnd]] * (1 - fI[bnd]))[:, None]).sum(axis=0)
    assert np.abs(lhs - rhs).max() <= 1e-9 * np.abs(rhs).max()


def test_poisson_against_dense_solve(oracle, tmp_path):
    """Brute force: the FV operator as a dense matrix; the PCG result solves it."""
    case = dict(name="pot_4", ndiv=(4, 3, 5), seed=3)
    mesh, U, p, kp, vp = potential_inputs(tmp_path, case)
    om = oracle.OracleMesh(mesh)
    st, it, r0, r = oracle.fv_poisson(om, kp, vp.reshape(-1), U.reshape(-1).copy(), p, 1e-14, 1000)
    n = mesh.info.nbcs
    FO, FN, fD = mesh.array("FO"), mesh.array("FN"), mesh.array("fD")
    fN = mesh.array("fN").reshape(-1, 3)
    A = np.zeros((n, n))
    b = np.zeros(n)
    xmin = mesh.boundaries.index("xmin")
    boff, bf = mesh.array("boff"), mesh.array("bfaces")
    dir_faces = set(bf[boff[xmin]:boff[xmin + 1]].tolist())
    Uc = U.reshape(-1, 3)
    for k in range(mesh.info.nfacets):
        c1, c2 = FO[k], FN[k]
        flux = 0.5 * (Uc[c1] + Uc[c2]) @ fN[k] if c2 < n else Uc[c2] @ fN[k]
        b[c1] -= flux
        if c2 < n:
            b[c2] += flux
            A[c1, c1] += fD[k]; A[c2, c2] += fD[k]
            A[c1, c2] -= fD[k]; A[c2, c1] -= fD[k]
        elif k in dir_faces:
            A[c1, c1] += fD[k]  # Dirichlet p = 0 on xmin
    x = np.linalg.solve(A, b)
    assert np.allclose(p[:n], x, rtol=1e-9, atol=1e-11 * np.abs(x).max())


def _log_numbers(lines):
    """Courant max/min/avg and mass/energy/volume losses from the reference's
    last two per-step log lines (euler.cpp:273-279)."""
    c = [ln for ln in lines if "Courant number" in ln][-1].split()
    m = [ln for ln in lines if "Mass loss" in ln][-1].split()
    cour = [c[c.index("Max:") + 1], c[c.index("Min:") + 1], c[c.index("Avg:") + 1]]
    loss = [m[m.index("Mass") + 2], m[m.index("Energy") + 2], m[m.index("Volume") + 2]]
    return cour, loss


@pytest.mark.parametrize("case", EULER_CASES, ids=[c["name"] for c in EULER_CASES])
def test_oracle_diagnostics_match_reference_log(oracle, tmp_path, case):
    """Courant numbers (%g) and mass/energy/volume losses (%.12g) after the
    run's last step, printed by the reference build, reproduced digit for
    digit from the oracle's state and its diagnostics restatement."""
    mesh, om, prm, p_ref, rho_ref, rho, U, T, p = run_oracle_euler(oracle, tmp_path, case)
    d0 = oracle.euler_diagnostics(om, prm, p_ref, rho, U, T, p)
    steps = max(case["check_steps"])
    assert oracle.euler_split_steps(om, prm, p_ref, rho_ref, rho, U, T, p, steps) == 0
    d1 = oracle.euler_diagnostics(om, prm, p_ref, rho, U, T, p)
    cour, loss = _log_numbers(FX["euler"][case["name"]]["log_tail"])
    assert ["%g" % v for v in d1[:3]] == cour
    mine = ["%.12g" % ((d0[k] - d1[k]) / d0[k]) for k in (3, 4, 5)]
    assert mine[1:] == loss[1:], (mine, loss)
    # the mass loss is round-off (~1e-16): equal to the printed digits too
    assert mine[0] == loss[0], (mine, loss)
