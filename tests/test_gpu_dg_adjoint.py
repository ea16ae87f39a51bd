"""Tangent-linear model and discrete adjoint of the FP64 method-of-lines
SSP-RK3 step (no reference counterpart: pinned by finite differences of the
forward map and by the dot-product identity <J v, w> = <v, J^T w>).

Tolerances.  Dot-product identity: 1e-12 relative for one residual
evaluation, 1e-10 for several RK3 steps (both sides are exact derivatives of
the same FP64 map, they differ by rounding only).  Finite differences:
central differences with a step chosen per case, 1e-6 relative (truncation
O(h^2) and rounding O(eps/h) both well below that)."""
import numpy as np
import pytest

from _cases import AMR_CASES, EULER_CASES
from test_gpu_dg_split import setup

pytestmark = pytest.mark.gpu

# plus the inviscid 2:1 non-conforming decks (the mortar pre-pass in the
# tangent-linear stages, its transpose in the vector-Jacobian product)
CASES = ([c for c in EULER_CASES if c["name"] in ("cfg0_2d_n4", "e3_n3", "e3_wall_n3")]
         + [c for c in AMR_CASES if c.get("viscosity") is None])


def conserved(dg):
    rp, m, tp = dg.get_conserved()
    m = m.reshape(-1, 3)
    return np.stack([rp, m[:, 0], m[:, 1], m[:, 2], tp])


def set_q(dg, q):
    dg.set_conserved(q[0], np.ascontiguousarray(q[1:4].T), q[4])


def directions(q, seed):
    """Random directions scaled per conserved variable to the state's size."""
    rng = np.random.default_rng(seed)
    scale = np.array([max(np.abs(q[v]).max(), 1e-3) for v in range(5)])[:, None]
    return rng.standard_normal(q.shape) * scale, rng.standard_normal(q.shape)


def advanced(oracle, tmp_path, case, steps=2):
    mesh, om, prm, dg, _ = setup(oracle, tmp_path, case)
    dg.rk3_steps(steps, 64)   # a state with non-zero velocities
    return dg, conserved(dg)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_residual_dot_product(oracle, tmp_path, case):
    print(f"[gpu] adjoint residual dot-product {case['name']}", flush=True)
    dg, q = advanced(oracle, tmp_path, case)
    v, w = directions(q, 1)
    L, Jv = dg.residual_tlm(v)
    JTw = dg.residual_vjp(w)
    lhs, rhs = np.vdot(Jv, w), np.vdot(v, JTw)
    print(f"  <Jv,w> {lhs:.16e}  <v,JTw> {rhs:.16e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs))
    dg.close()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_residual_tangent_matches_fd(oracle, tmp_path, case):
    print(f"[gpu] adjoint residual FD {case['name']}", flush=True)
    dg, q = advanced(oracle, tmp_path, case)
    v, _ = directions(q, 2)
    L0, Jv = dg.residual_tlm(v)
    h = 1e-6
    set_q(dg, q + h * v)
    Lp, _ = dg.residual_tlm(v)
    set_q(dg, q - h * v)
    Lm, _ = dg.residual_tlm(v)
    fd = (Lp - Lm) / (2 * h)
    err = np.linalg.norm(fd - Jv) / np.linalg.norm(Jv)
    print(f"  |fd - Jv| / |Jv| = {err:.2e}", flush=True)
    assert err <= 1e-6
    dg.close()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_rk3_dot_product(oracle, tmp_path, case):
    print(f"[gpu] adjoint rk3 dot-product {case['name']}", flush=True)
    dg, q = advanced(oracle, tmp_path, case)
    v, w = directions(q, 3)
    n = 5
    set_q(dg, q)
    Mv = dg.rk3_tlm(n, v)
    set_q(dg, q)
    MTw = dg.rk3_adjoint(n, w, checkpoint=2)
    lhs, rhs = np.vdot(Mv, w), np.vdot(v, MTw)
    print(f"  <Mv,w> {lhs:.16e}  <v,MTw> {rhs:.16e}", flush=True)
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), abs(rhs))
    # the same adjoint again: the buffers the VJP reuses between launches
    # (2:1 mortar terms, boundary-side terms) start clean on every call; the
    # mortar terms are summed with atomics, so equal to rounding
    set_q(dg, q)
    MTw2 = dg.rk3_adjoint(n, w, checkpoint=2)
    assert np.max(np.abs(MTw2 - MTw)) <= 1e-12 * np.max(np.abs(MTw))
    dg.close()


def test_rk3_gradient_matches_fd(oracle, tmp_path):
    """J(q0) = 1/2 |theta'(q_N) - theta_obs|^2 after N steps; the adjoint
    gradient against central differences along a random direction."""
    case = [c for c in EULER_CASES if c["name"] == "cfg0_2d_n4"][0]
    print(f"[gpu] adjoint gradient FD {case['name']}", flush=True)
    dg, q = advanced(oracle, tmp_path, case, steps=1)
    pr, rr = dg.background()
    T0 = 300.0
    n = 6

    def theta(qn):
        return (qn[4] + rr * T0) / (rr + qn[0]) - T0

    rng = np.random.default_rng(7)
    obs = theta(q) + 1e-3 * rng.standard_normal(q.shape[1])

    def J(q0):
        set_q(dg, q0)
        dg.rk3_steps(n, 64)
        return 0.5 * np.sum((theta(conserved(dg)) - obs) ** 2)

    set_q(dg, q)
    dg.rk3_steps(n, 64)
    qn = conserved(dg)
    r = theta(qn) - obs
    rho = rr + qn[0]
    lam = np.zeros_like(qn)
    lam[4] = r / rho
    lam[0] = -r * (qn[4] + rr * T0) / rho ** 2
    set_q(dg, q)
    g = dg.rk3_adjoint(n, lam, checkpoint=4)
    v, _ = directions(q, 4)
    h = 1e-6
    fd = (J(q + h * v) - J(q - h * v)) / (2 * h)
    ad = np.vdot(g, v)
    print(f"  FD {fd:.12e}  adjoint {ad:.12e}", flush=True)
    assert abs(fd - ad) <= 1e-6 * abs(ad)
    dg.close()


def test_autograd_function_gradient(oracle, tmp_path):
    """The same gradient through torch.autograd (RK3Steps), with respect to the
    initial potential-temperature perturbation at constant pressure."""
    torch = pytest.importorskip("torch")
    from nebulasem_amd.autograd import RK3Steps
    case = [c for c in EULER_CASES if c["name"] == "cfg0_2d_n4"][0]
    print(f"[gpu] adjoint autograd {case['name']}", flush=True)
    mesh, om, prm, dg, (p_ref, rho_ref, rho, U, T, p) = setup(oracle, tmp_path, case)
    nb = mesh.info.nbcs_field
    pr, rr = (torch.from_numpy(x) for x in dg.background())
    P0, T0, cp, cv = 101325.0, 300.0, 1004.67, 715.5
    R, gam = cp - cv, cp / cv
    n = 4

    def q_of(th):
        r = (P0 / (R * (th + T0))) * (pr / P0) ** (1 / gam)
        z = torch.zeros_like(th)
        return torch.stack([r - rr, z, z, z, r * (th + T0) - rr * T0])

    def objective(th0, obs):
        qn = RK3Steps.apply(q_of(th0), dg, n, 2)
        thn = (qn[4] + rr * T0) / (rr + qn[0]) - T0
        return 0.5 * torch.sum((thn - obs) ** 2)

    th0 = torch.from_numpy(T[:nb].copy())
    rng = np.random.default_rng(11)
    obs = th0 + torch.from_numpy(1e-3 * rng.standard_normal(nb))
    x = th0.clone().requires_grad_(True)
    objective(x, obs).backward()
    v = torch.from_numpy(rng.standard_normal(nb))
    h = 1e-5
    with torch.no_grad():
        fd = (objective(th0 + h * v, obs) - objective(th0 - h * v, obs)) / (2 * h)
    ad = torch.dot(x.grad, v)
    print(f"  FD {float(fd):.12e}  autograd {float(ad):.12e}", flush=True)
    assert abs(float(fd) - float(ad)) <= 1e-6 * abs(float(ad))
    dg.close()
