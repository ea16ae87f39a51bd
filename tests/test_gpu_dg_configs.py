"""The benchmark configurations' DG cases as stated, and the error paths of
the DG entry points.

cfg0 (2D bubble, 16x1x16 elements, N = 4) with SSP-RK3 at dt = 0.02 s: the
oracle's composition of the reference's operators goes non-finite within
100 steps (as the reference's own forward Euler does at 0.02 and 0.01), so
the device must report NSEM_ENONFINITE, not return NaN fields.  At
dt = 0.01 s (the largest step of the form 0.02 / 2^k that the oracle runs
for 100 steps) the 100 FP64 steps are compared with the oracle.

cfg3 (adjoint of the 2D bubble, 64x1x64 elements, N = 5, 50 SSP-RK3 steps,
checkpoints every 10): the gradient of J = 1/2 |theta'(T) - theta_obs|^2
with respect to the initial theta' against central differences, and the
tangent-linear / adjoint dot-product identity over the 50 steps."""
import os

import numpy as np
import pytest

from _cases import CFG0
from _euler_setup import GRAVITY
from nebulasem_amd import DGEuler, Mesh, NsemError, euler_params, read_field
from test_gpu_dg_rk3 import check_rk3_against_oracle
from test_gpu_dg_split import relerr, setup

pytestmark = pytest.mark.gpu

ENONFINITE = 5


def test_cfg0_rk3_at_stated_dt_reports_nonfinite(oracle, tmp_path):
    print("[gpu] cfg0 ssprk3 dt=0.02 x100", flush=True)
    case = dict(CFG0, dt=0.02)
    mesh, om, prm, dg, (p_ref, rho_ref, rho, U, T, p) = setup(oracle, tmp_path, case)
    assert oracle.euler_ssprk3_steps(om, prm, p_ref, rho_ref, rho, U, T, p, 100) == ENONFINITE
    with pytest.raises(NsemError) as ei:
        dg.rk3_steps(100, 64)
    assert ei.value.code == ENONFINITE
    # the flag is cleared: a fresh state runs again
    mesh2, om2, prm2, dg2, st2 = setup(oracle, tmp_path / "b", dict(CFG0, dt=0.005))
    dg2.rk3_steps(2, 64)
    dg.close()
    dg2.close()


def test_cfg0_rk3_100_steps_matches_oracle(oracle, tmp_path):
    print("[gpu] cfg0 ssprk3 dt=0.01 x100", flush=True)
    case = dict(CFG0, dt=0.01)
    mesh, om, prm, dg, (p_ref, rho_ref, rho, U, T, p) = setup(oracle, tmp_path, case)
    nb = mesh.info.nbcs_field
    check_rk3_against_oracle(oracle, om, prm, dg, nb, p_ref, rho_ref, rho, U, T, p, 100)
    dg.close()


def test_split_step_reports_nonfinite(oracle, tmp_path):
    """The reference's forward Euler at dt = 0.02 on cfg0 goes NaN (it runs
    on and writes NaN fields); the device split step returns
    NSEM_ENONFINITE."""
    print("[gpu] cfg0 split dt=0.02 x100", flush=True)
    case = dict(CFG0, dt=0.02)
    mesh, om, prm, dg, (p_ref, rho_ref, rho, U, T, p) = setup(oracle, tmp_path, case)
    assert oracle.euler_split_steps(om, prm, p_ref, rho_ref, rho, U, T, p, 100) == ENONFINITE
    with pytest.raises(NsemError) as ei:
        dg.split_steps(100)
    assert ei.value.code == ENONFINITE
    dg.close()


def test_split_and_rk3_steps_alternate(oracle, tmp_path):
    """Split steps, then SSP-RK3 steps, then split steps again: each form is
    rebuilt on the device from the one advanced last, as the oracle does by
    carrying one primitive state through both (its split steps with the
    correctly rounded pow the device uses, its SSP-RK3 in the device's
    cancellation-free forms)."""
    print("[gpu] split / rk3 alternation", flush=True)
    case = dict(CFG0, dt=0.005)
    mesh, om, prm, dg, (p_ref, rho_ref, rho, U, T, p) = setup(oracle, tmp_path, case)
    nb = mesh.info.nbcs_field
    prm.mol_exact = 1
    oracle.euler_split_steps(om, prm, p_ref, rho_ref, rho, U, T, p, 3, cr_pow=True)
    oracle.euler_ssprk3_steps(om, prm, p_ref, rho_ref, rho, U, T, p, 3)
    oracle.euler_split_steps(om, prm, p_ref, rho_ref, rho, U, T, p, 3, cr_pow=True)
    dg.split_steps(3)
    dg.rk3_steps(3, 64)
    q_mid = dg.get_conserved()
    dg.split_steps(3)
    g_rho, g_U, g_T, g_p = dg.get_state()
    # rho, theta' to 1e-12; U and p' follow Theta' (see check_rk3_against_oracle):
    # U to 1e-10 of itself, p' to 1e-13 of the full pressure
    for name, a, b, tol in (("rho", g_rho, rho[:nb], 1e-12), ("U", g_U, U[:3 * nb], 1e-10),
                            ("T", g_T, T[:nb], 1e-12)):
        l2, linf = relerr(a, b)
        print(f"  {name}: rel L2 {l2:.2e} rel max {linf:.2e}", flush=True)
        assert l2 <= tol, (name, l2)
    assert np.linalg.norm(g_p - p[:nb]) / np.linalg.norm(p_ref[:nb]) <= 1e-13
    # the conserved state read between the phases is the RK state
    assert np.abs(q_mid[1]).max() > 0
    # diagnostics follow the current (split) state
    diag = dg.diagnostics()
    od = oracle.euler_diagnostics(om, prm, p_ref, rho, U, T, p, cr_pow=True)
    assert abs(diag[3] - od[3]) <= 1e-12 * abs(od[3])
    dg.close()


def test_set_conserved_clears_flag(oracle, tmp_path):
    """A non-finite run leaves no stale flag behind: after the failure,
    setting a finite conserved state runs cleanly."""
    case = dict(CFG0, dt=0.02)
    mesh, om, prm, dg, state = setup(oracle, tmp_path, case)
    q = dg.get_conserved()
    with pytest.raises(NsemError):
        dg.rk3_steps(100, 64)
    dg.set_conserved(*q)
    dg.rk3_steps(1, 64)
    dg.close()


# --------------------------------------------------------------- cfg3 ----
def _cfg3_problem(tmp_path, dt):
    """64x1x64 elements over 1000 x 100 x 1000 m, N = 5 in x and z, the
    cfg0 bubble (gaussian 0.5 K, radius 250 m at x = 500, z = 350)."""
    case = dict(CFG0, name="cfg3_2d_n5", ndiv=(64, 1, 64), order=(5, 0, 5), dt=dt)
    from _cases import write_euler_case
    d = os.path.join(str(tmp_path), case["name"])
    write_euler_case(d, case)
    mesh = Mesh.from_spec(os.path.join(d, "spec"), order=case["order"])
    f = {k: read_field(mesh, os.path.join(d, k + "0.txt"), nc, GRAVITY)
         for k, nc in (("T", 1), ("U", 3), ("p", 1), ("rho", 1))}
    return mesh, f


def test_cfg3_adjoint_gradient_and_dot_product(tmp_path):
    print("[gpu] cfg3 adjoint 64x1x64 N=5, 50 steps", flush=True)
    nsteps, checkpoint, dt = 50, 10, 0.001
    mesh, f = _cfg3_problem(tmp_path, dt)
    nb = mesh.info.nbcs_field
    assert nb == 64 * 64 * 36
    dg = DGEuler.from_fields(mesh, euler_params(dt), f)
    pr, rr = dg.background()
    P0, T0, cp, cv = 101325.0, 300.0, 1004.67, 715.5
    R, gam = cp - cv, cp / cv
    rng = np.random.default_rng(20261203)
    bubble = np.ascontiguousarray(f["T"][0][:nb, 0])
    th0 = bubble + 1e-3 * rng.standard_normal(nb)
    th_obs0 = bubble + 1e-3 * rng.standard_normal(nb)

    def q_of(th):  # the state at constant (hydrostatic) pressure
        r = (P0 / (R * (th + T0))) * (pr / P0) ** (1 / gam)
        z = np.zeros(nb)
        return np.stack([r - rr, z, z, z, r * (th + T0) - rr * T0])

    def set_q(q):
        dg.set_conserved(q[0], np.ascontiguousarray(q[1:4].T), q[4])

    def run(q):
        set_q(q)
        dg.rk3_steps(nsteps, 64)
        rp, m, tp = dg.get_conserved()
        m = m.reshape(-1, 3)
        return np.stack([rp, m[:, 0], m[:, 1], m[:, 2], tp])

    def theta(q):
        return (q[4] + rr * T0) / (rr + q[0]) - T0

    obs = theta(run(q_of(th_obs0)))

    def J(th):
        return 0.5 * float(np.sum((theta(run(q_of(th))) - obs) ** 2))

    q0 = q_of(th0)
    qn = run(q0)
    r = theta(qn) - obs
    rho = rr + qn[0]
    lam = np.zeros_like(qn)
    lam[4] = r / rho
    lam[0] = -r * (qn[4] + rr * T0) / rho ** 2
    set_q(q0)
    gq = dg.rk3_adjoint(nsteps, lam, checkpoint)
    rr0 = (P0 / (R * (th0 + T0))) * (pr / P0) ** (1 / gam)
    dr = -rr0 / (th0 + T0)
    grad = gq[0] * dr + gq[4] * (dr * (th0 + T0) + rr0)
    v = rng.standard_normal(nb)
    h = 1e-4
    fd = (J(th0 + h * v) - J(th0 - h * v)) / (2 * h)
    ad = float(np.dot(grad, v))
    fd_err = abs(fd - ad) / abs(ad)
    scale = np.array([max(np.abs(q0[k]).max(), 1e-3) for k in range(5)])[:, None]
    vq = rng.standard_normal(q0.shape) * scale
    wq = rng.standard_normal(q0.shape)
    set_q(q0)
    Mv = dg.rk3_tlm(nsteps, vq)
    set_q(q0)
    MTw = dg.rk3_adjoint(nsteps, wq, checkpoint)
    lhs, rhs = float(np.vdot(Mv, wq)), float(np.vdot(vq, MTw))
    dot_err = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"  J {0.5 * float(np.sum(r ** 2)):.6e}  FD {fd:.10e} adjoint {ad:.10e} rel {fd_err:.2e}; "
          f"dot-product rel {dot_err:.2e}", flush=True)
    assert fd_err <= 1e-6
    assert dot_err <= 1e-10
    dg.close()
