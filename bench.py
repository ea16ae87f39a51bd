"""Benchmark of the MI355X dGSEM Euler residual and FV pressure PCG.

Workload (one GPU): 3D dGSEM compressible Euler, 64x64x32 hex elements over
10 x 10 x 5 km, N = 7 (512 GLL nodes per element, 67,108,864 nodes, 335.5M
DOF), hydrostatic theta0 = 300 K plus four seeded Gaussian warm bubbles
(dtheta in [0.1, 1] K, radii 0.5-2 km), no-flux walls, SSP-RK3.

One "step" = one SSP-RK3 step (three fused residual+update stages) over the
whole mesh, inputs resident in HBM.  value = DOF*stage/s of the
FP32-perturbation form (FP32 residual and stage increments over the FP64
background, FP64 base state and accumulation); the FP64 form, the error of
FP32 against FP64 after warmup+steps steps, the reference-mode split step and
the FV Poisson PCG time to |b-Ax|/|b| <= 1e-10 are reported beside it.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-fv]
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12               # MI355X HBM3E, bytes/s
NQ = 5


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes_per_node(s_in, s_base, s_out, s_geo, N):
    """Algorithmic bytes per node per stage of the DG residual+update, the
    figure of the path's measurement contract: n_q (s_in + s_base + s_out)
    + 10 s_geo + 2*8 + 3/(N+1) (3 s_geo + 8)."""
    return NQ * (s_in + s_base + s_out) + 10 * s_geo + 2 * 8 + 3.0 / (N + 1) * (3 * s_geo + 8)


def compulsory_bytes_per_node(s_inc, s_geo, N, affine=False):
    """Compulsory HBM bytes per node of one SSP-RK3 stage launch, averaged
    over the three stages (each array element read or written once):
    FP64 base state 5 x 8 every stage; increment in 5 s_inc on stages 2 and 3;
    output 5 s_inc (stages 1, 2) or the FP64 state 5 x 8 (stage 3);
    background p_ref, rho_ref 2 x 8; metric JcV, cV 10 s_geo per node, or one
    80-byte record per element with affine elements; per element the six
    face entries (6 x 12 B) and its three owned facet normals (3 x 24 B).
    Neighbour traces are re-reads of state counted once already."""
    npe = (N + 1) ** 3
    state = 5 * 8 + 5 * s_inc * 2 / 3 + (2 * 5 * s_inc + 5 * 8) / 3
    metric = 80.0 / npe if affine else 10 * s_geo
    return state + 2 * 8 + metric + (6 * 12 + 3 * 24) / npe


# Box corner points and the block's hex vertex order over them (x, then y,
# then z fastest for ndiv = (nx, ny, nz)), as a block description file writes
# them: points list, then `8{0 1 5 4 3 2 6 7}`.
BLOCK_ORDER = (0, 1, 5, 4, 3, 2, 6, 7)


def box_points(lx, ly, lz):
    return [(0, 0, 0), (lx, 0, 0), (lx, 0, lz), (0, 0, lz), (0, ly, 0), (lx, ly, 0), (lx, ly, lz), (0, ly, lz)]


def box_patches(lx, ly, lz):
    """Block corners in hex vertex order and the named wall patches."""
    c = box_points(lx, ly, lz)
    planes = [("top", (3, 2, 6, 7)), ("bottom", (0, 1, 5, 4)), ("front", (0, 1, 2, 3)),
              ("back", (4, 5, 6, 7)), ("sides", (1, 5, 6, 2)), ("sides", (0, 4, 7, 3))]
    return [c[i] for i in BLOCK_ORDER], [(n, [c[i] for i in g]) for n, g in planes]


def bubbles(seed, lx, ly, lz, k=4):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k):
        dth = rng.uniform(0.1, 1.0)
        r = rng.uniform(500.0, 2000.0)
        cx, cy, cz = rng.uniform(0, lx), rng.uniform(0, ly), rng.uniform(0, lz)
        out.append(f"gaussian 0 {dth!r} {cx!r} {cy!r} {cz!r} {r!r} {r!r} {r!r}")
    return out


def field_file(path, ncomp, gens, patches, kind):
    names = sorted(set(p for p in patches))
    with open(path, "w") as f:
        f.write(f"size {ncomp}\ninternal {len(gens)}\n{{\n")
        for g in gens:
            f.write("    " + g + "\n")
        f.write(f"}}\nboundary {len(names)}\n{{\n")
        for n in names:
            f.write(f"    {n} {{\n        type {kind}\n    }}\n")
        f.write("}\n")


def setup_dg(ndiv, size, order, dt, seed, tmp):
    from nebulasem_amd import DGEuler, Mesh, euler_params, read_field
    lx, ly, lz = size
    corners, patches = box_patches(lx, ly, lz)
    t = time.time()
    mesh = Mesh.box(corners, ndiv, patches=patches, order=order)
    log(f"mesh {ndiv} order {order}: {mesh.info.nbcs_field} nodes, {time.time() - t:.1f}s")
    names = [n for n, _ in patches]
    g = (0.0, 0.0, -9.80606)
    field_file(os.path.join(tmp, "T0.txt"), 1, bubbles(seed, lx, ly, lz), names, "NEUMANN")
    field_file(os.path.join(tmp, "U0.txt"), 3, ["uniform 0 0 0"], names, "SYMMETRY")
    field_file(os.path.join(tmp, "p0.txt"), 1, ["uniform 0"], names, "NEUMANN")
    field_file(os.path.join(tmp, "rho0.txt"), 1, ["uniform 0"], names, "NEUMANN")
    t = time.time()
    f = {k: read_field(mesh, os.path.join(tmp, k + "0.txt"), nc, g)
         for k, nc in (("T", 1), ("U", 3), ("p", 1), ("rho", 1))}
    nb = mesh.info.nbcs_field
    dg = DGEuler.from_fields(mesh, euler_params(dt), f)
    T = np.ascontiguousarray(f["T"][0][:nb, 0])
    p = np.ascontiguousarray(f["p"][0][:nb, 0])
    U = np.ascontiguousarray(f["U"][0][:nb])
    rho = dg.initial_density(T, p)
    del f
    log(f"fields + device set-up {time.time() - t:.1f}s")
    return mesh, dg, (rho, U, T, p)


def amr_bench(steps=20, dt=0.002):
    """3D AMR dGSEM Euler: a 32x32x16 hex block over 10x10x5 km, N=5, refined
    twice (2:1) around a 1 K bubble (centre 5000 5000 2000 m, radius
    1500 m) by the reference's prepare tool (the grid is committed,
    xz-compressed: 47,380 elements on three levels); initial temperature the
    deck's Gaussian evaluated on the refined nodes.  FP32-perturbation and
    FP64 SSP-RK3, mortar pre-pass included in every stage."""
    import lzma
    import shutil
    from nebulasem_amd import DGEuler, Mesh, euler_params, read_field
    src = os.path.join(ROOT, "tests", "golden", "amr", "amr3_32x32x16_n5_grid_0.bin.xz")
    with tempfile.TemporaryDirectory() as tmp:
        grid = os.path.join(tmp, "grid_0.bin")
        with lzma.open(src) as fi, open(grid, "wb") as fo:
            shutil.copyfileobj(fi, fo)
        t = time.time()
        mesh = Mesh.load(grid, order=(5, 5, 5))
        log(f"amr mesh: {mesh.info.nbcs_field} nodes, {time.time() - t:.1f}s")
        names = mesh.boundaries
        g = (0.0, 0.0, -9.80606)
        field_file(os.path.join(tmp, "T0.txt"), 1, ["gaussian 0 1.0 5000 5000 2000 1500 1500 1500"], names,
                   "NEUMANN")
        field_file(os.path.join(tmp, "U0.txt"), 3, ["uniform 0 0 0"], names, "SYMMETRY")
        field_file(os.path.join(tmp, "p0.txt"), 1, ["uniform 0"], names, "NEUMANN")
        field_file(os.path.join(tmp, "rho0.txt"), 1, ["uniform 0"], names, "NEUMANN")
        f = {k: read_field(mesh, os.path.join(tmp, k + "0.txt"), nc, g)
             for k, nc in (("T", 1), ("U", 3), ("p", 1), ("rho", 1))}
    nb = mesh.info.nbcs_field
    fmc = mesh.array("FMC")
    dg = DGEuler.from_fields(mesh, euler_params(dt), f)
    T = np.ascontiguousarray(f["T"][0][:nb, 0])
    p = np.ascontiguousarray(f["p"][0][:nb, 0])
    U = np.ascontiguousarray(f["U"][0][:nb])
    rho = dg.initial_density(T, p)
    del f
    out = {"workload": "3D AMR dGSEM Euler, 32x32x16 base hex over 10x10x5 km refined twice (2:1) around "
                       "a 1 K bubble, N=5, SSP-RK3 dt=%g" % dt,
           "elements": nb // 216, "nodes": nb, "dof": NQ * nb, "mortar_faces": int((fmc > 0).sum()),
           "steps": steps, "stage_kernel": "rk_stage_kernel_fx<R,S,6,6,6,W,DIN,AFF,false,true> + mortar pre-pass"}
    for prec in (32, 64):
        dg.set_state(rho, U, T, p)
        dg.rk3_steps(2, prec)
        dg.synchronize()
        dg.set_state(rho, U, T, p)
        sec = dg.time_rk3(steps, prec)
        out[f"fp{prec}"] = {"value": NQ * nb * 3 * steps / sec, "unit": "DOF*stage/s",
                            "ms_per_step": sec / steps * 1e3}
    out["fp32_over_fp64"] = out["fp32"]["value"] / out["fp64"]["value"]
    out["fp32_vs_fp64_error"] = fp32_error(dg, (rho, U, T, p), steps)
    # discrete adjoint over the same steps (FP64, checkpoints every 10):
    # J = 1/2 |Theta'(T)|^2, its gradient with respect to the initial
    # conserved state; cost relative to the forward run, a central
    # difference along a random direction, and the tangent-linear / adjoint
    # dot-product identity over the whole trajectory
    dg.set_state(rho, U, T, p)
    rp, m, tp = dg.get_conserved()
    m = m.reshape(-1, 3)
    q0 = np.stack([rp, m[:, 0], m[:, 1], m[:, 2], tp])

    def set_q(qq):
        dg.set_conserved(qq[0], np.ascontiguousarray(qq[1:4].T), qq[4])

    def run(qq):
        set_q(qq)
        dg.rk3_steps(steps, 64)
        a_, b_, c_ = dg.get_conserved()
        b_ = b_.reshape(-1, 3)
        return np.stack([a_, b_[:, 0], b_[:, 1], b_[:, 2], c_])

    set_q(q0)
    t_fwd = dg.time_rk3(steps, 64)
    qn = run(q0)
    lam = np.zeros_like(qn)
    lam[4] = qn[4]
    # every step's stage states held in HBM (1.2 GB per step at this size,
    # 25 GB for the 20 steps: no segment is recomputed), and beside it the
    # checkpoint-every-10 schedule (half the steps recomputed)
    held = {}
    for ck in (steps, 10):
        set_q(q0)
        dg.rk3_adjoint(steps, lam, ck)  # warm-up: workspace allocation
        set_q(q0)
        held[ck] = dg.time_rk3_adjoint(steps, lam, ck)
    gq, t_adj = held[steps]
    t_adj10 = held[10][1]
    rng = np.random.default_rng(20261204)
    scale = np.array([max(np.abs(q0[k]).max(), 1e-3) for k in range(5)])[:, None]
    v = rng.standard_normal(q0.shape) * scale
    h = 1e-5

    def Jof(qq):
        return 0.5 * float(np.sum(run(qq)[4] ** 2))

    fd = (Jof(q0 + h * v) - Jof(q0 - h * v)) / (2 * h)
    ad = float(np.vdot(gq, v))
    w = rng.standard_normal(q0.shape)
    set_q(q0)
    Mv = dg.rk3_tlm(steps, v)
    set_q(q0)
    MTw = dg.rk3_adjoint(steps, w, 10)
    lhs, rhs = float(np.vdot(Mv, w)), float(np.vdot(v, MTw))
    out["adjoint"] = {"objective": "1/2 |Theta'(T)|^2 after the FP64 steps, gradient w.r.t. the initial state",
                      "checkpoint": steps, "forward_s": t_fwd, "adjoint_s": t_adj,
                      "timing": "HIP events on the library's stream, host transfers excluded; the adjoint "
                                "includes its own forward sweep",
                      "adjoint_over_forward": t_adj / t_fwd,
                      "backward_over_forward": (t_adj - t_fwd) / t_fwd,
                      "checkpoint_10": {"adjoint_s": t_adj10, "adjoint_over_forward": t_adj10 / t_fwd},
                      "grad_fd": fd, "grad_adjoint": ad,
                      "grad_rel_err": abs(fd - ad) / abs(ad), "dot_tlm": lhs, "dot_adjoint": rhs,
                      "dot_rel_err": abs(lhs - rhs) / max(abs(lhs), abs(rhs))}
    log(f"amr adjoint: {t_adj / t_fwd:.2f}x forward, its backward sweep {(t_adj - t_fwd) / t_fwd:.2f}x "
        f"(states held; checkpoint every 10: {t_adj10 / t_fwd:.2f}x), "
        f"FD rel err {out['adjoint']['grad_rel_err']:.2e}, "
        f"dot-product rel err {out['adjoint']['dot_rel_err']:.2e}")
    dg.close()
    return out


def adjoint_bench(nsteps=50, checkpoint=10, dt=0.001):
    """Discrete adjoint of the 2D bubble: 64x64 elements (1 thick in y) over
    1000 x 1000 m, N = 5, Gaussian dtheta = 0.5 K (radius 250 m, centre
    x = 500, z = 350) plus seeded 1e-3 K noise; J = 1/2 |theta'(T) - theta_obs|^2
    after nsteps FP64 SSP-RK3 steps, theta_obs from a forward run with other
    noise; gradient with respect to the initial theta' (at constant pressure)
    with checkpoints every `checkpoint` steps.  Reports the adjoint's cost
    relative to the forward run, the tangent-linear / adjoint dot-product
    identity and a central-difference check of the gradient."""
    from nebulasem_amd import DGEuler, Mesh, euler_params, read_field
    lx, ly, lz = 1000.0, 100.0, 1000.0
    c = box_points(lx, ly, lz)
    patches = [("top", [c[i] for i in (3, 2, 6, 7)]), ("bottom", [c[i] for i in (0, 1, 5, 4)]),
               ("delete", [c[i] for i in (0, 1, 2, 3)]), ("delete", [c[i] for i in (4, 5, 6, 7)]),
               ("sides", [c[i] for i in (1, 5, 6, 2)]), ("sides", [c[i] for i in (0, 4, 7, 3)])]
    mesh = Mesh.box([c[i] for i in BLOCK_ORDER], (64, 1, 64), patches=patches, order=(5, 0, 5))
    names = [n for n, _ in patches if n != "delete"]
    g = (0.0, 0.0, -9.80606)
    with tempfile.TemporaryDirectory() as tmp:
        field_file(os.path.join(tmp, "T0.txt"), 1, ["gaussian 0 0.5 500 50 350 250 1000 250"], names, "NEUMANN")
        field_file(os.path.join(tmp, "U0.txt"), 3, ["uniform 0 0 0"], names, "SYMMETRY")
        field_file(os.path.join(tmp, "p0.txt"), 1, ["uniform 0"], names, "NEUMANN")
        field_file(os.path.join(tmp, "rho0.txt"), 1, ["uniform 0"], names, "NEUMANN")
        f = {k: read_field(mesh, os.path.join(tmp, k + "0.txt"), nc, g)
             for k, nc in (("T", 1), ("U", 3), ("p", 1), ("rho", 1))}
    nb = mesh.info.nbcs_field
    dg = DGEuler.from_fields(mesh, euler_params(dt), f)
    pr, rr = dg.background()
    P0, T0, cp, cv = 101325.0, 300.0, 1004.67, 715.5
    R, gam = cp - cv, cp / cv
    rng = np.random.default_rng(20261203)
    bubble = np.ascontiguousarray(f["T"][0][:nb, 0])
    th0 = bubble + 1e-3 * rng.standard_normal(nb)
    th_obs0 = bubble + 1e-3 * rng.standard_normal(nb)

    def q_of(th):                       # constant-pressure initial state
        r = (P0 / (R * (th + T0))) * (pr / P0) ** (1 / gam)
        z = np.zeros(nb)
        return np.stack([r - rr, z, z, z, r * (th + T0) - rr * T0])

    def dq_dth(th):                     # derivative of q_of, per node
        r = (P0 / (R * (th + T0))) * (pr / P0) ** (1 / gam)
        dr = -r / (th + T0)
        return dr, dr * (th + T0) + r

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
    set_q(q0)
    t_fwd = dg.time_rk3(nsteps, 64)
    qn = run(q0)
    r = theta(qn) - obs
    rho = rr + qn[0]
    lam = np.zeros_like(qn)
    lam[4] = r / rho
    lam[0] = -r * (qn[4] + rr * T0) / rho ** 2
    set_q(q0)
    dg.rk3_adjoint(nsteps, lam, checkpoint)  # warm-up: workspace allocation
    set_q(q0)
    gq, t_adj = dg.time_rk3_adjoint(nsteps, lam, checkpoint)
    dr, dtp = dq_dth(th0)
    grad = gq[0] * dr + gq[4] * dtp
    v = rng.standard_normal(nb)
    h = 1e-4
    fd = (J(th0 + h * v) - J(th0 - h * v)) / (2 * h)
    ad = float(np.dot(grad, v))
    # dot-product identity over the whole trajectory
    scale = np.array([max(np.abs(q0[k]).max(), 1e-3) for k in range(5)])[:, None]
    vq = rng.standard_normal(q0.shape) * scale
    wq = rng.standard_normal(q0.shape)
    set_q(q0)
    Mv = dg.rk3_tlm(nsteps, vq)
    set_q(q0)
    MTw = dg.rk3_adjoint(nsteps, wq, checkpoint)
    lhs, rhs = float(np.vdot(Mv, wq)), float(np.vdot(vq, MTw))
    dg.close()
    out = {"workload": f"2D bubble 64x1x64 N=5 ({nb} nodes), {nsteps} FP64 SSP-RK3 steps, dt={dt}, "
                       f"checkpoint every {checkpoint}",
           "forward_s": t_fwd, "adjoint_s": t_adj, "adjoint_over_forward": t_adj / t_fwd,
           "backward_over_forward": (t_adj - t_fwd) / t_fwd,
           "timing": "HIP events on the library's stream, host transfers excluded",
           "J": 0.5 * float(np.sum(r ** 2)), "grad_fd": fd, "grad_adjoint": ad,
           "grad_rel_err": abs(fd - ad) / abs(ad),
           "dot_tlm": lhs, "dot_adjoint": rhs, "dot_rel_err": abs(lhs - rhs) / max(abs(lhs), abs(rhs))}
    log(f"adjoint: forward {t_fwd * 1e3:.1f} ms, adjoint {t_adj * 1e3:.1f} ms "
        f"({out['adjoint_over_forward']:.2f}x; backward sweep {out['backward_over_forward']:.2f}x), "
        f"FD rel err {out['grad_rel_err']:.2e}, "
        f"dot-product rel err {out['dot_rel_err']:.2e}")
    return out


def profiled_traffic(key):
    """HBM bytes per launch of the dominant kernel, from the committed counter
    profile (calibrated FETCH_SIZE + WRITE_SIZE), or None."""
    path = os.path.join(ROOT, "profiles", "rk_stage_64x64x32_n7_hbm_bytes.md")
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_max(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def fp32_error(dg, state, nsteps=100):
    """The FP32-perturbation form against FP64 after nsteps SSP-RK3 steps
    from the same initial state (the FP64 device path being the one checked
    against the oracle's FP64 composition to 1e-12).  Per field, relative L2
    and max-norm, each relative to the field itself: the reference's fields
    rho' = rho - rho_ref, U, theta', p' and the conserved state rho', m,
    Theta' = rho theta - rho_ref T0; Theta' and p' also relative to the
    background Theta_ref = rho_ref T0 and p_ref they perturb.  Theta' and p'
    are small differences (at constant pressure rho' T0 and rho theta'
    cancel), so their error relative to themselves is the largest."""
    rho, U, T, p = state
    out = {}
    for prec in (64, 32):
        dg.set_state(rho, U, T, p)
        dg.rk3_steps(nsteps, prec)
        out[prec] = (dg.get_conserved(), dg.get_state())
    p_ref, rho_ref = dg.background()
    (q64, s64), (q32, s32) = out[64], out[32]
    fields = {"rho_p": (s32[0] - rho_ref, s64[0] - rho_ref), "U": (s32[1], s64[1]),
              "theta_p": (s32[2], s64[2]), "p_p": (s32[3], s64[3]),
              "m": (q32[1], q64[1]), "Theta_p": (q32[2], q64[2])}
    err = {k: {"rel_l2": rel_l2(a, b), "rel_max": rel_max(a, b)} for k, (a, b) in fields.items()}
    err["Theta_p"]["rel_l2_of_Theta_ref"] = float(np.linalg.norm(q32[2] - q64[2]) / np.linalg.norm(rho_ref * 300.0))
    err["p_p"]["rel_l2_of_p_ref"] = float(np.linalg.norm(s32[3] - s64[3]) / np.linalg.norm(p_ref))
    err["steps"] = nsteps
    err["prognostic_max_rel_l2"] = max(err[k]["rel_l2"] for k in ("rho_p", "U", "theta_p"))
    return err


def host_cpu():
    """The host's CPU model and the threads this process may use (the
    scheduler affinity, or OMP_NUM_THREADS where the environment caps it)."""
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    n = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        n = min(n, int(os.environ["OMP_NUM_THREADS"]))
    return model, max(n, 1)


def _ref_deck(d, ndiv, steps):
    from _cases import PATCHES_3D, _spec, euler_controls, field_file as ref_field_file
    lx, ly, lz = 10000.0 * ndiv[0] / 64, 10000.0 * ndiv[1] / 64, 5000.0 * ndiv[2] / 32
    case = dict(ndiv=ndiv, size=(lx, ly, lz), order=(7, 7, 7), patches=PATCHES_3D,
                dt=0.005, check_steps=[steps], T=[f"gaussian 0 0.5 {lx / 2} {ly / 2} {lz / 2} 400 400 400"])
    names = sorted(n for n, _ in PATCHES_3D)
    with open(os.path.join(d, "spec"), "w") as f:
        f.write(_spec(case))
    with open(os.path.join(d, "controls"), "w") as f:
        f.write(euler_controls(case).replace("write_interval               1\n",
                                             "write_interval               100000\n"))
    for fn, nc, gens, kind in (("T0.txt", 1, case["T"], "NEUMANN"), ("U0.txt", 3, ["uniform 0 0 0"], "SYMMETRY"),
                               ("p0.txt", 1, ["uniform 0"], "NEUMANN"), ("rho0.txt", 1, ["uniform 0"], "NEUMANN")):
        with open(os.path.join(d, fn), "w") as f:
            f.write(ref_field_file(nc, gens, [(n, kind) for n in names]))


def cpu_reference(threads, ndiv, steps):
    """The reference's own CPU path (apps/euler, its OpenMP build compiled
    from the reference sources into oracle/_ref by oracle/ref.mk) on a
    bounded sample of the workload: ndiv elements at N = 7 (the workload's
    element shape and order), split forward-Euler steps (1 stage each),
    `threads` OpenMP threads.  Step time = spacing of the per-step log
    lines (set-up excluded), median over the steps."""
    import subprocess
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    exe = os.path.join(ROOT, "oracle", "_ref", "euler_omp")
    mesher = os.path.join(ROOT, "oracle", "_ref", "mesh")
    if not (os.path.exists(exe) and os.path.exists(mesher)):
        return None
    with tempfile.TemporaryDirectory() as d:
        _ref_deck(d, ndiv, steps)
        subprocess.run([mesher, "spec", "-o", "grid_0.bin"], cwd=d, check=True, capture_output=True)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        t = time.time()
        r = subprocess.run([exe, "controls"], cwd=d, env=env, capture_output=True, text=True, timeout=300)
        wall = time.time() - t
        if r.returncode != 0:
            log(f"reference run failed: {r.stderr[-300:]}")
            return None
        stamps = [int(ln.split()[0]) for ln in r.stdout.splitlines() if " Time " in ln]
    dts = np.diff(stamps) * 1e-3
    if len(dts) < 1:
        return None
    step = float(np.median(dts))
    nodes = ndiv[0] * ndiv[1] * ndiv[2] * 512
    return {"value": NQ * nodes / step, "unit": "DOF*stage/s", "cores": threads, "kind": "reference",
            "sample": f"reference apps/euler (OpenMP build, {threads} threads), split forward-Euler step "
                      f"(1 stage/step), {ndiv[0]}x{ndiv[1]}x{ndiv[2]} elements N=7 = {nodes} nodes, median of "
                      f"{len(dts)} steps {step * 1e3:.0f} ms/step ({wall:.1f}s run incl. setup)"}


def cpu_reference_pcg(threads, n=128):
    """The reference's own Poisson solve (apps/potential, OpenMP build,
    SYMM-FULL-DIAG-PCG to its criterion 1e-10) on an n^3 unit cube with the
    benchmark's boundary conditions and seeded divergent velocity; time from
    its 'Starting iterations' log line to the solver's line (divergence,
    assembly and PCG), per cell-iteration."""
    import subprocess
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _cases import write_potential_case
    exe = os.path.join(ROOT, "oracle", "_ref", "potential_omp")
    mesher = os.path.join(ROOT, "oracle", "_ref", "mesh")
    if not (os.path.exists(exe) and os.path.exists(mesher)):
        return None
    with tempfile.TemporaryDirectory() as d:
        write_potential_case(d, dict(name=f"pot_{n}", ndiv=(n, n, n)))
        subprocess.run([mesher, "spec", "-o", "grid_0.bin"], cwd=d, check=True, capture_output=True)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        r = subprocess.run([exe, "controls"], cwd=d, env=env, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            log(f"reference potential failed: {r.stderr[-300:]}")
            return None
        t0 = t1 = its = None
        for ln in r.stdout.splitlines():
            if "Starting iterations" in ln:
                t0 = int(ln.split()[0])
            if "PCG" in ln and "Iterations" in ln:
                t1 = int(ln.split()[0])
                its = int(ln.split("Iterations")[1].split()[0])
    if t0 is None or t1 is None:
        return None
    sec = (t1 - t0) * 1e-3
    return {"seconds": sec, "iterations": its, "cells": n ** 3, "cores": threads, "kind": "reference",
            "cell_iterations_per_s": n ** 3 * its / sec,
            "sample": f"reference apps/potential (OpenMP build, {threads} threads), {n}^3 cells, DIAG-PCG to "
                      f"its criterion 1e-10: {its} iterations in {sec:.2f}s (divergence + assembly + PCG)"}


def cpu_baseline(seconds_budget=20.0):
    """The oracle (single-threaded C restatement of the reference's split
    step) on a bounded sample: 16x16x8 elements at N = 7 (1,048,576 nodes,
    the workload's element shape and order), timed over whole steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as oracle
    from nebulasem_amd import Mesh, read_field
    with tempfile.TemporaryDirectory() as tmp:
        lx, ly, lz = 10000.0 / 4, 10000.0 / 4, 5000.0 / 4
        corners, patches = box_patches(lx, ly, lz)
        mesh = Mesh.box(corners, (16, 16, 8), patches=patches, order=(7, 7, 7))
        names = [n for n, _ in patches]
        g = (0.0, 0.0, -9.80606)
        field_file(os.path.join(tmp, "T0.txt"), 1, bubbles(7, lx, ly, lz), names, "NEUMANN")
        field_file(os.path.join(tmp, "U0.txt"), 3, ["uniform 0 0 0"], names, "SYMMETRY")
        field_file(os.path.join(tmp, "p0.txt"), 1, ["uniform 0"], names, "NEUMANN")
        field_file(os.path.join(tmp, "rho0.txt"), 1, ["uniform 0"], names, "NEUMANN")
        f = {k: read_field(mesh, os.path.join(tmp, k + "0.txt"), nc, g)
             for k, nc in (("T", 1), ("U", 3), ("p", 1), ("rho", 1))}
        om = oracle.OracleMesh(mesh)
        prm = oracle.make_params(0.005, {k: (f[k][1], f[k][2].reshape(-1)) for k in ("rho", "U", "T", "p")})
        T = f["T"][0].reshape(-1).copy()
        p = f["p"][0].reshape(-1).copy()
        U = f["U"][0].reshape(-1).copy()
        p_ref, rho_ref, rho = oracle.euler_hydrostatic(om, prm, T, p)
        nb = mesh.info.nbcs_field
        steps, t0 = 0, time.time()
        while True:
            oracle.euler_split_steps(om, prm, p_ref, rho_ref, rho, U, T, p, 1)
            steps += 1
            el = time.time() - t0
            if el > seconds_budget or steps >= 20:
                break
        rate = NQ * nb * steps / el
        return {"value": rate, "unit": "DOF*stage/s", "cores": 1, "kind": "port",
                "sample": f"oracle split step (reference algorithm, forward Euler, 1 stage/step), "
                          f"16x16x8 elements N=7 = {nb} nodes, {steps} steps in {el:.1f}s, 1 thread"}


def fv_bench(n, tol=1e-10):
    from nebulasem_amd import BC_DIRICHLET, BC_NEUMANN, FVPoisson, Mesh
    corners = box_points(1.0, 1.0, 1.0)
    planes = [("xmin", (0, 4, 7, 3)), ("xmax", (1, 5, 6, 2)), ("ymin", (0, 1, 2, 3)),
              ("ymax", (4, 5, 6, 7)), ("zmin", (0, 1, 5, 4)), ("zmax", (3, 2, 6, 7))]
    t = time.time()
    mesh = Mesh.box([corners[i] for i in BLOCK_ORDER], (n, n, n),
                    patches=[(nm, [corners[i] for i in g]) for nm, g in planes])
    kinds = np.array([BC_DIRICHLET if b == "xmin" else BC_NEUMANN for b in mesh.boundaries], dtype=np.int32)
    fv = FVPoisson(mesh, kinds, np.zeros(len(kinds)))
    rng = np.random.default_rng(20261202)
    U = rng.uniform(-1.0, 1.0, size=(mesh.info.nbcs, 3))
    fv.set_rhs_divergence(U, np.zeros(len(kinds), dtype=np.int32), np.zeros((len(kinds), 3)))
    log(f"FV {n}^3 set-up {time.time() - t:.1f}s (structured={fv.structured}, uniform couplings={fv.uniform})")
    out = {}
    for mode, name in ((1, "fp64"), (2, "fp32_refined")):
        fv.solve(mode=mode, tol=tol)  # warm-up (kernel load)
        x, res, rc = fv.solve(mode=mode, tol=tol)
        out[name] = {"seconds": res.seconds, "iterations": res.iterations,
                     "outer": res.outer_iterations, "true_rel_residual": res.res_true}
        log(f"FV {name}: {res.seconds * 1e3:.1f} ms, {res.iterations} it, |b-Ax|/|b| {res.res_true:.2e}")
    # algorithmic bytes per cell per iteration of the measurement contract
    # (stored 7-coefficient operator): SpMV+dot 72 + update 56 + p update 32
    # = 160 B in FP64, 80 B for the FP32 inner iterations
    for name, b_iter in (("fp64", 160.0), ("fp32_refined", 80.0)):
        t_iter = out[name]["seconds"] / max(out[name]["iterations"], 1)
        out[name]["GBps_per_iter"] = b_iter * n ** 3 / t_iter / 1e9
        out[name]["frac_hbm_peak"] = out[name]["GBps_per_iter"] * 1e9 / HBM_PEAK
    # the bytes the kernels move per cell per iteration when the couplings
    # are uniform (taken as constants, so the SpMV reads only p and the
    # diagonal): SpMV 16 read + 8 written, x/r/z update 40 + 24, p update
    # 16 + 8 = 112 B in FP64, 56 B for the FP32 inner iterations (the
    # solver's counter profile measures them to 0.5%); the FP32 figure's
    # time also holds the FP64 refinement steps
    if fv.uniform:
        for name, b_moved in (("fp64", 112.0), ("fp32_refined", 56.0)):
            t_iter = out[name]["seconds"] / max(out[name]["iterations"], 1)
            out[name]["moved_GBps_per_iter"] = b_moved * n ** 3 / t_iter / 1e9
            out[name]["moved_frac_hbm_peak"] = out[name]["moved_GBps_per_iter"] * 1e9 / HBM_PEAK
    out["operator"] = ("structured, uniform couplings" if fv.uniform else
                       "structured" if fv.structured else "general")
    out["cells"] = n ** 3
    fv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ndiv", type=str, default="64,64,32")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fv", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-amr", action="store_true")
    # for profiling one leg at a time: skip the 100-step FP32-vs-FP64 error
    # runs and the reference-order split step
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-split", action="store_true")
    ap.add_argument("--fv-n", type=int, default=256)
    ap.add_argument("--copy-only", type=int, default=0, metavar="MIB",
                    help="only time the plain copy kernel over MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    from nebulasem_amd import stream_copy
    if args.copy_only:
        nbytes = args.copy_only << 20
        t = stream_copy(nbytes, max(args.steps, 1))
        print(json.dumps({"kernel": "stream_copy_kernel", "bytes_read": nbytes, "bytes_written": nbytes,
                          "seconds_per_copy": t, "GBps": 2 * nbytes / t / 1e9}))
        return
    t_copy = stream_copy(1 << 30, 10)
    stream_bw = 2 * (1 << 30) / t_copy
    log(f"stream copy 1 GiB: {stream_bw / 1e9:.0f} GB/s (read + write)")
    ndiv = tuple(int(x) for x in args.ndiv.split(","))
    N = 7
    dt = 0.005
    with tempfile.TemporaryDirectory() as tmp:
        mesh, dg, (rho, U, T, p) = setup_dg(ndiv, (10000.0, 10000.0, 5000.0), (N, N, N), dt, 20261202, tmp)
    nb = mesh.info.nbcs_field
    K, W = args.steps, args.warmup
    res = {}
    xfer = {}
    for prec in (64, 32):
        t0 = time.time()
        dg.set_state(rho, U, T, p)
        xfer.setdefault("set_state_s", time.time() - t0)
        dg.rk3_steps(W, prec)
        dg.synchronize()
        t = dg.time_rk3(K, prec)
        t0 = time.time()
        q = dg.get_conserved()
        xfer.setdefault("get_conserved_s", time.time() - t0)
        res[prec] = {"seconds": t, "rate": NQ * nb * 3 * K / t, "stage_s": t / (3 * K), "q": q}
        log(f"rk3 fp{prec}: {t * 1e3:.1f} ms for {K} steps, {res[prec]['rate'] / 1e9:.2f} G DOF*stage/s")
    err = None
    if not args.no_error:
        err = fp32_error(dg, (rho, U, T, p), 100)
        log("fp32 vs fp64 after 100 steps (rel L2): " +
            ", ".join(f"{k} {v['rel_l2']:.2e}" for k, v in err.items() if isinstance(v, dict)))
    # the reference's split forward-Euler step (1 stage/step)
    affine = dg.affine()
    ks = max(5, K // 5)
    ts = split_rate = None
    if not args.no_split:
        dg.set_state(rho, U, T, p)
        ts = dg.time_split(ks)
        split_rate = NQ * nb * ks / ts
        log(f"split step fp64: {ts / ks * 1e3:.2f} ms/step, {split_rate / 1e9:.2f} G DOF*stage/s")
    dg.close()
    del res[64]["q"], res[32]["q"]

    b32 = algorithmic_bytes_per_node(4, 8, 8, 4, N)
    b64 = algorithmic_bytes_per_node(8, 8, 8, 8, N)
    c32 = compulsory_bytes_per_node(4, 4, N, affine)
    c64 = compulsory_bytes_per_node(8, 8, N, affine)
    achieved = b32 * nb / res[32]["stage_s"]
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK,
                "traffic": profiled_traffic("traffic_bytes_per_launch_fp32") if ndiv == (64, 64, 32) else None,
                "traffic_source": "profiles/rk_stage_64x64x32_n7_hbm_bytes.md (mean of the 3 stages)",
                "kernel": "rk_stage_kernel_fx<float,float,8,8,8,6,DIN,%s,false>" % ("true" if affine else "false"),
                "bytes_per_node": b32, "nodes_per_launch": nb,
                # the exact mean over the three stages of what must move
                # (per-element metric on affine meshes), beside the contract
                # figure above
                "compulsory_bytes_per_node": c32,
                "compulsory_GBps": c32 * nb / res[32]["stage_s"] / 1e9,
                "compulsory_frac": c32 * nb / res[32]["stage_s"] / HBM_PEAK,
                "fp64_compulsory_bytes_per_node": c64,
                "metric_layout": "per element (affine)" if affine else "per node",
                "fp64_kernel_GBps": b64 * nb / res[64]["stage_s"] / 1e9,
                "fp64_frac": b64 * nb / res[64]["stage_s"] / HBM_PEAK,
                "stream_copy_GBps": stream_bw / 1e9,
                "frac_of_stream_copy": achieved / stream_bw}
    adj = None if args.no_adjoint else adjoint_bench()
    amr = None if args.no_amr else amr_bench()
    fv = None
    if not args.no_fv:
        fv = fv_bench(args.fv_n)
    cpu = None
    if not args.no_cpu:
        model, ncores = host_cpu()
        port = cpu_baseline(seconds_budget=10.0)
        cpu = cpu_reference(ncores, (16, 16, 8), 6) or port
        if cpu is not port:
            one = cpu_reference(1, (8, 8, 4), 4)
            if one:
                cpu["reference_single_thread"] = one
            cpu["port_single_thread"] = port
            pcg = cpu_reference_pcg(ncores)
            if pcg:
                cpu["pcg"] = pcg
                if fv:
                    cpu["pcg"]["gpu_fp64_cell_iterations_per_s"] = (
                        fv["cells"] * fv["fp64"]["iterations"] / fv["fp64"]["seconds"])
        cpu["cpu_model"] = model
        cpu["host_threads_available"] = ncores
        log(f"cpu baseline: {cpu['value'] / 1e6:.1f} M DOF*stage/s on {cpu['cores']} threads ({model})")
    line = {
        "metric": "dGSEM DOF*stage/s (FP32-perturbation SSP-RK3, N=7)",
        "value": res[32]["rate"],
        "unit": "DOF*stage/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": res[32]["seconds"] / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 residual / f64 state",
        "data": "synthetic: hydrostatic theta0=300K + 4 seeded Gaussian bubbles",
        "config": {"workload": f"3D dGSEM Euler {ndiv[0]}x{ndiv[1]}x{ndiv[2]} hex, 10x10x5 km, N=7, "
                               f"SSP-RK3 dt={dt}", "nodes": nb, "dof": NQ * nb},
        "fp64": {"value": res[64]["rate"], "ms_per_step": res[64]["seconds"] / K * 1e3},
        "fp32_over_fp64": res[32]["rate"] / res[64]["rate"],
        "fp32_vs_fp64_error": err,
        "split_reference_step_fp64": ({"value": split_rate, "ms_per_step": ts / ks * 1e3}
                                      if ts else None),
        # host <-> device state transfers through the C-ABI (host arrays,
        # pageable memory, layout conversion included), once per run
        "host_transfer": dict(xfer, state_bytes=6 * 8 * nb,
                              fp32_rate_incl_one_round_trip=NQ * nb * 3 * K / (
                                  res[32]["seconds"] + xfer["set_state_s"] + xfer["get_conserved_s"])),
        "roofline": roofline,
        "pcg": fv,
        "adjoint": adj,
        "amr": amr,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
