"""Test cases shared by the golden-fixture generator and the tests: each case
is an input deck in the reference's formats (block spec for the mesh tool,
controls file, field files)."""
import os

import numpy as np

BOX_CORNERS = lambda lx, ly, lz: [(0, 0, 0), (lx, 0, 0), (lx, 0, lz), (0, 0, lz),  # noqa: E731
                                  (0, ly, 0), (lx, ly, 0), (lx, ly, lz), (0, ly, lz)]


def _spec(case):
    lx, ly, lz = case["size"]
    nx, ny, nz = case["ndiv"]
    pts = "\n".join("%r %r %r" % tuple(c) for c in case.get("corners", BOX_CORNERS(lx, ly, lz)))
    spacing = case.get("spacing", "linear 3{%d %d %d}" % (nx, ny, nz))
    lines = ["8", "{", pts, "}", "8{0 1 5 4 3 2 6 7} " + spacing]
    for name, groups in case["patches"]:
        lines.append(name + " " + " ".join("4{%s}" % " ".join(map(str, g)) for g in groups))
    return "\n".join(lines) + "\n"


PATCHES_2D = [("top", [(3, 2, 6, 7)]), ("bottom", [(0, 1, 5, 4)]),
              ("delete", [(0, 1, 2, 3), (4, 5, 6, 7)]), ("sides", [(1, 5, 6, 2), (0, 4, 7, 3)])]
PATCHES_3D = [("top", [(3, 2, 6, 7)]), ("bottom", [(0, 1, 5, 4)]), ("front", [(0, 1, 2, 3)]),
              ("back", [(4, 5, 6, 7)]), ("sides", [(1, 5, 6, 2), (0, 4, 7, 3)])]
PATCHES_BOX = [("xmin", [(0, 4, 7, 3)]), ("xmax", [(1, 5, 6, 2)]), ("ymin", [(0, 1, 2, 3)]),
               ("ymax", [(4, 5, 6, 7)]), ("zmin", [(0, 1, 5, 4)]), ("zmax", [(3, 2, 6, 7)])]


def patch_names(case):
    return sorted(n for n, _ in case["patches"] if n != "delete")


def euler_controls(case):
    o = case["order"]
    return """general
{
    solver                       euler
    mesh                         grid
    state                        TRANSIENT
    start_step                   0
    end_step                     %d
    write_interval               1
    dt                           %r
    convection_scheme            RUSANOV
    time_scheme                  BDF1
    method                       PCG
    preconditioner               DIAG
    gravity                      0 0 -9.80606
    npx                          %d
    npy                          %d
    npz                          %d
    viscosity                    %r
    write_format                 %s
}
euler
{
    diffusion                    %s
    buoyancy                     YES
}
""" % (max(case.get("check_steps", [1])), case.get("dt", 0.005), o[0], o[1], o[2],
       case.get("viscosity", 1.568e-5), case.get("write_format", "BINARY"),
       "NO" if case.get("viscosity") is None else "YES") + refinement_controls(case)


def refinement_controls(case):
    """The prepare tool's refinement block (field.cpp:474-482) for AMR decks."""
    r = case.get("refine")
    if not r:
        return ""
    return """prepare
{
    fields    4 { U T p rho }
}
refinement
{
    direction       %d %d %d
    field           %s
    field_min       %r
    field_max       %r
    max_level       %d
    limit           100000
}
""" % (tuple(r["direction"]) + (r["field"], r["field_min"], r["field_max"], r["max_level"]))


def field_file(ncomp, gens, bcs):
    """bcs: (patch, type) or (patch, type, [extra lines such as "value 1"])."""
    body = ["size %d" % ncomp, "internal %d" % len(gens), "{"] + ["    " + g for g in gens] + ["}"]
    body += ["boundary %d" % len(bcs), "{"]
    for bc in bcs:
        body += ["    %s {" % bc[0], "        type %s" % bc[1]] + ["        " + x for x in (bc[2] if len(bc) > 2 else [])]
        body += ["    }"]
    body += ["}"]
    return "\n".join(body) + "\n"


def write_euler_case(d, case):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "spec"), "w") as f:
        f.write(_spec(case))
    with open(os.path.join(d, "controls"), "w") as f:
        f.write(euler_controls(case))
    names = patch_names(case)
    fields = {"T0.txt": (1, case.get("T", ["uniform 0"]), "NEUMANN"),
              "U0.txt": (3, ["uniform 0 0 0"], "SYMMETRY"),
              "p0.txt": (1, ["uniform 0"], "NEUMANN"),
              "rho0.txt": (1, ["uniform 0"], "NEUMANN")}
    for fn, (nc, gens, kind) in fields.items():
        with open(os.path.join(d, fn), "w") as f:
            f.write(field_file(nc, gens, [(n, kind) for n in names]))


def write_potential_case(d, case):
    os.makedirs(d, exist_ok=True)
    n = case["ndiv"]
    c = dict(case, size=(1, 1, 1), patches=PATCHES_BOX, order=(0, 0, 0))
    with open(os.path.join(d, "spec"), "w") as f:
        f.write(_spec(c))
    with open(os.path.join(d, "controls"), "w") as f:
        f.write("""general
{
    solver                       potential
    mesh                         grid
    state                        STEADY
    start_step                   0
    end_step                     1
    write_interval               1
    convection_scheme            HYBRID
    nonortho_scheme              OVER_RELAXED
    method                       PCG
    preconditioner               DIAG
    tolerance                    %r
    max_iterations               100000
    gravity                      0 0 0
    npx                          0
    npy                          0
    npz                          0
}
potential
{
    n_ORTHO                      %d
}
""" % (case.get("tol", 1e-10), case.get("n_ortho", 0)))
    U = potential_velocity(case)
    names = patch_names(c)
    with open(os.path.join(d, "U0.txt"), "w") as f:
        f.write("size 3\ninternal %d\n{\n" % len(U))
        for u in U:
            f.write("%.17g %.17g %.17g\n" % tuple(u))
        f.write("}\n" + field_file(3, [], [(nm, "NEUMANN") for nm in names]).split("}\n", 1)[1])
    robin = case.get("robin", (None, []))
    with open(os.path.join(d, "p0.txt"), "w") as f:
        f.write(field_file(1, ["uniform 0"], [(nm, "ROBIN", robin[1]) if nm == robin[0] else
                                              (nm, "DIRICHLET" if nm == "xmin" else "NEUMANN") for nm in names]))


def potential_velocity(case):
    """Seeded U ~ uniform[-1,1]^3 per interior cell, in the reference's cell order."""
    nx, ny, nz = case["ndiv"]
    rng = np.random.default_rng(case.get("seed", 20261202))
    return rng.uniform(-1.0, 1.0, size=(nx * ny * nz, 3))


# ------------------------------------------------------------------ cases --
CFG0 = dict(name="cfg0_2d_n4", ndiv=(16, 1, 16), size=(1000, 100, 1000), order=(4, 0, 4),
            patches=PATCHES_2D, dt=0.005, check_steps=[1, 10, 100],
            T=["gaussian 0 0.5 500 50 350 250 1000 250"])

EULER_CASES = [
    CFG0,
    dict(name="e3_n3", ndiv=(3, 2, 2), size=(3000, 2000, 2000), order=(3, 3, 3),
         patches=PATCHES_3D, dt=0.01, check_steps=[1, 5], keep_fields=True,
         T=["gaussian 0 1.0 1500 1000 700 800 800 800"]),
    dict(name="e3_n7", ndiv=(2, 2, 2), size=(1000, 1000, 500), order=(7, 7, 7),
         patches=PATCHES_3D, dt=0.002, check_steps=[1, 3],
         T=["gaussian 0 0.5 400 600 200 300 300 300"]),
    dict(name="e3_visc_n3", ndiv=(3, 2, 2), size=(3000, 2000, 2000), order=(3, 3, 3),
         patches=PATCHES_3D, dt=0.01, check_steps=[1, 4], viscosity=1.5,
         T=["gaussian 0 1.0 1500 1000 700 800 800 800"]),
    dict(name="cfg0_2d_visc", ndiv=(16, 1, 16), size=(1000, 100, 1000), order=(4, 0, 4),
         patches=PATCHES_2D, dt=0.005, check_steps=[1, 10], viscosity=1.5,
         T=["gaussian 0 0.5 500 50 350 250 1000 250"]),
    dict(name="e3_aniso_bubbles", ndiv=(4, 3, 2), size=(1000, 900, 500), order=(2, 3, 4),
         patches=PATCHES_3D, dt=0.004, check_steps=[2],
         T=["gaussian 0 0.7 300 300 200 250 250 250", "gaussian 0 0.3 700 600 300 200 200 200",
            "gaussian 0 1.0 500 450 250 150 150 150"]),
    # non-affine elements: the top face tilted (z = 600 at x = 0, 700 at
    # x = 1000), so every element is a trapezoid in x-z and Jinv varies
    # inside it
    dict(name="e3_trilinear_n3", ndiv=(3, 3, 2), size=(1000, 800, 600), order=(3, 3, 3),
         corners=[(0, 0, 0), (1000, 0, 0), (1000, 0, 700), (0, 0, 600),
                  (0, 800, 0), (1000, 800, 0), (1000, 800, 700), (0, 800, 600)],
         patches=PATCHES_3D, dt=0.004, check_steps=[1, 3],
         T=["gaussian 0 0.5 500 400 250 200 200 200"]),
    dict(name="e3_wall_n3", ndiv=(5, 4, 6), size=(1000, 800, 600), order=(3, 3, 3),
         spacing="wall 3{5 4 6} 3{2 3 0.5}", patches=PATCHES_3D, dt=0.002, check_steps=[2],
         T=["gaussian 0 0.5 500 400 300 200 200 200"]),
]

# 2:1 non-conforming meshes: the base block refined by the reference's
# prepare tool (`prepare controls -refine`, once per pass, field.cpp:625-955)
# around the bubble, two levels; the refined grid file is a committed
# fixture (tests/golden/amr/<name>_grid_0.bin) and the initial fields are
# the deck's generators evaluated on it, as the reference's euler does (it
# reads <field>0.txt before the prepare tool's refined .bin, field.cpp:567).
AMR_CASES = [
    dict(name="amr3_n2", ndiv=(4, 4, 2), size=(1000, 1000, 500), order=(2, 2, 2),
         patches=PATCHES_3D, dt=0.002, check_steps=[1, 2], keep_fields=True,
         T=["gaussian 0 1.0 500 500 250 200 200 200"],
         refine=dict(direction=(1, 1, 1), field="T", field_min=0.05, field_max=0.4, max_level=2, passes=2)),
    dict(name="amr3_n5", ndiv=(2, 2, 2), size=(1000, 1000, 500), order=(5, 5, 5),
         patches=PATCHES_3D, dt=0.001, check_steps=[1], T=["gaussian 0 1.0 450 550 200 250 250 250"],
         refine=dict(direction=(1, 1, 1), field="T", field_min=0.05, field_max=0.4, max_level=2, passes=2)),
    # the first deck with its refined cells listed in reverse order (a text
    # grid): a refined cell then precedes its coarse neighbour and owns the
    # mortar faces (FMC 1), the other orientation of every mortar path
    dict(name="amr3_n2_rev", ndiv=(4, 4, 2), size=(1000, 1000, 500), order=(2, 2, 2),
         patches=PATCHES_3D, dt=0.002, check_steps=[1, 2], reverse_cells=True,
         T=["gaussian 0 1.0 500 500 250 200 200 200"],
         refine=dict(direction=(1, 1, 1), field="T", field_min=0.05, field_max=0.4, max_level=2, passes=2)),
    dict(name="amr3_n3_visc", ndiv=(4, 3, 2), size=(1000, 750, 500), order=(3, 3, 3),
         patches=PATCHES_3D, dt=0.002, check_steps=[1, 2], viscosity=1.5,
         T=["gaussian 0 1.0 400 375 250 200 200 200"],
         refine=dict(direction=(1, 1, 1), field="T", field_min=0.05, field_max=0.4, max_level=2, passes=2)),
]

GEOM_CASES = [dict(c) for c in EULER_CASES] + [
    dict(name="g2d_n4_small", ndiv=(6, 1, 5), size=(600, 100, 500), order=(4, 0, 4),
         patches=PATCHES_2D, keep=("FO", "FN", "faceid", "cell_faces")),
    dict(name="g3d_geometric_n2", ndiv=(5, 4, 6), size=(1000, 800, 600), order=(2, 2, 2),
         spacing="geometric 3{5 4 6} 3{2 3 0.5}", patches=PATCHES_3D),
    dict(name="gfv_543", ndiv=(5, 4, 3), size=(1, 1, 1), order=(0, 0, 0), patches=PATCHES_BOX),
]

POTENTIAL_CASES = [
    dict(name="pot_8", ndiv=(8, 8, 8), keep_fields=True),
    dict(name="pot_16", ndiv=(16, 16, 16)),
    dict(name="pot_12x8x6", ndiv=(12, 8, 6), seed=7),
]

# a ROBIN patch (applyImplicitBCs / applyExplicitBCs, field.h:2571-2573,
# 2677-2680): p on xmax mixes the value 1 and the slope 0.2 half and half.
# Run through the field API only (the C-ABI pressure solver and the oracle
# take NEUMANN and DIRICHLET patches).
POTENTIAL_ROBIN_CASES = [
    dict(name="pot_robin", ndiv=(10, 8, 6), seed=3, keep_fields=True,
         robin=("xmax", ["value 1", "shape 0.5", "tvalue 0.2"])),
]

# a sheared unit cube (x shifted by 0.3 z, y by 0.2 x): every face is
# non-orthogonal, so the laminar correction of lap (field.h:3676-3699) is
# live; with n_ORTHO 2 the pressure system is assembled three times, the
# later ones with the correction of the previous solution
POTENTIAL_SKEW_CASES = [
    dict(name="pot_skew", ndiv=(10, 8, 6), seed=5, n_ortho=2, keep_fields=True,
         corners=[(0, 0, 0), (1, 0.2, 0), (1.3, 0.2, 1), (0.3, 0, 1),
                  (0, 1, 0), (1, 1.2, 0), (1.3, 1.2, 1), (0.3, 1, 1)]),
]


# --------------------------------------------------------------- PISO --
# lid-driven cavity decks for the piso solver (apps/piso/piso.cpp), after
# the reference's examples/cavity: rho 1, viscosity 0.1, HYBRID convection,
# OVER_RELAXED non-orthogonal correction, laminar (turbulence_model NONE),
# two PISO correctors; the lid moves at U = (1, 0, 0).  Each case picks a
# linear-solver method / preconditioner (the default preconditioner is
# SSOR) and whether the momentum equation is solved (momentum_predictor,
# a non-symmetric system: biconjugate gradient).
PISO_PATCHES_3D = [("top_WALL", [(3, 2, 6, 7)]),
                   ("all_WALL", [(0, 1, 5, 4), (0, 1, 2, 3), (4, 5, 6, 7), (1, 5, 6, 2), (0, 4, 7, 3)])]
PISO_PATCHES_2D = [("top_WALL", [(3, 2, 6, 7)]), ("delete", [(0, 1, 2, 3), (4, 5, 6, 7)]),
                   ("all_WALL", [(1, 5, 6, 2), (0, 1, 5, 4), (0, 4, 7, 3)])]


def write_piso_case(d, case):
    os.makedirs(d, exist_ok=True)
    c = dict(case, order=(0, 0, 0))
    with open(os.path.join(d, "spec"), "w") as f:
        f.write(_spec(c))
    extra = "".join("    %-28s %s\n" % kv for kv in case.get("general", {}).items())
    piso = "".join("    %-28s %s\n" % kv for kv in case.get("piso", {}).items())
    with open(os.path.join(d, "controls"), "w") as f:
        f.write("""general
{
    solver                       piso
    mesh                         grid
    rho                          1
    viscosity                    1e-1
    state                        %s
    start_step                   0
    end_step                     %d
    write_interval               1
    dt                           %r
    convection_scheme            HYBRID
    nonortho_scheme              OVER_RELAXED
    blend_factor                 0.2
    method                       %s
    tolerance                    %r
    max_iterations               6400
    SOR_omega                    1.7
    gravity                      0 0 0
    npx                          0
    npy                          0
    npz                          0
%s}
piso
{
    turbulence_model             NONE
    velocity_UR                  0.7
    pressure_UR                  0.3
    n_PISO                       2
    n_ORTHO                      0
%s}
""" % (case.get("state", "TRANSIENT"), max(case["check_steps"]), case.get("dt", 0.005),
       case.get("method", "PCG"), case.get("tol", 1e-5), extra, piso))
    names = sorted(n for n, _ in case["patches"] if n != "delete")
    with open(os.path.join(d, "U0.txt"), "w") as f:
        f.write("size 3\ninternal 1\n{\n    uniform 0 0 0\n}\nboundary %d\n{\n" % len(names))
        for n in names:
            v = "1.0 0 0" if n == "top_WALL" else "0 0 0"
            f.write("    %s {\n        type DIRICHLET\n        value %s\n    }\n" % (n, v))
        f.write("}\n")
    with open(os.path.join(d, "p0.txt"), "w") as f:
        f.write(field_file(1, ["uniform %r" % case.get("p0", 0.0)], [(n, "NEUMANN") for n in names]))


# p starts at 0 and the tolerance is 1e-8, so that every solve iterates
# (the residual is relative to the solution's norm)
_CAV3D = dict(ndiv=(6, 6, 6), size=(1, 1, 1), patches=PISO_PATCHES_3D, check_steps=[1, 3], tol=1e-8)
PISO_CASES = [
    # pressure: conjugate gradient with the default SSOR preconditioner
    dict(_CAV3D, name="cav3d_ssor", keep_fields=True),
    # momentum solved (biconjugate gradient) and pressure, both DILU
    dict(_CAV3D, name="cav3d_dilu_mp", general={"preconditioner": "DILU"},
         piso={"momentum_predictor": "YES"}),
    # DIAG: pressure by the library's conjugate gradient, momentum by BiCG
    dict(_CAV3D, name="cav3d_diag_mp", general={"preconditioner": "DIAG"},
         piso={"momentum_predictor": "YES"}),
    # no preconditioner, BiCG momentum
    dict(_CAV3D, name="cav3d_none_mp", general={"preconditioner": "NONE"},
         piso={"momentum_predictor": "YES"}),
    # SOR and Jacobi iterations for both equations
    dict(_CAV3D, name="cav3d_sor_mp", method="SOR", piso={"momentum_predictor": "YES"}),
    dict(_CAV3D, name="cav3d_jac_mp", method="JAC", piso={"momentum_predictor": "YES"}),
    # steady state: under-relaxation (Relax) instead of the time derivative
    dict(_CAV3D, name="cav3d_steady", state="STEADY", check_steps=[1, 4]),
    # the reference's 2D cavity (the y faces deleted), 20 x 20 in x-z
    dict(name="cav2d_20", ndiv=(20, 1, 20), size=(1, 0.1, 1), patches=PISO_PATCHES_2D, check_steps=[1, 5],
         keep_fields=True, p0=101325.0, tol=1e-5),
    # the same with p from 0 and a tight tolerance
    dict(name="cav2d_20_tight", ndiv=(20, 1, 20), size=(1, 0.1, 1), patches=PISO_PATCHES_2D, check_steps=[1, 5],
         tol=1e-8),
]
