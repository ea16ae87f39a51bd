"""ctypes front end of the oracle (oracle/liboracle.so) — test infrastructure.

The oracle is the plain-C CPU restatement of the reference algorithms; it is
the checker for the HIP path and never part of the product."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

OM_NEUMANN, OM_SYMMETRY, OM_DIRICHLET = 0, 1, 2


ORACLE_ULP_SO = os.path.join(ROOT, "oracle", "liboracle_ulp.so")
ORACLE_CR_SO = os.path.join(ROOT, "oracle", "liboracle_cr.so")


def build_oracle():
    src = os.path.join(ROOT, "oracle", "nsem_oracle.c")
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-o",
                           ORACLE_SO, src, "-lm"])
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-DOM_PERTURB_POW",
                           "-o", ORACLE_ULP_SO, src, "-lm"])
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-DOM_CR_POW",
                           "-o", ORACLE_CR_SO, src, "-lquadmath", "-lm"])
    return ORACLE_SO


_lib = None


_lib_ulp = None


def _stale(so):
    src = os.path.join(ROOT, "oracle", "nsem_oracle.c")
    return not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src)


def lib():
    global _lib
    if _lib is None:
        if _stale(ORACLE_SO):
            build_oracle()
        _lib = C.CDLL(ORACLE_SO)
    return _lib


_lib_cr = None


def lib_cr():
    """The oracle built with a correctly rounded pow (see nsem_oracle.c)."""
    global _lib_cr
    if _lib_cr is None:
        if _stale(ORACLE_CR_SO):
            build_oracle()
        _lib_cr = C.CDLL(ORACLE_CR_SO)
    return _lib_cr


def lib_ulp():
    """The oracle built with the 1-ulp pow perturbation (see nsem_oracle.c)."""
    global _lib_ulp
    if _lib_ulp is None:
        if _stale(ORACLE_ULP_SO):
            build_oracle()
        _lib_ulp = C.CDLL(ORACLE_ULP_SO)
    return _lib_ulp


P = C.POINTER
PD = P(C.c_double)
PU = P(C.c_uint32)
PI = P(C.c_int32)


class OmMesh(C.Structure):
    _fields_ = [("npx", C.c_int32), ("npy", C.c_int32), ("npz", C.c_int32), ("np", C.c_int32),
                ("npf", C.c_int32), ("nbcs", C.c_uint32), ("ncells", C.c_uint32),
                ("nfacets", C.c_uint32), ("nbf", C.c_uint64), ("naf", C.c_uint64),
                ("FO", PU), ("FN", PU), ("fN", PD), ("fI", PD), ("fD", PD), ("cV", PD),
                ("cC", PD), ("Jinv", PD), ("psi", PD * 3), ("dpsi", PD * 3),
                ("allFaces", PU), ("fi0", PU), ("fi1", PU), ("nbnd", C.c_int32),
                ("boff", PU), ("bfaces", PU),
                ("FMC", PU), ("FOC", PU), ("FNC", PU), ("cell_off", PU), ("cell_faces", PU),
                ("faceid", PU), ("gFC", PD), ("psiRef", PD * 3), ("psiCor", PD * 3)]


class OmEuler(C.Structure):
    _fields_ = [("dt", C.c_double), ("P0", C.c_double), ("T0", C.c_double), ("cp", C.c_double),
                ("cv", C.c_double), ("g", C.c_double * 3), ("buoyancy", C.c_int32),
                ("bc_rho", PI), ("bc_U", PI), ("bc_T", PI), ("bc_p", PI),
                ("val_rho", PD), ("val_U", PD), ("val_T", PD), ("val_p", PD),
                ("diffusion", C.c_int32), ("viscosity", C.c_double), ("Pr", C.c_double),
                ("mol_exact", C.c_int32)]


def _p(a, t):
    return a.ctypes.data_as(t)


class OracleMesh:
    """Holds the numpy arrays a Mesh exports and the OmMesh view of them."""

    def __init__(self, mesh):
        inf = mesh.info
        self.keep = {}
        for name in ("FO", "FN", "fN", "fI", "fD", "cV", "cC", "allFaces", "faceIndices0",
                     "faceIndices1", "boff", "bfaces"):
            self.keep[name] = np.ascontiguousarray(mesh.array(name))
        if inf.np > 1:
            self.keep["Jinv"] = np.ascontiguousarray(mesh.array("Jinv"))
        else:
            self.keep["Jinv"] = np.zeros(9)
        for d in range(3):
            self.keep[f"psi{d}"] = mesh.array(f"psi{d}")
            self.keep[f"dpsi{d}"] = mesh.array(f"dpsi{d}")
        k = self.keep
        s = OmMesh()
        s.npx, s.npy, s.npz, s.np, s.npf = inf.npx, inf.npy, inf.npz, inf.np, inf.npf
        s.nbcs, s.ncells, s.nfacets = inf.nbcs, inf.ncells, inf.nfacets
        s.nbf, s.naf = inf.nbcs_field, inf.nall_field
        s.FO, s.FN = _p(k["FO"], PU), _p(k["FN"], PU)
        s.fN, s.fI, s.fD = _p(k["fN"], PD), _p(k["fI"], PD), _p(k["fD"], PD)
        s.cV, s.cC, s.Jinv = _p(k["cV"], PD), _p(k["cC"], PD), _p(k["Jinv"], PD)
        for d in range(3):
            s.psi[d] = _p(k[f"psi{d}"], PD)
            s.dpsi[d] = _p(k[f"dpsi{d}"], PD)
        s.allFaces, s.fi0, s.fi1 = _p(k["allFaces"], PU), _p(k["faceIndices0"], PU), _p(k["faceIndices1"], PU)
        s.nbnd = len(mesh.boundaries)
        s.boff, s.bfaces = _p(k["boff"], PU), _p(k["bfaces"], PU)
        # 2:1 non-conforming faces: the mortar projections run only where a
        # face is flagged (FMC 1 or 2)
        for name in ("FMC", "FOC", "FNC", "cell_off", "cell_faces", "faceid", "gFC"):
            k[name] = np.ascontiguousarray(mesh.array(name))
        if inf.np > 1 and k["FMC"].any():
            s.FMC, s.FOC, s.FNC = _p(k["FMC"], PU), _p(k["FOC"], PU), _p(k["FNC"], PU)
            s.cell_off, s.cell_faces, s.faceid = _p(k["cell_off"], PU), _p(k["cell_faces"], PU), _p(k["faceid"], PU)
            s.gFC = _p(k["gFC"], PD)
            for d in range(3):
                k[f"psiRef{d}"] = np.ascontiguousarray(mesh.array(f"psiRef{d}"))
                k[f"psiCor{d}"] = np.ascontiguousarray(mesh.array(f"psiCor{d}"))
                s.psiRef[d] = _p(k[f"psiRef{d}"], PD)
                s.psiCor[d] = _p(k[f"psiCor{d}"], PD)
        self.s = s
        self.info = inf


def euler_hydrostatic(om, params, T, p):
    n = om.info.nall_field
    p_ref, rho_ref, rho = np.empty(n), np.empty(n), np.empty(n)
    lib().om_euler_hydrostatic(C.byref(om.s), C.byref(params), _p(T, PD), _p(p, PD),
                               _p(p_ref, PD), _p(rho_ref, PD), _p(rho, PD))
    return p_ref, rho_ref, rho


def euler_diagnostics(om, params, p_ref, rho, U, T, p, cr_pow=False):
    """(Courant max, min, avg, mass, energy, volume) as euler.cpp:261-283."""
    out = np.empty(6)
    (lib_cr() if cr_pow else lib()).om_euler_diagnostics(C.byref(om.s), C.byref(params), _p(p_ref, PD), _p(rho, PD), _p(U, PD),
                               _p(T, PD), _p(p, PD), _p(out, PD))
    return out


def make_params(dt, bcs, P0=101325.0, T0=300.0, cp=1004.67, cv=715.5, g=(0, 0, -9.80606),
                buoyancy=1, viscosity=None, Pr=0.9, mol_exact=0):
    """bcs: dict field -> (kinds int32 array, values float64 array);
    viscosity: None for `diffusion NO`, else the artificial viscosity;
    mol_exact: the method-of-lines residual with the reference's expressions
    (0: p' = P0 (rho theta R/P0)^gamma - p_ref, gradient of theta' + T0) or
    their cancellation-free forms (1: p' = p_ref expm1(gamma log1p(Theta'/
    Theta_ref)), gradient of theta'), the device's forms."""
    e = OmEuler()
    e.dt, e.P0, e.T0, e.cp, e.cv = dt, P0, T0, cp, cv
    e.g[:] = list(g)
    e.buoyancy = buoyancy
    e.diffusion = 0 if viscosity is None else 1
    e.viscosity = 0.0 if viscosity is None else viscosity
    e.Pr = Pr
    e.mol_exact = int(mol_exact)
    keep = []
    for f in ("rho", "U", "T", "p"):
        kinds, vals = bcs[f]
        kinds = np.ascontiguousarray(kinds, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        keep += [kinds, vals]
        setattr(e, "bc_" + f, _p(kinds, PI))
        setattr(e, "val_" + f, _p(vals, PD))
    e._keep = keep
    return e


def euler_split_steps(om, params, p_ref, rho_ref, rho, U, T, p, nsteps, perturb_pow=False, cr_pow=False):
    """perturb_pow: the build whose equation-of-state pow is moved by +-1 ulp;
    cr_pow: the build with a correctly rounded pow."""
    L = lib_ulp() if perturb_pow else (lib_cr() if cr_pow else lib())
    st = L.om_euler_split_steps(C.byref(om.s), C.byref(params), _p(p_ref, PD), _p(rho_ref, PD),
                                    _p(rho, PD), _p(U, PD), _p(T, PD), _p(p, PD), int(nsteps))
    return st


def fv_poisson(om, bc_kind, bc_val, U, p, tol, max_it):
    bc_kind = np.ascontiguousarray(bc_kind, dtype=np.int32)
    bc_val = np.ascontiguousarray(bc_val, dtype=np.float64)
    it = C.c_int32()
    r0, r = C.c_double(), C.c_double()
    st = lib().om_fv_poisson(C.byref(om.s), _p(bc_kind, PI), _p(bc_val, PD), _p(U, PD), _p(p, PD),
                             C.c_double(tol), C.c_int32(max_it), C.byref(it), C.byref(r0), C.byref(r))
    return st, it.value, r0.value, r.value


def fv_correct_velocity(om, bc_kind, bc_val, p, U):
    """U -= gradf(p, true) on the interior cells (p: all cells, ghosts set
    from bc_kind/bc_val first; U: all cells xyz, updated in place)."""
    bc_kind = np.ascontiguousarray(bc_kind, dtype=np.int32)
    bc_val = np.ascontiguousarray(bc_val, dtype=np.float64)
    lib().om_fv_correct_velocity(C.byref(om.s), _p(bc_kind, PI), _p(bc_val, PD), _p(p, PD), _p(U, PD))


def euler_ssprk3_steps(om, params, p_ref, rho_ref, rho, U, T, p, nsteps):
    return lib().om_euler_ssprk3_steps(C.byref(om.s), C.byref(params), _p(p_ref, PD), _p(rho_ref, PD),
                                       _p(rho, PD), _p(U, PD), _p(T, PD), _p(p, PD), int(nsteps))


def euler_mol_rhs(om, params, p_ref, rho_ref, rho, U, T, Tp=None):
    """The method-of-lines residual at the state rho, U, theta' (every node;
    ghosts set in place from the boundary conditions): L (naf, 5) and the
    Rusanov divergences r (naf, 5) of the density, momentum and rho theta
    fluxes.  Tp: Theta' on the interior nodes (for mol_exact)."""
    n = om.info.nall_field
    L, r = np.empty(5 * n), np.empty(5 * n)
    tp = None if Tp is None else _p(np.ascontiguousarray(Tp, dtype=np.float64), PD)
    lib().om_euler_mol_rhs(C.byref(om.s), C.byref(params), _p(p_ref, PD), _p(rho_ref, PD), _p(rho, PD),
                           _p(U, PD), _p(T, PD), tp, _p(L, PD), _p(r, PD))
    return L.reshape(n, 5), r.reshape(n, 5)
