"""The field API library (nebulasem_amd/field) is built and links against
the C-ABI library; the reference's solver apps compiled against it, where
the reference's sources were present at build time, resolve every symbol
from it.  No GPU calls."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "nebulasem_amd")
LIB = os.path.join(PKG, "libnebulasem_field.so")
APPS = os.path.join(PKG, "field", "_apps")


def _symbols(path):
    out = subprocess.run(["nm", "-DC", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return out


def test_field_library_exports_the_api():
    assert os.path.exists(LIB), "run __graft_entry__.build()"
    syms = _symbols(LIB)
    for s in ("Solve(MeshMatrix<double, double, double> const&)", "Mesh::LoadMesh(", "Solver::Initialize(",
              "Util::read_params(", "Mesh::calc_courant(", "Prepare::refineMesh(", "nsem_field::div_flux(",
              "nsem_field::divf_volume(", "nsem_field::solve_fv_pcg("):
        assert s in syms, s


def test_field_library_links_the_c_abi():
    out = subprocess.run(["ldd", LIB], capture_output=True, text=True, check=True).stdout
    assert "libnebulasem_amd.so" in out and "not found" not in out


@pytest.mark.parametrize("app", ["euler", "potential"])
def test_reference_apps_resolve_against_the_api(app):
    exe = os.path.join(APPS, app)
    if not os.path.exists(exe):
        pytest.skip("built only where the reference's sources are present")
    out = subprocess.run(["ldd", exe], capture_output=True, text=True, check=True).stdout
    assert "libnebulasem_field.so" in out and "not found" not in out
