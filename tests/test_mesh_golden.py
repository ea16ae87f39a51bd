"""The host mesh pipeline (grid generator, LoadMesh fix-ups, DG/FV node
geometry) reproduces the reference's arrays bit for bit: SHA-256 of every
array against the reference build's dumps (tests/golden/)."""
import os

import numpy as np
import pytest

from _cases import AMR_CASES, GEOM_CASES, write_euler_case
from _euler_setup import amr_grid
from _golden import arrays, fixtures, sha
from nebulasem_amd import Mesh

FX = fixtures()["geometry"]
AMR_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "amr")


@pytest.mark.parametrize("case", AMR_CASES, ids=[c["name"] for c in AMR_CASES])
def test_amr_geometry_bit_exact(case):
    """A 2:1 non-conforming mesh (two levels refined by the reference's
    prepare tool, grid file committed): LoadMesh's fix-ups with split cell
    sides (merged outlines, mortar flags FMC), the DG node geometry and face
    maps, and the 2:1 mortar matrices psiRef/psiCor agree bit for bit."""
    mesh = Mesh.load(amr_grid(case), order=case["order"])
    fx = FX[case["name"]]
    assert mesh.boundaries == fx["bnames"]
    bad = []
    for name, h in fx["sha256"].items():
        a = mesh.array(name)
        if a.size != fx["sizes"][name] or sha(a) != h:
            bad.append(name)
    assert not bad, f"arrays differing from the reference: {bad}"
    # mortar faces are present; the prepare tool numbers a coarse cell
    # before its refined neighbours, so the coarse cell owns them (FMC 2);
    # with the cells listed in reverse the refined cell owns them (FMC 1)
    fmc = mesh.array("FMC")
    m = 1 if case.get("reverse_cells") else 2
    assert (fmc == m).sum() > 0 and set(np.unique(fmc)) <= {0, m}


@pytest.mark.parametrize("case", GEOM_CASES, ids=[c["name"] for c in GEOM_CASES])
def test_geometry_bit_exact(tmp_path, case):
    write_euler_case(str(tmp_path), case)
    mesh = Mesh.from_spec(os.path.join(tmp_path, "spec"), order=case["order"])
    fx = FX[case["name"]]
    assert mesh.boundaries == fx["bnames"]
    bad = []
    for name, h in fx["sha256"].items():
        a = mesh.array(name)
        if a.size != fx["sizes"][name] or sha(a) != h:
            bad.append(name)
    assert not bad, f"arrays differing from the reference: {bad}"


def test_grid_roundtrip_text_and_binary(tmp_path):
    case = GEOM_CASES[1]
    write_euler_case(str(tmp_path), case)
    m0 = Mesh.from_spec(os.path.join(tmp_path, "spec"), order=case["order"])
    for ext in (".txt", ".bin"):
        path = os.path.join(tmp_path, "grid_0" + ext)
        m0.write_grid(path)
        m1 = Mesh.load(path, order=case["order"])
        for name in ("FO", "FN", "cell_faces", "faceid", "Jinv", "cV"):
            a, b = m0.array(name), m1.array(name)
            if ext == ".bin" or a.dtype == np.uint32:
                assert np.array_equal(a, b), (ext, name)
            else:  # text grids keep 12 significant digits
                assert np.allclose(a, b, rtol=1e-10, atol=1e-9), (ext, name)


def test_small_case_arrays_match_stored_reference(tmp_path):
    case = [c for c in GEOM_CASES if c["name"] == "g2d_n4_small"][0]
    write_euler_case(str(tmp_path), case)
    mesh = Mesh.from_spec(os.path.join(tmp_path, "spec"), order=case["order"])
    arr = arrays()
    for name in case["keep"]:
        assert np.array_equal(mesh.array(name), arr[f"geom_{case['name']}_{name}"]), name


def test_lgl_closed_forms():
    """LGL nodes/weights against closed forms; D exact on polynomials <= N."""
    from _cases import BOX_CORNERS, PATCHES_BOX
    corners = BOX_CORNERS(1, 1, 1)
    patches = [(n, [corners[i] for i in g[0]]) for n, g in PATCHES_BOX]
    m = Mesh.box(corners, (1, 1, 1), patches=patches, order=(4, 4, 6))
    x, w = m.array("xgl0"), m.array("wgl0")
    s = np.sqrt(3.0 / 7.0)
    assert np.allclose(x, [-1, -s, 0, s, 1], atol=1e-15)
    assert np.allclose(w, [0.1, 49 / 90, 32 / 45, 49 / 90, 0.1], atol=1e-15)
    # N=6: weights 2/(N(N+1) P_N(x)^2) and sum 2
    x6, w6 = m.array("xgl2"), m.array("wgl2")
    from numpy.polynomial import legendre as L
    PN = L.legval(x6, [0] * 6 + [1])
    assert np.allclose(w6, 2.0 / (6 * 7 * PN ** 2), rtol=1e-13)
    assert abs(w6.sum() - 2.0) < 1e-14
    # derivative matrix dpsi[s*n+i] = l_i'(x_s): exact on x^k, k <= N
    D = m.array("dpsi2").reshape(7, 7)
    for k in range(7):
        assert np.allclose(D @ x6 ** k, k * x6 ** max(k - 1, 0) if k else 0 * x6, atol=1e-12)
    assert np.allclose(D.sum(axis=1), 0, atol=1e-13)
    assert np.isclose(D[0, 0], -6 * 7 / 4) and np.isclose(D[-1, -1], 6 * 7 / 4)
