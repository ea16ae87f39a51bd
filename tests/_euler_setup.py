"""Shared set-up of an Euler case: product mesh + product field reader,
initial state as the reference's euler app builds it."""
import os
import shutil

import numpy as np

from _cases import write_euler_case
from nebulasem_amd import Mesh, read_field

GRAVITY = (0.0, 0.0, -9.80606)


AMR_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "amr")


def amr_grid(case):
    """The committed refined grid of an AMR deck (text for the reordered one)."""
    ext = "txt" if case.get("reverse_cells") else "bin"
    return os.path.join(AMR_DIR, f"{case['name']}_grid_0.{ext}")


def load_case(tmpdir, case):
    """The deck's mesh (the grid generator, or for an AMR deck the committed
    refined grid) and its initial fields."""
    d = os.path.join(str(tmpdir), case["name"])
    write_euler_case(d, case)
    if case.get("refine"):
        mesh = Mesh.load(amr_grid(case), order=case["order"])
        if not case.get("reverse_cells"):
            # the prepare tool's refined temperature, which the reference's
            # euler reads before the deck's T0.txt
            shutil.copy(os.path.join(AMR_DIR, f"{case['name']}_T0.bin"), os.path.join(d, "T0.bin"))
            os.remove(os.path.join(d, "T0.txt"))
    else:
        mesh = Mesh.from_spec(os.path.join(d, "spec"), order=case["order"])
    f = {}
    for name, nc in (("T", 1), ("U", 3), ("p", 1), ("rho", 1)):
        path = os.path.join(d, name + "0.bin")
        if not os.path.exists(path):
            path = os.path.join(d, name + "0.txt")
        f[name] = read_field(mesh, path, nc, GRAVITY)
    return mesh, f


def bc_table(f):
    return {k: (f[k][1], f[k][2].reshape(-1)) for k in ("rho", "U", "T", "p")}
