"""Shared set-up of an Euler case: product mesh + product field reader,
initial state as the reference's euler app builds it."""
import os

import numpy as np

from _cases import write_euler_case
from nebulasem_amd import Mesh, read_field

GRAVITY = (0.0, 0.0, -9.80606)


def load_case(tmpdir, case):
    d = os.path.join(str(tmpdir), case["name"])
    write_euler_case(d, case)
    mesh = Mesh.from_spec(os.path.join(d, "spec"), order=case["order"])
    f = {}
    for name, nc in (("T", 1), ("U", 3), ("p", 1), ("rho", 1)):
        f[name] = read_field(mesh, os.path.join(d, name + "0.txt"), nc, GRAVITY)
    return mesh, f


def bc_table(f):
    return {k: (f[k][1], f[k][2].reshape(-1)) for k in ("rho", "U", "T", "p")}
