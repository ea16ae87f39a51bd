"""Regenerate the method-of-lines residual fixtures from the reference build.

Runs in this container only (it needs oracle/_ref/euler_tendency, built by
`make -f oracle/ref.mk`, and the reference's mesh tool).  For each case:

  1. the deck's initial state (the oracle's hydrostatic set-up, bit-identical
     to the reference's), advanced `steps` SSP-RK3 steps by the oracle, then
     taken to the second stage's state q + dt L(q): a mid-RK state that no
     split step produces;
  2. that state (interior rho, U, theta') written to state.bin and handed to
     the reference's own divf / rusanov / gradf<true> through
     oracle/ref_tendency.cpp, which writes the Rusanov divergences r of the
     density, momentum and rho theta fluxes and p';
  3. recorded: SHA-256 of the state and of each r array (interior nodes),
     their norms, and for the smaller cases the arrays themselves.

Usage:  python tests/golden/make_tendency.py
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)
from _tendency import TENDENCY_CASES, mid_rk_state, prepare_case  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
OUT_JSON = os.path.join(HERE, "tendency_fixtures.json")
OUT_NPZ = os.path.join(HERE, "tendency_arrays.npz")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run_reference(d, state, nb):
    rho, U, T = state
    np.concatenate([rho[:nb], U[:3 * nb], T[:nb]]).astype(np.float64).tofile(os.path.join(d, "state.bin"))
    r = subprocess.run([os.path.join(REF, "euler_tendency"), "controls", "state.bin", "tend.bin"], cwd=d,
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"euler_tendency failed ({r.returncode}):\n{r.stdout}\n{r.stderr}")
    out = np.fromfile(os.path.join(d, "tend.bin"), dtype=np.float64)
    assert out.size == 6 * nb
    return {"r_rho": out[:nb], "r_m": out[nb:4 * nb], "r_T": out[4 * nb:5 * nb], "p": out[5 * nb:]}


def make_grid(case, d):
    if case.get("refine"):
        from _euler_setup import amr_grid
        src = amr_grid(case)
        shutil.copy(src, os.path.join(d, "grid_0" + os.path.splitext(src)[1]))
    else:
        r = subprocess.run([os.path.join(REF, "mesh"), "spec", "-o", "grid_0.bin"], cwd=d,
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stdout + r.stderr)


def main():
    import _oracle as oracle
    oracle.build_oracle()
    fx, arrays = {}, {}
    for case in TENDENCY_CASES:
        with tempfile.TemporaryDirectory() as tmp:
            d, mesh, om, prm, state0 = prepare_case(oracle, tmp, case)
            make_grid(case, d)
            nb = mesh.info.nbcs_field
            state = mid_rk_state(oracle, om, prm, state0, case)
            ref = run_reference(d, state, nb)
            rec = {"nb": int(nb), "state_sha256": sha(np.concatenate([a[:n] for a, n in zip(state, (nb, 3 * nb, nb))])),
                   "sha256": {k: sha(v) for k, v in ref.items()},
                   "l2": {k: float(np.linalg.norm(v)) for k, v in ref.items()}}
            fx[case["name"]] = rec
            if nb <= 2000:
                for k, v in ref.items():
                    arrays[f"{case['name']}_{k}"] = v
            print(case["name"], nb, {k: f"{v:.6e}" for k, v in rec["l2"].items()})
    with open(OUT_JSON, "w") as f:
        json.dump(fx, f, indent=1, sort_keys=True)
    np.savez_compressed(OUT_NPZ, **arrays)


if __name__ == "__main__":
    main()
