"""Device diagnostics (Courant max/min/avg, mass, energy, volume; the
reference's per-step reductions, euler.cpp:261-283) against the oracle's
serial restatement, which itself reproduces the reference build's printed
values digit for digit (test_oracle.py).

Tolerance.  Volume, mass and energy: 1e-12 relative (summation order
differs).  Courant max and avg: 1e-12 relative against the oracle with the
correctly rounded pow the device uses (the split states then agree to
1e-14); Courant min: within 1e-12 of the max (the minimum speed is
round-off sized).  The energy loss
after the run is checked against the reference build's printed value to
1e-4 relative plus 2e-14 absolute (a difference of two sums, each good to
~1e-15 in a different summation order)."""
import numpy as np
import pytest

from _cases import EULER_CASES
from test_gpu_dg_split import setup
from test_oracle import FX, _log_numbers

pytestmark = pytest.mark.gpu

CASES = [c for c in EULER_CASES if c["name"] in ("cfg0_2d_n4", "e3_n3", "e3_wall_n3")]


def check(dev, ref):
    assert abs(dev[5] - ref[5]) <= 1e-12 * abs(ref[5]), ("volume", dev[5], ref[5])
    assert abs(dev[3] - ref[3]) <= 1e-12 * abs(ref[3]), ("mass", dev[3], ref[3])
    assert abs(dev[4] - ref[4]) <= 1e-12 * abs(ref[4]), ("energy", dev[4], ref[4])
    for k in (0, 2):
        assert abs(dev[k] - ref[k]) <= 1e-12 * abs(ref[k]) + 1e-300, ("courant", k, dev[k], ref[k])
    assert abs(dev[1] - ref[1]) <= 1e-12 * abs(ref[0]), ("courant min", dev[1], ref[1])


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_split_diagnostics_match_oracle(oracle, tmp_path, case):
    print(f"[gpu] diagnostics split {case['name']}", flush=True)
    mesh, om, prm, dg, (p_ref, rho_ref, rho, U, T, p) = setup(oracle, tmp_path, case)
    d0 = dg.diagnostics()
    r0 = oracle.euler_diagnostics(om, prm, p_ref, rho, U, T, p)
    check(d0, r0)
    steps = max(case["check_steps"])
    dg.split_steps(steps)
    oracle.euler_split_steps(om, prm, p_ref, rho_ref, rho, U, T, p, steps, cr_pow=True)
    d1 = dg.diagnostics()
    r1 = oracle.euler_diagnostics(om, prm, p_ref, rho, U, T, p, cr_pow=True)
    check(d1, r1)
    _, loss = _log_numbers(FX["euler"][case["name"]]["log_tail"])
    e_loss = (d0[4] - d1[4]) / d0[4]
    # each energy sum is good to ~1e-15 relative (summation order), so a loss
    # of round-off size is held to that absolute level
    assert abs(e_loss - float(loss[1])) <= 1e-4 * abs(float(loss[1])) + 2e-14, (e_loss, loss[1])
    dg.close()


def test_rk3_diagnostics_match_oracle(oracle, tmp_path):
    case = [c for c in EULER_CASES if c["name"] == "e3_n3"][0]
    print(f"[gpu] diagnostics rk3 {case['name']}", flush=True)
    mesh, om, prm, dg, (p_ref, rho_ref, rho, U, T, p) = setup(oracle, tmp_path, case)
    dg.rk3_steps(3, 64)
    prm.mol_exact = 1
    oracle.euler_ssprk3_steps(om, prm, p_ref, rho_ref, rho, U, T, p, 3)
    check(dg.diagnostics(), oracle.euler_diagnostics(om, prm, p_ref, rho, U, T, p))
    dg.close()
