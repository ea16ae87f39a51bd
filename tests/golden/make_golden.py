"""Regenerate the golden fixtures from the reference build (oracle/_ref).

Runs in this container only (it needs /root/reference built by
`make -f oracle/ref.mk`).  For each case it writes the reference's input
deck into a temporary directory, runs the reference's own executables
(mesh, geomdump, euler, potential) there, and records:

  * SHA-256 of every loaded-mesh / geometry array the reference produces
    (geomdump), to pin the mesh pipeline bit for bit;
  * SHA-256 plus norms of the reference's output fields after N steps
    (euler: rho U T p; potential: p and the PCG iteration count);
  * for the small cases, the output arrays themselves (npz), so that a
    tolerance comparison is possible if a hash ever differs.

Usage:  python tests/golden/make_golden.py
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
from _cases import (AMR_CASES, EULER_CASES, GEOM_CASES, PISO_CASES, POTENTIAL_CASES,  # noqa: E402
                    POTENTIAL_ROBIN_CASES, POTENTIAL_SKEW_CASES, write_euler_case, write_piso_case, write_potential_case)
from _fieldio import read_field_bin  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
FIELDS = os.path.join(HERE, "fields")
AMR = os.path.join(HERE, "amr")
GEOM_SKIP = {"header", "bnames"}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def read_dump(path):
    out = {}
    with open(path, "rb") as f:
        while True:
            nm = f.read(32)
            if len(nm) < 32:
                break
            name = nm.split(b"\0")[0].decode()
            kind = int(np.frombuffer(f.read(4), np.uint32)[0])
            cnt = int(np.frombuffer(f.read(8), np.uint64)[0])
            dt = {0: np.uint32, 1: np.float64}[kind]
            out[name] = np.frombuffer(f.read(cnt * np.dtype(dt).itemsize), dt)
    return out


def run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{cmd} failed:\n{r.stdout}\n{r.stderr}")
    return r.stdout


def geometry_hashes(case, d, keep_arrays, make_grid=True):
    if make_grid:
        run([os.path.join(REF, "mesh"), "spec", "-o", "grid_0.bin"], d)
    run([os.path.join(REF, "geomdump"), "controls", "geom.bin"], d)
    dump = read_dump(os.path.join(d, "geom.bin"))
    out = {"sha256": {}, "sizes": {}}
    for k, v in dump.items():
        if k in GEOM_SKIP:
            continue
        if k == "fD" and tuple(case["order"]) != (0, 0, 0):
            continue  # DG penalty factor: only the implicit DG laplacian uses it
        out["sha256"][k] = sha(v)
        out["sizes"][k] = int(v.size)
    out["bnames"] = bytes(dump["bnames"].astype(np.uint8)).decode().split()
    arrays = {k: v for k, v in dump.items() if k in keep_arrays}
    return out, arrays


def euler_record(case, d, arrays):
    """Run the reference euler in d and hash its output fields."""
    log = run([os.path.join(REF, "euler"), "controls"], d)
    rec = {"steps": {}}
    for s in case["check_steps"]:
        rec["steps"][str(s)] = {}
        for name in ("rho", "U", "T", "p"):
            v = read_field_bin(os.path.join(d, f"{name}{s}.bin")).reshape(-1)
            rec["steps"][str(s)][name] = {"sha256": sha(v), "l2": float(np.linalg.norm(v)),
                                          "max": float(np.max(np.abs(v))), "n": int(v.size)}
            if case.get("keep_fields"):
                arrays[f"euler_{case['name']}_{name}{s}"] = v
    rec["log_tail"] = log.strip().splitlines()[-3:]
    return rec


def reverse_cells(path):
    """Rewrite a text grid (MeshObject::writeTextMesh, mesh.h:200-216) with
    its cells in reverse order; vertices, facets and boundaries unchanged."""
    text = open(path).read()
    # blocks: vertices "n\n{ ... }", facets "n\n{ ... }", cells "n\n{ ... }", boundaries
    pos = 0
    starts = []
    for _ in range(3):
        b = text.index("{", pos)
        depth, i = 0, b
        while True:
            if text[i] == "{":
                depth += 1
            elif text[i] == "}":
                depth -= 1
                if depth == 0:
                    break
            i += 1
        starts.append((b, i))
        pos = i + 1
    b, e = starts[2]
    lines = [ln for ln in text[b + 1:e].split("\n") if ln.strip()]
    new = text[:b + 1] + "\n" + "\n".join(reversed(lines)) + "\n" + text[e:]
    open(path, "w").write(new)


def refine(case, d):
    """Base grid from the mesh tool, then the prepare tool's refinement
    passes; the refined grid is kept as a fixture."""
    if case.get("reverse_cells"):
        # the prepare tool writes its refined grid as text
        write_euler_case(d, dict(case, write_format="TEXT"))
    run([os.path.join(REF, "mesh"), "spec", "-o", "grid_0.bin"], d)
    out = []
    for _ in range(case["refine"]["passes"]):
        log = run([os.path.join(REF, "prepare"), "controls", "-refine", "-start", "0"], d)
        out += [ln.strip() for ln in log.splitlines() if ln.startswith(("Level", "Refining "))]
    os.makedirs(AMR, exist_ok=True)
    name = case["name"]
    if case.get("reverse_cells"):
        # text grid, cells reversed (the reference reads a grid's .txt
        # before its .bin, field.cpp:107); the deck's generators for the
        # fields, evaluated on the nodes, independent of the cell order.  The
        # refinement tree names cells by index and is not kept.
        reverse_cells(os.path.join(d, "grid_0.txt"))
        shutil.copy(os.path.join(d, "grid_0.txt"), os.path.join(AMR, f"{name}_grid_0.txt"))
        for f in os.listdir(d):
            if f[:-4] in ("T0", "U0", "p0", "rho0", "amrTree_0"):
                os.remove(os.path.join(d, f))
        write_euler_case(d, case)
        return out
    else:
        shutil.copy(os.path.join(d, "grid_0.bin"), os.path.join(AMR, f"{name}_grid_0.bin"))
        # the prepare tool's refined temperature (refineField), which euler
        # reads in preference to the deck's T0.txt (field.h:1579-1584); the
        # other fields are uniform zeros, the same either way
        shutil.copy(os.path.join(d, "T0.bin"), os.path.join(AMR, f"{name}_T0.bin"))
    # the refinement tree the prepare tool wrote beside the grid
    shutil.copy(os.path.join(d, "amrTree_0.bin"), os.path.join(AMR, f"{name}_amrTree_0.bin"))
    return out


def main():
    # with case names as arguments, only those cases are rerun and merged
    # into the committed fixtures
    only = set(sys.argv[1:])
    fixtures = {"geometry": {}, "euler": {}, "potential": {}}
    arrays = {}
    if only:
        with open(os.path.join(HERE, "reference_fixtures.json")) as f:
            fixtures = json.load(f)
        with np.load(os.path.join(HERE, "reference_arrays.npz")) as z:
            arrays = {k: z[k] for k in z.files}
    pick = (lambda cases: [c for c in cases if c["name"] in only]) if only else (lambda cases: cases)
    tmp = tempfile.mkdtemp(prefix="nsem_golden_")
    try:
        for case in pick(GEOM_CASES):
            d = os.path.join(tmp, "g_" + case["name"])
            write_euler_case(d, case)
            fx, arr = geometry_hashes(case, d, keep_arrays=case.get("keep", ()))
            fixtures["geometry"][case["name"]] = fx
            for k, v in arr.items():
                arrays[f"geom_{case['name']}_{k}"] = v
            print("geometry", case["name"], len(fx["sha256"]), "arrays")
        for case in pick(EULER_CASES):
            d = os.path.join(tmp, "e_" + case["name"])
            write_euler_case(d, case)
            run([os.path.join(REF, "mesh"), "spec", "-o", "grid_0.bin"], d)
            log = run([os.path.join(REF, "euler"), "controls"], d)
            rec = {"steps": {}}
            for s in case["check_steps"]:
                rec["steps"][str(s)] = {}
                for name in ("rho", "U", "T", "p"):
                    v = read_field_bin(os.path.join(d, f"{name}{s}.bin")).reshape(-1)
                    rec["steps"][str(s)][name] = {"sha256": sha(v), "l2": float(np.linalg.norm(v)),
                                                  "max": float(np.max(np.abs(v))), "n": int(v.size)}
                    if case.get("keep_fields"):
                        arrays[f"euler_{case['name']}_{name}{s}"] = v
            rec["log_tail"] = log.strip().splitlines()[-3:]
            fixtures["euler"][case["name"]] = rec
            if case.get("keep_fields"):
                # the reference's own field files, binary and text (the same
                # deck rerun with write_format TEXT), for the writer's tests
                s1 = case["check_steps"][0]
                os.makedirs(FIELDS, exist_ok=True)
                for name in ("T", "U"):
                    shutil.copy(os.path.join(d, f"{name}{s1}.bin"),
                                os.path.join(FIELDS, f"{case['name']}_{name}{s1}.bin"))
                dt = os.path.join(tmp, "t_" + case["name"])
                write_euler_case(dt, dict(case, write_format="TEXT", check_steps=[s1]))
                run([os.path.join(REF, "mesh"), "spec", "-o", "grid_0.bin"], dt)
                run([os.path.join(REF, "euler"), "controls"], dt)
                for name in ("T", "U"):
                    shutil.copy(os.path.join(dt, f"{name}{s1}.txt"),
                                os.path.join(FIELDS, f"{case['name']}_{name}{s1}.txt"))
                # the mesh tool's grid file, binary and text
                shutil.copy(os.path.join(d, "grid_0.bin"), os.path.join(FIELDS, f"{case['name']}_grid_0.bin"))
                run([os.path.join(REF, "mesh"), "spec", "-o", "grid_t.txt"], dt)
                shutil.copy(os.path.join(dt, "grid_t.txt"), os.path.join(FIELDS, f"{case['name']}_grid_0.txt"))
            print("euler", case["name"], "ok")
        for case in pick(AMR_CASES):
            d = os.path.join(tmp, "a_" + case["name"])
            write_euler_case(d, case)
            passes = refine(case, d)
            fx, arr = geometry_hashes(case, d, keep_arrays=(), make_grid=False)
            fx["refine_log"] = passes
            fixtures["geometry"][case["name"]] = fx
            rec = euler_record(case, d, arrays)
            fixtures["euler"][case["name"]] = rec
            print("amr", case["name"], passes, rec["log_tail"][-1])
        for case in pick(POTENTIAL_CASES + POTENTIAL_SKEW_CASES + POTENTIAL_ROBIN_CASES):
            d = os.path.join(tmp, "p_" + case["name"])
            write_potential_case(d, case)
            run([os.path.join(REF, "mesh"), "spec", "-o", "grid_0.bin"], d)
            log = run([os.path.join(REF, "potential"), "controls"], d)
            line = [ln for ln in log.splitlines() if "PCG" in ln][-1]
            toks = line.split()
            it = int(toks[[t.endswith("Iterations") for t in toks].index(True) + 1])
            res = float(toks[toks.index("Final") + 2])
            p = read_field_bin(os.path.join(d, "p1.bin")).reshape(-1)
            # the corrected velocity U -= gradf(p) (potential.cpp:66-67)
            u = read_field_bin(os.path.join(d, "U1.bin")).reshape(-1)
            fixtures["potential"][case["name"]] = {
                "iterations": it, "final_residual": res, "log_line": line.split("]", 1)[-1].strip(),
                "p_sha256": sha(p), "p_l2": float(np.linalg.norm(p)), "n": int(p.size),
                "U_sha256": sha(u), "U_l2": float(np.linalg.norm(u)), "U_n": int(u.size)}
            if case.get("keep_fields"):
                arrays[f"potential_{case['name']}_p"] = p
                arrays[f"potential_{case['name']}_U"] = u
            print("potential", case["name"], it, res)
        for case in pick(PISO_CASES):
            d = os.path.join(tmp, "s_" + case["name"])
            write_piso_case(d, case)
            run([os.path.join(REF, "mesh"), "spec", "-o", "grid_0.bin"], d)
            log = run([os.path.join(REF, "piso"), "controls"], d)
            rec = {"steps": {}, "solver_lines": [ln.split("]", 1)[-1].strip() for ln in log.splitlines()
                                                 if "Iterations" in ln]}
            for s in case["check_steps"]:
                rec["steps"][str(s)] = {}
                for name in ("U", "p"):
                    v = read_field_bin(os.path.join(d, f"{name}{s}.bin")).reshape(-1)
                    rec["steps"][str(s)][name] = {"sha256": sha(v), "l2": float(np.linalg.norm(v)),
                                                  "max": float(np.max(np.abs(v))), "n": int(v.size)}
                    if case.get("keep_fields"):
                        arrays[f"piso_{case['name']}_{name}{s}"] = v
            fixtures.setdefault("piso", {})[case["name"]] = rec
            print("piso", case["name"], len(rec["solver_lines"]), "solves")
    finally:
        shutil.rmtree(tmp)
    with open(os.path.join(HERE, "reference_fixtures.json"), "w") as f:
        json.dump(fixtures, f, indent=1, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "reference_arrays.npz"), **arrays)


if __name__ == "__main__":
    main()
