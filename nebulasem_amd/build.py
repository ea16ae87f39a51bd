"""Build the native libraries in-tree: host C++ with g++, HIP kernels with
hipcc for gfx950, linked into nebulasem_amd/libnebulasem_amd.so (the C-ABI);
the field API into nebulasem_amd/libnebulasem_field.so; and, where the
reference's source tree is present, its solver apps compiled unchanged
against the field API."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libnebulasem_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
              "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
HIP_FLAGS = ["-std=c++17", "-O3", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _sources():
    host, hip = [], []
    for d, _, files in os.walk(CSRC):
        for f in sorted(files):
            p = os.path.join(d, f)
            if f.endswith(".cpp"):
                host.append(p)
            elif f.endswith(".hip"):
                hip.append(p)
    return sorted(host), sorted(hip)


def _obj(src):
    rel = os.path.relpath(src, CSRC).replace(os.sep, "_")
    return os.path.join(OBJ, rel + ".o")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    host, hip = _sources()
    jobs = [["g++"] + HOST_FLAGS + ["-c", s, "-o", _obj(s)] for s in host]
    jobs += [[HIPCC] + HIP_FLAGS + ["-c", s, "-o", _obj(s)] for s in hip]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp]
         + [_obj(s) for s in host + hip] + ["-lpthread"])
    os.replace(tmp, LIB)
    return LIB


ASAN_LIB = os.path.join(HERE, "_obj_asan", "libnebulasem_amd_asan.so")


def build_sanitized():
    """The C-ABI library with its host C++ (mesh, geometry, grid and field
    files, the C-ABI glue) compiled with -fsanitize=address,undefined by
    g++; the HIP sources are compiled as usual (no device sanitizer) and
    linked in.  For the CPU test suite run under the sanitizers:

        NEBULASEM_AMD_LIB=<this library> LD_PRELOAD="$(g++ -print-file-name=libasan.so)
        $(g++ -print-file-name=libubsan.so)" ASAN_OPTIONS=detect_leaks=0 python -m pytest tests -m "not gpu"
    """
    out = os.path.dirname(ASAN_LIB)
    os.makedirs(out, exist_ok=True)
    host, hip = _sources()
    san = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined", "-g"]
    objs = []
    jobs = []
    for s in host:
        o = os.path.join(out, os.path.relpath(s, CSRC).replace(os.sep, "_") + ".o")
        jobs.append(["g++"] + HOST_FLAGS + san + ["-c", s, "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        list(ex.map(_run, jobs))
    missing = [s for s in hip if not os.path.exists(_obj(s))]
    if missing:
        build()
    _run(["g++", "-shared", "-o", ASAN_LIB] + san + objs + [_obj(s) for s in hip]
         + ["-L/opt/rocm/lib", "-lamdhip64", "-lpthread", "-Wl,-rpath,/opt/rocm/lib"])
    return ASAN_LIB


FIELD = os.path.join(HERE, "field")
FIELD_LIB = os.path.join(HERE, "libnebulasem_field.so")
APPS = os.path.join(FIELD, "_apps")
REFERENCE = "/root/reference"
# the reference's solvers compiled unchanged against the field API
API_APPS = ("euler", "potential", "piso")


def build_field(verbose=False):
    """The source-compatible field API (nebulasem_amd/field): its operator
    kernels and run-time, linked against the C-ABI library."""
    flags = HIP_FLAGS + ["-I" + os.path.join(CSRC, "hip"), "-I" + FIELD]
    srcs = sorted(os.path.join(FIELD, f) for f in os.listdir(FIELD) if f.endswith(".hip"))
    objs = [os.path.join(OBJ, "field_" + os.path.basename(s) + ".o") for s in srcs]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for r in ex.map(_run, [[HIPCC] + flags + ["-c", s, "-o", o] for s, o in zip(srcs, objs)]):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr)
    tmp = FIELD_LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
         + ["-L" + HERE, "-lnebulasem_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FIELD_LIB)
    return FIELD_LIB


def build_apps(reference=REFERENCE):
    """The reference's apps/<name>/<name>.cpp, read where they lie and
    compiled without change against the field API's headers, into
    nebulasem_amd/field/_apps/<name>.  Needs the reference's source tree."""
    built = []
    if not os.path.isdir(reference):
        return built
    os.makedirs(APPS, exist_ok=True)
    flags = HIP_FLAGS + ["-x", "hip", "-O2", "-I" + os.path.join(CSRC, "hip"), "-I" + FIELD]
    jobs = []
    for a in API_APPS:
        src = os.path.join(reference, "apps", a, a + ".cpp")
        jobs.append([HIPCC] + flags + ["-c", src, "-o", os.path.join(OBJ, "app_" + a + ".o")])
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        list(ex.map(_run, jobs))
    for a in API_APPS:
        out = os.path.join(APPS, a)
        _run([HIPCC, "--offload-arch=gfx950", "-o", out + ".tmp", os.path.join(OBJ, "app_" + a + ".o"),
              "-L" + HERE, "-lnebulasem_field", "-lnebulasem_amd", "-Wl,-rpath,$ORIGIN/../.."])
        os.replace(out + ".tmp", out)
        built.append(out)
    return built


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_field(verbose="-v" in sys.argv))
    print(build_apps())
