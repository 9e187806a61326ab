"""The reference's C++ library interface (include/scc.h:49-53, include/load.h:6)
as exported by libscc_amd.so: a caller compiled against the reference's own
prototypes (tests/dropin_caller.cpp) links unchanged, loads a fixture through
loadFullGraph and receives the reference layout, and sees failures as a thrown
`const char *` (src/scc.cu:104-108).  The GPU tests run the drivers through the
same caller.  The caller binary is built per session into a temporary folder."""
import os
import subprocess

import numpy as np
import pytest

import detectingscc_amd as scc
from oracle import pyoracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "detectingscc_amd", "lib")
CALLER = os.path.join(ROOT, "tests", "dropin_caller.cpp")


def _fnv(a: np.ndarray) -> str:
    h = 1469598103934665603
    for x in a.astype(np.uint64).tolist():
        h = ((h ^ x) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return f"{h:016x}"


@pytest.fixture(scope="session")
def caller(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("dropin") / "dropin_caller")
    r = subprocess.run(["g++", "-O1", "-std=c++11", CALLER, "-o", exe, f"-L{LIBDIR}", "-lscc_amd",
                        f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath,/opt/rocm/lib"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    return exe


def run_caller(exe, *args, env=None, timeout=120):
    e = dict(os.environ)
    e.pop("SCC_PARTITION_FILE", None)
    e.update(env or {})
    return subprocess.run([exe, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=timeout, env=e)


def test_reference_declarations_link(caller):
    assert os.path.exists(caller)


def test_dropin_header_declares_reference_prototypes(tmp_path):
    """include/scc_dropin.h compiles in C++ and names the same six functions
    (same mangled symbols) as the reference-shaped caller."""
    src = tmp_path / "use_header.cpp"
    src.write_text('#include "scc_dropin.h"\n'
                   'int main(int c, char **v) {\n'
                   '  uint32_t E, V, *Fc, *Fr, *Bc, *Br, *Pr;\n'
                   '  loadFullGraph(v[1], &E, &V, &Fc, &Fr, &Bc, &Br, &Pr);\n'
                   '  printf("%u %u\\n", V, E); return atoi("0");\n}\n')
    exe = str(tmp_path / "use_header")
    r = subprocess.run(["g++", str(src), f"-I{os.path.join(ROOT, 'include')}", "-o", exe, f"-L{LIBDIR}",
                        "-lscc_amd", f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath,/opt/rocm/lib"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout


@pytest.mark.parametrize("name", ["sample0", "sample128", "sample2"])
def test_loadfullgraph_reference_layout(caller, known, fixture_path, name):
    path = fixture_path(known["graphs"][name]["file"])
    r = run_caller(caller, "load", path)
    assert r.returncode == 0, r.stdout + r.stderr
    o = orc.load_gtgraph(path)
    want = (f"V={o.V} E={o.E} Fr={_fnv(o.Fr)} Fc={_fnv(o.Fc)} Br={_fnv(o.Br)} Bc={_fnv(o.Bc)} "
            f"Pr=null")
    assert r.stdout.strip() == want


def test_loadfullgraph_reads_partition_from_env(caller, fixture_path):
    r = run_caller(caller, "load", fixture_path("sample128.gr"),
                   env={"SCC_PARTITION_FILE": fixture_path("MeTiSInput128.graph.part.10")})
    assert r.returncode == 0 and r.stdout.strip().endswith("Pr=set"), r.stdout


def test_loadfullgraph_missing_file(caller):
    r = run_caller(caller, "load", "/nonexistent.gr")
    assert r.returncode == 3
    assert r.stdout.splitlines() == ["Error Reading graph file", "empty V=0 E=0"]


def test_loadfullgraph_bad_partition(caller, fixture_path):
    r = run_caller(caller, "load", fixture_path("sample128.gr"),
                   env={"SCC_PARTITION_FILE": fixture_path("MeTiSInput0.graph.part.2")})
    assert r.returncode == 3 and r.stdout.startswith("Error Reading metis graph file")


def test_driver_failure_is_thrown(caller, fixture_path):
    """detectSCC without a partition vector fails on any machine (with or
    without a device), and the failure reaches the caller as a thrown
    const char * message."""
    r = run_caller(caller, "run", "d", fixture_path("sample0.gr"))
    assert r.returncode == 1
    assert r.stdout.startswith("caught: ") and len(r.stdout.strip()) > len("caught: ")


def test_c_abi_drivers_are_declared():
    L = scc.lib()
    for n in ["scc_vHong", "scc_wHong", "scc_vSlota", "scc_wSlota", "scc_detectSCC"]:
        assert hasattr(L, n), n


@pytest.mark.gpu
@pytest.mark.parametrize("algo,warp", [("g", 1), ("x", 64), ("h", 1), ("y", 8)])
def test_reference_caller_runs_drivers(caller, fixture_path, algo, warp):
    r = run_caller(caller, "run", algo, fixture_path("sample128.gr"), str(warp))
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split("\n")
    if algo == "g":
        assert lines[0].startswith("numberOf1Sccs: ")
        lines = lines[1:]
    assert lines[0] == "" and lines[1] == "Number Of Sccs : 13" and lines[2].startswith("Time : ")


@pytest.mark.gpu
def test_reference_caller_detect_with_partition(caller, fixture_path):
    r = run_caller(caller, "run", "d", fixture_path("sample128.gr"),
                   env={"SCC_PARTITION_FILE": fixture_path("MeTiSInput128.graph.part.10")})
    assert r.returncode == 0, r.stdout + r.stderr
    assert "\nNumber Of Sccs : 13\nTime : " in r.stdout
