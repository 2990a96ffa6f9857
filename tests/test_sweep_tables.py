"""CPU check of the dataflow sweep tables that the batch refinement and the
one-workgroup solve run on (build_df in glu-hcu_amd/csrc/glu_plan.hip):
tests/cpp/sweep_tables_check.cpp emulates the kernel's loads and task order
on the host — every load in bounds, every dependency finished before use,
every row once, result equal to the sequential solve bit for bit — for MNA
and random patterns, with and without long lane rows."""

import os
import shutil
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

from helpers import mna_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "glu-hcu_amd")
BUILD = os.path.join(PKG, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    objs = [os.path.join(BUILD, f) for f in sorted(os.listdir(BUILD))
            if f.startswith("host_") or f == "glu_kernels.o"] if os.path.isdir(BUILD) else []
    if not os.path.exists(HIPCC) or not any(o.endswith("glu_kernels.o") for o in objs):
        pytest.skip("hipcc or the library's objects are missing (run __graft_entry__.build())")
    d = tmp_path_factory.mktemp("sweep")
    obj, exe = str(d / "check.o"), str(d / "check")
    flags = ["-O2", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950", "-w",
             "-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "csrc")]
    subprocess.check_call([HIPCC, *flags, "-c", os.path.join(ROOT, "tests", "cpp", "sweep_tables_check.cpp"),
                           "-o", obj])
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", obj, *objs, "-o", exe])
    yield exe
    shutil.rmtree(d, ignore_errors=True)


def _dump(path, an):
    with open(path, "wb") as f:
        for a in (np.array([an.n], np.uint32), an.csr_r_ptr, an.csr_c_idx, an.csr_diag_ptr):
            a = np.ascontiguousarray(a, dtype=np.uint32)
            f.write(np.uint64(a.size).tobytes())
            f.write(a.tobytes())


def _random(t):
    rng = np.random.default_rng(2000 + t)
    n = int(rng.integers(5, 2500))
    A = sp.random(n, n, density=float(rng.uniform(0.002, 0.03)), random_state=t, format="csc")
    A = A + sp.csc_matrix((rng.uniform(0.5, 2, n), (rng.permutation(n), np.arange(n))), shape=(n, n))
    A = sp.csc_matrix(A)
    A.sum_duplicates()
    A.sort_indices()
    return A.data, A.indices.astype(np.int32), A.indptr.astype(np.int32), n


@pytest.mark.parametrize("case", ["mna600", "mna2k", "small", "add32", "rand0", "rand1", "rand2"])
@pytest.mark.parametrize("long_lanes", [0, 1])
def test_sweep_tables(checker, tmp_path, case, long_lanes):
    from pyglu.analysis import Analysis
    mat = _random(int(case[4:])) if case.startswith("rand") else mna_case(case)
    an = Analysis(*mat)
    path = tmp_path / "pattern.bin"
    _dump(path, an)
    r = subprocess.run([checker, str(path), str(long_lanes)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bit for bit" in r.stdout
