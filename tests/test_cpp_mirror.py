"""The C++ host mirror (include/symbolic.h, numeric.h, preprocess.h) and the
lu_cmd front end, built against libglu_hip.so."""

import os
import subprocess

import numpy as np
import pytest

from helpers import add32

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "glu-hcu_amd")


def write_mtx(path, data, indices, indptr, n, banner=False):
    """Column-sorted 1-based triplets, the format of the reference's
    src/add32.mtx (no banner) and its reader."""
    cols = np.repeat(np.arange(n), np.diff(indptr))
    with open(path, "w") as f:
        if banner:
            f.write("%%MatrixMarket matrix coordinate real general\n% comment\n")
        f.write(f"{n} {n} {data.size}\n")
        for r, c, v in zip(indices, cols, data):
            f.write(f"{r + 1} {c + 1} {float(v):.17g}\n")


@pytest.fixture(scope="module")
def mirror_exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("cpp")
    exe = str(d / "mirror_check")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "cpp", "mirror_check.cpp"), "-L", PKG, "-lglu_hip",
                           f"-Wl,-rpath,{PKG}", "-o", exe])
    return exe


def test_cpp_mirror_symbolic_matches_c_abi(mirror_exe, tmp_path):
    d, i, p, n = add32()
    f = tmp_path / "add32.mtx"
    write_mtx(f, d, i, p, n, banner=True)
    r = subprocess.run([mirror_exe, str(f)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("OK n=4960 nnz=23884")


def test_triplet_reader(tmp_path):
    import ctypes as C
    from pyglu import _capi
    L = _capi.lib()
    L.glu_read_triplet_column.argtypes = [C.c_char_p] + [C.c_void_p] * 5
    d, i, p, n = add32()
    f = tmp_path / "a.mtx"
    write_mtx(f, d, i, p, n)
    nn, nz = C.c_uint(), C.c_uint()
    ax, ai, ap = C.POINTER(C.c_double)(), C.POINTER(C.c_uint)(), C.POINTER(C.c_uint)()
    assert L.glu_read_triplet_column(str(f).encode(), C.byref(nn), C.byref(nz), C.byref(ax), C.byref(ai),
                                     C.byref(ap)) == 0
    assert nn.value == n and nz.value == d.size
    assert np.array_equal(np.ctypeslib.as_array(ap, (n + 1,)), p)
    assert np.array_equal(np.ctypeslib.as_array(ai, (d.size,)), i)
    assert np.array_equal(np.ctypeslib.as_array(ax, (d.size,)), d)
    # missing file / unsorted columns
    assert L.glu_read_triplet_column(str(tmp_path / "none.mtx").encode(), C.byref(nn), C.byref(nz), C.byref(ax),
                                     C.byref(ai), C.byref(ap)) == -4
    bad = tmp_path / "bad.mtx"
    bad.write_text("2 2 2\n1 2 1.0\n2 1 1.0\n")
    assert L.glu_read_triplet_column(str(bad).encode(), C.byref(nn), C.byref(nz), C.byref(ax), C.byref(ai),
                                     C.byref(ap)) == -7


@pytest.mark.gpu
def test_lu_cmd_on_add32(tmp_path):
    """The CLI flow of src/lu_cmd.cpp on add32: preprocess, symbolic,
    LUonDevice (with the ABFT check around it, -a), solve with b = 1, x.dat —
    x equals the oracle's FP32 solve with the oracle's FP32 factors (to the 6
    digits x.dat holds), and the factors pass the ABFT check, as the
    oracle's restatement of src/symbolic.cc:172-197 says they do."""
    from oracle import oracle as O
    from pyglu.analysis import Analysis
    d, i, p, n = add32()
    f = tmp_path / "add32.mtx"
    write_mtx(f, d, i, p, n)
    r = subprocess.run([os.path.join(PKG, "lu_cmd"), "-i", str(f), "-a"], capture_output=True, text=True,
                       cwd=tmp_path, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    for key in ("Symbolic nonzero:", "Number of levels:", "Device 0:", "Total GPU time:",
                "Results passed ABFT check."):
        assert key in r.stdout
    x = np.loadtxt(tmp_path / "x.dat")
    an = Analysis(d, i, p, n)
    pat = O.SymPattern(an)
    lu, _ = O.factor_f32(pat, an.val)
    assert O.abft_check(pat, lu, O.abft_cca(pat, an.val))[0] == -1
    want = O.solve_f32(pat, lu, an.col_perm, an.row_perm_inv, an.col_scale_perm, an.row_scale,
                       np.ones(n, np.float32))
    assert np.allclose(x, want, rtol=1e-5, atol=1e-6 * np.abs(want).max())
