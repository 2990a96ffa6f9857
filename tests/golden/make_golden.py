"""Writes tests/golden/analysis_sha256.json from the reference's own host
code (NICSLU analysis + my_DumpA + Symbolic_Matrix, compiled from the
reference sources into oracle/_ref by oracle/Makefile).

For every input matrix it records the SHA-256 of each array the reference
produces (permutations, scale factors, the permuted/scaled matrix, the LU
pattern, the predictLU values and the level sets) and of the reference's
FP32 solve output for the oracle's FP32 factors and a seeded right-hand
side, and of the refactor values (NICSLU's _I_NicsLU_MC64ScaleForRefact,
my_DumpA, predictLU) of a drifted value set.  Run from the repository root:
python tests/golden/make_golden.py
"""

from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "glu-hcu_amd"), ROOT, os.path.dirname(HERE)]

from oracle import oracle as O  # noqa: E402
from oracle import ref  # noqa: E402

from helpers import add32  # noqa: E402
from pyglu import mna  # noqa: E402

ARRAYS = ["row_perm", "row_perm_inv", "col_perm", "col_perm_inv", "row_scale", "col_scale_perm",
          "dump_ax", "dump_ai", "dump_ap", "sym_c_ptr", "sym_r_idx", "l_col_ptr", "csr_r_ptr",
          "csr_c_idx", "csr_diag_ptr", "val", "level_ptr", "level_idx"]
REFACT_SEED = 9000   # drift of the refactor value set: data * (1 + U[-0.1, 0.1])


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def inputs():
    yield "add32", add32()
    yield "mna_small", mna.mna_csc(mna.SPEC_SMALL, mna.SEED_SMALL)
    yield "mna_2k", mna.mna_csc(mna.MNASpec(n_nodes=1980, n_vsrc=20, n_rails=3, rail_deg=120), 7)
    rng = np.random.default_rng(2024)
    for t in range(8):
        n = int(rng.integers(20, 300))
        A = sp.random(n, n, density=float(rng.uniform(0.01, 0.06)), random_state=int(rng.integers(1 << 30)),
                      format="csc")
        if t % 2 == 0:
            A.data[:] = 1.0  # equal magnitudes: MC64/AMD tie-breaks decide everything
        A = A + sp.csc_matrix((np.ones(n), (rng.permutation(n), np.arange(n))), shape=(n, n))
        A = sp.csc_matrix(A)
        A.sum_duplicates()
        A.sort_indices()
        yield f"random{t}", (A.data.copy(), A.indices.astype(np.int32), A.indptr.astype(np.int32), n)


def main():
    out = {}
    for name, (d, i, p, n) in inputs():
        r = ref.analyze(d, i, p, n)
        rec = {"n": int(n), "nnz": int(d.size), "nnz_lu": int(r.nnz_lu), "num_lev": int(r.num_lev)}
        rec["inputs_sha256"] = sha(np.concatenate([d.view(np.uint8), i.astype(np.int32).view(np.uint8),
                                                   p.astype(np.int32).view(np.uint8)]))
        for a in ARRAYS:
            rec[a] = sha(getattr(r, a))
        pat = O.SymPattern(r)
        lu32, _ = O.factor_f32(pat, r.val)
        rhs = mna.rhs(n, 77).astype(np.float32)
        rec["oracle_lu32"] = sha(lu32)
        rec["ref_solve_f32"] = sha(r.solve(lu32, rhs))
        # refactor values of a drifted value set (NICSLU's refactor scaling,
        # my_DumpA, predictLU)
        dump, val = r.refact_values(mna.perturb_values(d, REFACT_SEED))
        rec["refact_dump_ax"] = sha(dump)
        rec["refact_val"] = sha(val)
        out[name] = rec
        print(name, n, r.nnz_lu, r.num_lev)
    with open(os.path.join(HERE, "analysis_sha256.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
