"""Direct-solve fixture for the 64x16x16 Hex8 cantilever (about 56k dofs).

K from the CPU oracle (the reference's assembly order), x = 0 face clamped,
uniform downward (-z) traction of total 1 on the x = 4 face as consistent
nodal loads, E = 1, nu = 0.3.  K_ff u = rhs is solved with scipy's SuperLU
(a sparse direct solve, the semantics of the reference's backslash) and a
compact summary is written: ||u_f||_2, u_f at 64 seeded sample indices, the
mean z-displacement of the loaded face, and the reaction force sum.

Run:  python tests/golden/make_cfg2_direct.py   (about 2 minutes)
"""
import json
import sys
from pathlib import Path

import numpy as np
import scipy.sparse.linalg as spla

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT / "oracle"), str(ROOT / "smearfem.jl_amd")]

import oracle as O  # noqa: E402
from smearfem.mesh import clamp_x0, face_load, structured_mesh  # noqa: E402


def main():
    m = structured_mesh((64, 16, 16), lengths=(4.0, 1.0, 1.0))
    K = O.assemble_system(0, m.NodeList, m.IEN, 3, "Q1", 3, m.ID, 1.0, 0.3, n_elem=m.n_elem)
    fixed = clamp_x0(m)
    f = face_load(m, 0, 1, 2, -1.0)
    Kff, rhs = O.eliminate(K, fixed, f)
    u = spla.spsolve(Kff.tocsc(), rhs)
    q = O.recombine(fixed, None, u)
    R = K @ q - f
    idx = np.sort(np.random.Generator(np.random.PCG64(2024)).choice(u.size, 64, replace=False))
    gi = m.node_grid_index()
    tip = np.repeat(gi[0] == 64, 3)
    out = {
        "generator": "tests/golden/make_cfg2_direct.py",
        "dims": [64, 16, 16], "lengths": [4.0, 1.0, 1.0], "E": 1.0, "nu": 0.3,
        "load": "uniform -z traction, total 1, on x=4 face", "clamp": "x=0 face, all dofs",
        "n_free": int(u.size), "nnz_K": int(K.nnz), "nnz_Kff": int(Kff.nnz),
        "norm_u": float(np.linalg.norm(u)),
        "sample_idx": idx.tolist(), "sample_u": u[idx].tolist(),
        "tip_mean_uz": float(q[tip][2::3].mean()),
        "reaction_sum": R[fixed == 1].reshape(-1, 3).sum(0).tolist(),
        "relres_direct": float(np.linalg.norm(Kff @ u - rhs) / np.linalg.norm(rhs)),
    }
    (Path(__file__).resolve().parent / "cfg2_direct.json").write_text(json.dumps(out, indent=1))
    print(json.dumps({k: out[k] for k in ("n_free", "norm_u", "tip_mean_uz", "relres_direct")}))


if __name__ == "__main__":
    main()
