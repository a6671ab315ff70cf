"""Direct-solve fixtures for the largest batch problems and large served requests.

Two sets, both solved on the CPU from the oracle's K (the reference's assembly
order) with scipy's SuperLU (a sparse direct solve, the semantics of the
reference's backslash; minimum-degree ordering on K + K', symmetric mode):

* ``batch``: the three largest of the 512 uneven Hex8 cantilevers of the
  batch workload (``uneven_hex8_problems(512, seed=4)``: dims U{4..32}^3,
  E log-U[0.1, 10] and nu U[0.2, 0.4] per problem, x = 0 clamped, unit
  downward end load), about 89k-95k free dofs each;
* ``serve``: requests of the serving workload (``random_request``: Q4 or
  Hex8 cantilever, per-element E log-U[0.1, 10], nu 0.3, N(0, 1) loads on the
  loaded face) drawn with DOF counts of 1e5 and more.

A full solution is too large for a fixture, so each problem keeps a compact
summary that a GPU solution is compared with: ||u_f||_2, u_f at 128 seeded
sample indices, 16 projections g_k . u_f on seeded N(0, 1) vectors g_k (for
an error e, g_k . e has standard deviation ||e||_2, so the projections bound
the whole error vector, not only the samples), the reaction sum and the
direct solve's relative residual.

Run:  python tests/golden/make_large_direct.py   (about 15 minutes on 8 cores)
"""
import json
import sys
import time
from pathlib import Path

import numpy as np
import scipy.sparse.linalg as spla

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT / "oracle"), str(ROOT / "smearfem.jl_amd")]

import oracle as O  # noqa: E402
from smearfem.batch import uneven_hex8_problems  # noqa: E402
from smearfem.serve import random_request  # noqa: E402

N_SAMPLE = 128
N_PROJ = 16
# (seed, dof_lo, dof_hi, p3d) of each large request: one generator per request
SERVE_SPECS = [(101, 1e5, 1e6, 0.0), (102, 3e5, 1e6, 0.0), (103, 1e5, 1.3e5, 1.0)]


def projections(u: np.ndarray, seed: int) -> np.ndarray:
    """g_k . u for N_PROJ vectors g_k ~ N(0, 1) of PCG64(seed), k in order."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.array([rng.standard_normal(u.size) @ u for _ in range(N_PROJ)])


def summary(p, seed: int) -> dict:
    m = p.mesh
    K = O.assemble_system(0, m.NodeList, m.IEN, m.ndim, "Q1", m.ndim, m.ID, p.young, p.nu[0], n_elem=m.n_elem)
    Kff, rhs = O.eliminate(K, p.fixed, p.f)
    A = Kff.tocsc()
    lu = spla.splu(A, permc_spec="MMD_AT_PLUS_A", options=dict(SymmetricMode=True))
    u = lu.solve(rhs)
    q = O.recombine(p.fixed, None, u)
    R = K @ q - p.f
    idx = np.sort(np.random.Generator(np.random.PCG64(seed)).choice(u.size, N_SAMPLE, replace=False))
    return {
        "dims": list(m.dims), "n_free": int(u.size), "nnz_K": int(K.nnz),
        "norm_u": float(np.linalg.norm(u)), "sample_idx": idx.tolist(), "sample_u": u[idx].tolist(),
        "proj_seed": seed, "proj": projections(u, seed).tolist(),
        "reaction_sum": R[p.fixed == 1].reshape(-1, m.ndim).sum(0).tolist(),
        "relres_direct": float(np.linalg.norm(A @ u - rhs) / np.linalg.norm(rhs)),
    }


def main():
    out = {"generator": "tests/golden/make_large_direct.py", "n_sample": N_SAMPLE, "n_proj": N_PROJ,
           "batch": {"nprob": 512, "seed": 4, "problems": {}}, "serve": []}
    probs = uneven_hex8_problems(512, 4)
    nfree = np.array([int((p.fixed == 0).sum()) for p in probs])
    for i in np.argsort(-nfree, kind="stable")[:3]:
        t0 = time.time()
        s = summary(probs[i], 7000 + int(i))
        out["batch"]["problems"][str(int(i))] = s
        print(f"batch problem {i} {s['dims']} n_free {s['n_free']} relres {s['relres_direct']:.1e} "
              f"{time.time() - t0:.0f}s", flush=True)
    for seed, lo, hi, p3d in SERVE_SPECS:
        t0 = time.time()
        p = random_request(np.random.Generator(np.random.PCG64(seed)), lo, hi, p3d)
        s = summary(p, 8000 + seed)
        s["spec"] = [seed, lo, hi, p3d]
        out["serve"].append(s)
        print(f"request seed {seed} {s['dims']} n_free {s['n_free']} relres {s['relres_direct']:.1e} "
              f"{time.time() - t0:.0f}s", flush=True)
    (Path(__file__).resolve().parent / "large_direct.json").write_text(json.dumps(out))


if __name__ == "__main__":
    main()
