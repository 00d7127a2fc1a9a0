"""Writes ground_energy_xxz5x5.json: the FP64 ground energy of the 5 x 5 periodic
Heisenberg XXZ torus (25 qubits, Jxy = 1, Jz ~ U[0.8, 1.2] per bond, h_i ~
U[-0.2, 0.2] on Z_i, seed 20260801) in the S_z = +1/2 sector (U1Sector(25, 12),
dimension 5,200,300).

The energy comes from SciPy's ARPACK ``eigsh(k=1, which="SA", tol=1e-12)`` over
a LinearOperator whose product is the oracle's restatement of the reference's
U(1) lazy scatter (core/sector.rs:579-642, ``oracle.u1_lazy_apply``) -- the
pattern of the reference's tests/test_scipy_linear_operator.py:104-124, which
pins eigenvalues of a restricted plan against an independent operator.  No GPU
code takes part.  The operator is real symmetric in this basis, so the product
runs on complex vectors whose imaginary part stays zero.

Each product splits the source ranks into contiguous ranges, one per worker
process, and sums the partial scatters (the oracle's scatter is serial; the
split only changes the FP64 summation order, far below the fixture's use at
1e-8).  The run also records the explicit residual ||H v - E v|| of the
returned vector and the second-lowest Ritz value of a k = 2 solve, so that a
test can tell how well separated the level is.  With six workers the whole
script takes 15-30 minutes.  Run from the repository root:

    python tests/golden/make_ground_energy.py
"""

from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from tencirpauli_amd import lattice  # noqa: E402

OUT = os.path.join(HERE, "ground_energy_xxz5x5.json")
SEED = 20260801


_WORK = {}


def _init(shm_name: str, dim: int, codes: np.ndarray, coef: np.ndarray) -> None:
    from multiprocessing import shared_memory

    shm = shared_memory.SharedMemory(name=shm_name)
    _WORK["shm"] = shm
    _WORK["x"] = np.ndarray((dim,), dtype=np.complex128, buffer=shm.buf)
    _WORK["terms"] = (codes, coef)


def _partial(rng):  # type: ignore[no-untyped-def]
    codes, coef = _WORK["terms"]
    return oracle.u1_lazy_apply(25, 12, codes, coef, _WORK["x"], sources=rng)


def main(workers: int = 6) -> None:
    from concurrent.futures import ProcessPoolExecutor
    from multiprocessing import shared_memory

    from scipy.sparse.linalg import LinearOperator, eigsh

    codes, coef = lattice.heisenberg_periodic(5, 5, seed=SEED)
    # canonical terms, as every plan of the library sees them
    codes, coef = oracle.canonicalize(25, codes, coef)
    dim = oracle.u1_dimension(25, 12)
    shm = shared_memory.SharedMemory(create=True, size=16 * dim)
    shared_x = np.ndarray((dim,), dtype=np.complex128, buffer=shm.buf)
    bounds = np.linspace(0, dim, workers + 1).astype(np.int64)
    ranges = [(int(bounds[i]), int(bounds[i + 1])) for i in range(workers)]
    pool = ProcessPoolExecutor(workers, initializer=_init, initargs=(shm.name, dim, codes, coef))
    calls = [0]

    def apply(x: np.ndarray) -> np.ndarray:
        shared_x[:] = x
        y = np.zeros(dim, dtype=np.complex128)
        for part in pool.map(_partial, ranges):
            y += part
        return y

    def matvec(x: np.ndarray) -> np.ndarray:
        calls[0] += 1
        y = apply(np.asarray(x, dtype=np.float64).reshape(-1))
        if calls[0] % 20 == 0:
            print(f"  {calls[0]} products", flush=True)
        return y.real.copy()

    linear = LinearOperator((dim, dim), matvec=matvec, dtype=np.float64)
    v0 = lattice.gaussian_state(dim, SEED + 4).real.copy()
    t0 = time.perf_counter()
    w, v = eigsh(linear, k=2, which="SA", v0=v0, tol=1e-12, ncv=24, maxiter=2000)
    seconds = time.perf_counter() - t0
    order = np.argsort(w)
    w, v = w[order], v[:, order]
    x = v[:, 0]
    hx = apply(x).real
    pool.shutdown()
    shm.close()
    shm.unlink()
    e0 = float(x @ hx / (x @ x))
    residual = float(np.linalg.norm(hx - w[0] * x) / np.linalg.norm(x))
    record = {
        "workload": "5x5 periodic XXZ, Jxy=1, Jz~U[0.8,1.2], h~U[-0.2,0.2], seed 20260801, "
                    "U1Sector(25, 12) (S_z=+1/2)",
        "dimension": int(dim),
        "terms": int(len(coef)),
        "energy": float(w[0]),
        "rayleigh_quotient": e0,
        "second_energy": float(w[1]),
        "residual_norm": residual,
        "method": "scipy.sparse.linalg.eigsh(k=2, which='SA', tol=1e-12, ncv=24) over oracle.u1_lazy_apply "
                  "(restatement of core/sector.rs:579-642)",
        "products": calls[0],
        "seconds": round(seconds, 1),
    }
    with open(OUT, "w") as fh:
        json.dump(record, fh, indent=1)
        fh.write("\n")
    print(json.dumps(record, indent=1))


if __name__ == "__main__":
    main()
