"""Train the learned Davidson preconditioner on 16-20-qubit lattices and save it.

The training family is the periodic spin-1/2 XXZ model of the 25-qubit target
(Jxy = 1, Jz ~ U[0.8, 1.2] per bond, field h ~ U[-0.2, 0.2] on Z): 4x4, 3x6 and
4x5 tori with several seeds, in their S_z = 0 and S_z = +-1 sectors.  Every
lattice contributes snapshots of plain Jacobi Davidson (see
``learned.jacobi_snapshots``); the loss is the one-step Ritz gain of the
learned correction, and training runs entirely in the library's HIP kernels
(MFMA forward and backward, loss gradient with an MVP, Adam).

    python -m tencirpauli_amd.train_preconditioner [--epochs 30] [--out PATH]
"""

from __future__ import annotations

import argparse
import os
import time
from typing import List, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "preconditioner_mlp.npz")

# (rows, cols, seed, particle_number offset from half filling)
TRAINING_LATTICES: Sequence[Tuple[int, int, int, int]] = (
    (4, 4, 101, 0), (4, 4, 102, 1), (4, 4, 103, 0),
    (3, 6, 201, 0), (3, 6, 202, -1),
    (4, 5, 301, 0), (4, 5, 302, 1),
)
SNAPSHOT_ITERATIONS = (1, 2, 3, 4, 6, 8, 11, 15, 20, 26, 34, 44)


def restricted_lattice(rows: int, cols: int, seed: int, offset: int = 0):
    import tencirpauli_amd as tcp
    from tencirpauli_amd import lattice

    codes, coef = lattice.heisenberg_periodic(rows, cols, seed=seed)
    op = tcp.PauliOperator.from_code_arrays(codes, coef)
    n = rows * cols
    return op.restrict_charge(tcp.U1Sector(n, n // 2 + offset))


def build_training_set(mlp, lattices=TRAINING_LATTICES, iterations=SNAPSHOT_ITERATIONS) -> List:
    from tencirpauli_amd import learned

    snaps = []
    for rows, cols, seed, offset in lattices:
        snaps += learned.jacobi_snapshots(restricted_lattice(rows, cols, seed, offset), mlp, iterations)
    return snaps


def train_default(epochs: int = 30, lr: float = 2e-3, seed: int = 20260805, verbose: bool = False):
    """Returns (mlp, loss history, seconds)."""
    from tencirpauli_amd import learned

    t0 = time.perf_counter()
    mlp = learned.MLP(seed=seed)
    snaps = build_training_set(mlp)
    history = learned.train(mlp, snaps, epochs=epochs, lr=lr, seed=seed)
    if verbose:
        print(f"{len(snaps)} snapshots; loss {history[0]:.4f} -> {history[-1]:.4f}", flush=True)
    return mlp, history, time.perf_counter() - t0


def save(mlp, path: str = DEFAULT_WEIGHTS, **meta) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=mlp.get_parameters(), **{k: np.asarray(v) for k, v in meta.items()})


def load(path: str = DEFAULT_WEIGHTS):
    from tencirpauli_amd import learned

    with np.load(path, allow_pickle=False) as data:
        params = data["params"]
    mlp = learned.MLP()
    mlp.set_parameters(params)
    return mlp


def main() -> None:
    parser = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    parser.add_argument("--epochs", type=int, default=30)
    parser.add_argument("--lr", type=float, default=2e-3)
    parser.add_argument("--out", default=DEFAULT_WEIGHTS)
    args = parser.parse_args()
    mlp, history, seconds = train_default(args.epochs, args.lr, verbose=True)
    save(mlp, args.out, history=history, seconds=seconds, epochs=args.epochs, lr=args.lr)
    print(f"saved {args.out} after {seconds:.1f} s", flush=True)


if __name__ == "__main__":
    main()
