"""Writes the golden fixtures of the MVP path (run from the repository root).

kats.json   known-answer vectors transcribed from the reference's own tests
            (each entry cites the file:line it comes from);
seeded.npz  seeded small operators (PCG64, seed 20260801) with outputs of the
            independent dense Kronecker oracle (oracle/dense.py), n <= 10.
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import dense  # noqa: E402


def kats() -> dict:
    return {
        "two_z_plus_identity": {
            "source": "tests/test_native_abi.py:84-121",
            "nqubits": 1,
            "terms": [["Z", [2.0, 0.0]], ["I", [1.0, 0.0]]],
            "dense": [[3.0, 0.0], [0.0, -1.0]],
            "state": [[1.0, 0.0], [2.0, 0.0]],
            "output": [[3.0, 0.0], [-2.0, 0.0]],
        },
        "x_on_plus": {
            "source": "crates/tencir-pauli-core/src/pauli_action.rs:188-206",
            "nqubits": 1,
            "terms": [["X", [1.0, 0.0]]],
            "state": [[0.7071067811865476, 0.0], [0.7071067811865476, 0.0]],
            "output": [[0.7071067811865476, 0.0], [0.7071067811865476, 0.0]],
        },
        "u1_basis": {
            "source": "tests/test_symmetry.py:159-183; crates/tencir-pauli-core/src/tests.rs:377-383",
            "cases": [[0, 0, [0]], [3, 0, [0]], [3, 1, [1, 2, 4]], [3, 2, [3, 5, 6]], [3, 3, [7]]],
            "rank_3_1_of_2": 1,
            "unrank_3_1_of_2": 4,
        },
        "u1_hopping": {
            "source": "crates/tencir-pauli-core/src/tests.rs:384-418",
            "nqubits": 3,
            "particle_number": 1,
            "terms": [["XXI", [1.0, 0.0]], ["YYI", [1.0, 0.0]]],
            "dense_flat_nonzero": {"5": 2.0, "7": 2.0},
            "state": [3.0, 5.0, 7.0],
            "output": [0.0, 14.0, 10.0],
        },
        "u1_leakage": {
            "source": "tests/test_symmetry.py:229-235",
            "nqubits": 3,
            "particle_number": 1,
            "conserving": [["XXI", 1.0], ["YYI", 1.0]],
            "leaking": [["XII", 1.0]],
        },
        "charge_three_modes_one_particle": {
            "source": "tests/test_charge.py:86-101",
            "nqubits": 3,
            "population": 1,
            "order": [[0, 0, 1], [0, 1, 0], [1, 0, 0]],
        },
        "charge_spin_and_number": {
            "source": "tests/test_charge.py:263-270",
            "levels": [[[0, 1], [0, 1]], [[0, 1], [0, -1]]],
            "targets": [1, 1],
            "dimension": 1,
            "order": [[1, 0]],
        },
        "charge_free_spectators": {
            "source": "tests/test_charge.py:855-861",
            "levels": [[[0, 1], [0, 0], [0, 0], [0, 0]]],
            "targets": [0],
            "dimension": 8,
        },
        "charge_wide_one_particle": {
            "source": "tests/test_charge.py:103-115",
            "modes": [62, 63],
            "number_dimension": "modes",
            "marked_dimension": 1,
            "marked_first": "mode 0 occupied, every other mode empty",
        },
        "charge_qubit_hopping": {
            "source": "tests/test_charge.py:310-345",
            "nqubits": 2,
            "terms": [["XX", [1.0, 0.0]], ["YY", [1.0, 0.0]]],
            "levels": [[[0, 1], [0, 1]]],
            "targets": [1],
            "dense": [[[0.0, 0.0], [2.0, 0.0]], [[2.0, 0.0], [0.0, 0.0]]],
            "state": [[0.5, -0.2], [-0.3, 0.7]],
            "transition_count": 2,
        },
        "charge_aggregated_cancellation": {
            "source": "tests/test_charge.py:369-385",
            "nqubits": 2,
            "terms": [["XX", [1.0, 0.0]], ["YY", [1.0, 0.0]]],
            "levels": [[[0, 1], [0, 1]]],
            "targets": [0],
            "dense": [[[0.0, 0.0]]],
            "transition_count": 0,
        },
        "charge_two_modes_zero_charge": {
            "source": "tests/test_execution_invariants.py:76-93",
            "order": [[0, 0], [0, 1], [1, 0], [1, 1]],
        },
        "first_last_qubit": {
            "source": "tests/test_hamiltonian.py:71-77",
            "nqubits": 3,
            "note": "code on qubit 0 acts on the most significant index bit",
        },
    }


def make_operator(nqubits: int):
    """tests/test_hamiltonian.py:16-25: codes (index + qubit) % 4, coefficient index+1 - i index/3."""
    count = min(10, 4 ** nqubits)
    codes = np.asarray([[(i + q) % 4 for q in range(nqubits)] for i in range(count)], dtype=np.uint8)
    coef = np.asarray([complex(i + 1, -i / 3) for i in range(count)], dtype=np.complex128)
    return codes.reshape(count, nqubits), coef


def mvp_test_operator(nqubits: int, term_count: int):
    """crates/tencir-pauli-core/src/tests.rs:25-37: code bit q of the term index, coef (i+1, -0.25 i)."""
    codes = np.asarray([[1 if (m >> q) & 1 else 0 for q in range(nqubits)] for m in range(term_count)],
                       dtype=np.uint8)
    coef = np.asarray([complex(i + 1.0, -i * 0.25) for i in range(term_count)], dtype=np.complex128)
    return codes.reshape(term_count, nqubits), coef


def seeded() -> dict:
    out = {}
    for n in (0, 1, 2, 5):
        codes, coef = make_operator(n)
        state = np.arange(1 << n, dtype=np.float64) + 1j * np.arange(1 << n)
        out[f"make_operator_{n}_codes"] = codes
        out[f"make_operator_{n}_coef"] = coef
        out[f"make_operator_{n}_state"] = state
        out[f"make_operator_{n}_output"] = dense.dense_operator(n, codes, coef) @ state
    codes, coef = mvp_test_operator(10, 64)
    state = np.asarray([complex(i * 0.5, 1.0 - i * 0.125) for i in range(1 << 10)])
    out["mvp_test_operator_codes"] = codes
    out["mvp_test_operator_coef"] = coef
    out["mvp_test_operator_state"] = state
    out["mvp_test_operator_output"] = dense.dense_operator(10, codes, coef) @ state
    # tests/test_hamiltonian.py:126-147 pattern: 24 random complex terms per n, seed 20260801
    rng = np.random.default_rng(20260801)
    for n in range(1, 9):
        codes = rng.integers(0, 4, size=(24, n)).astype(np.uint8)
        coef = rng.normal(size=24) + 1j * rng.normal(size=24)
        state = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        out[f"random_{n}_codes"] = codes
        out[f"random_{n}_coef"] = coef
        out[f"random_{n}_state"] = state
        out[f"random_{n}_output"] = dense.dense_operator(n, codes, coef) @ state
    return out


def main() -> None:
    with open(os.path.join(HERE, "kats.json"), "w") as fh:
        json.dump(kats(), fh, indent=1, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "seeded.npz"), **seeded())


if __name__ == "__main__":
    main()
