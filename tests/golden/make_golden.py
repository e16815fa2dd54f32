"""Write the golden amplitude fixtures from the oracle's brute-force state vector.

    python tests/golden/make_golden.py

grid5x5_d12.json: the 5x5 depth-12 grid circuit (seed 2026) -- amplitudes
<x|C|0...0> for 8 seeded bitstrings, plus the 32 amplitudes of the batch that
leaves the last grid column open (qubits 5, 10, 15, 20, 25; other bits from
the first seeded bitstring).  The 2^25 state vector is too slow to rebuild in
every test run, so the values are committed.
"""
import json
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from oracle import pq_oracle as O  # noqa: E402

ROWS, COLS, DEPTH, SEED = 5, 5, 12, 2026


def main():
    n = ROWS * COLS
    gates = O.grid_circuit_gates(ROWS, COLS, DEPTH, SEED)
    psi = O.statevector(n, gates)
    rng = np.random.Generator(np.random.PCG64(SEED + 1))
    bitstrings = [[int(b) for b in rng.integers(0, 2, n)] for _ in range(8)]
    amps = []
    for bits in bitstrings:
        idx = sum(b << q for q, b in enumerate(bits))
        amps.append([float(psi[idx].real), float(psi[idx].imag)])
    open_q = [r * COLS + (COLS - 1) for r in range(ROWS)]      # 0-based
    base = list(bitstrings[0])
    batch = []
    for j in range(2 ** len(open_q)):
        bits = list(base)
        for t, q in enumerate(open_q):
            bits[q] = (j >> t) & 1
        idx = sum(b << q for q, b in enumerate(bits))
        batch.append([float(psi[idx].real), float(psi[idx].imag)])
    out = {
        "rows": ROWS, "cols": COLS, "depth": DEPTH, "seed": SEED,
        "norm": float(np.vdot(psi, psi).real),
        "bitstrings": ["".join(map(str, b)) for b in bitstrings],
        "amplitudes": amps,
        "open_qubits_1based": [q + 1 for q in open_q],
        "batch_base": "".join(map(str, base)),
        "batch_amplitudes": batch,
    }
    (HERE / "grid5x5_d12.json").write_text(json.dumps(out, indent=1))
    print("wrote", HERE / "grid5x5_d12.json")


if __name__ == "__main__":
    main()
