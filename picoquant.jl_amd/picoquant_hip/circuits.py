"""Circuit generators as gate lists [(U, [qubits...]), ...] with 0-based qubits
and qiskit-ordered matrices (first target qubit = least significant bit).

  qft_gates            -- create_qft_circuit (src/algorithms.jl:14-32)
  grid_circuit_gates   -- random grid circuits: per cycle one of {sqrtX, sqrtY,
                          sqrtW} on every qubit (never the qubit's previous
                          gate), then CZ on one of four coupler patterns
"""
from __future__ import annotations

import cmath
import math

import numpy as np

SQ2 = 1.0 / math.sqrt(2.0)
I2 = np.eye(2, dtype=np.complex128)
X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
H = np.array([[1, 1], [1, -1]], dtype=np.complex128) * SQ2
W = (X + Y) * SQ2


def sqrt_involution(P):
    """Principal square root of a Hermitian involution P: eigenvalue 1 -> 1,
    -1 -> i, i.e. (1+i)/2 * I + (1-i)/2 * P."""
    return 0.5 * (1 + 1j) * I2 + 0.5 * (1 - 1j) * P


SQRT_X = sqrt_involution(X)
SQRT_Y = sqrt_involution(Y)
SQRT_W = sqrt_involution(W)
SQRT_GATES = (SQRT_X, SQRT_Y, SQRT_W)
CZ = np.diag([1, 1, 1, -1]).astype(np.complex128)
# CX on [control, target]: control is the first (least significant) qubit
CX = np.array([[1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0]], dtype=np.complex128)
SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=np.complex128)


def cphase(theta):
    """Controlled phase diag(1, 1, 1, e^{i theta}) (qiskit cu1; symmetric)."""
    return np.diag([1, 1, 1, cmath.exp(1j * theta)]).astype(np.complex128)


def qft_gates(n):
    """src/algorithms.jl:14-32: for i = 1..n: H on i-1, then cu1(pi/2^(j-i+1))
    on (j, i-1) for j = i..n-1; then SWAP(i-1, n-i) for i = 1..n/2.
    cu1(pi/2^(j-i+1)) is R_k = diag(1, e^{2 pi i / 2^k}) with k = j-i+2."""
    gates = []
    for i in range(1, n + 1):
        gates.append((H, [i - 1]))
        for j in range(i, n):
            gates.append((cphase(math.pi / 2 ** (j - i + 1)), [j, i - 1]))
    for i in range(1, n // 2 + 1):
        gates.append((SWAP, [i - 1, n - i]))
    return gates


GRID_PATTERNS = "ACBD"


def grid_couplers(rows, cols, pattern):
    """Couplers of one pattern on a rows x cols grid, qubit q = r*cols + c.
    A / B: horizontal pairs (c, c+1) with c even / odd; C / D: vertical pairs
    (r, r+1) with r even / odd."""
    out = []
    if pattern in ("A", "B"):
        par = 0 if pattern == "A" else 1
        for r in range(rows):
            for c in range(par, cols - 1, 2):
                out.append((r * cols + c, r * cols + c + 1))
    elif pattern in ("C", "D"):
        par = 0 if pattern == "C" else 1
        for r in range(par, rows - 1, 2):
            for c in range(cols):
                out.append((r * cols + c, (r + 1) * cols + c))
    else:
        raise ValueError(pattern)
    return out


def grid_circuit_gates(rows, cols, depth, seed):
    """Seeded random grid circuit of `depth` cycles.  Cycle t applies to every
    qubit q (in order) a gate drawn with numpy PCG64(seed): uniform over the
    three gates in the first cycle, afterwards uniform over the two gates that
    differ from the qubit's previous one; then CZ on the couplers of pattern
    'ACBD'[t % 4]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = rows * cols
    prev = [-1] * n
    gates = []
    for t in range(depth):
        for q in range(n):
            if prev[q] < 0:
                g = int(rng.integers(0, 3))
            else:
                choices = [x for x in range(3) if x != prev[q]]
                g = choices[int(rng.integers(0, 2))]
            prev[q] = g
            gates.append((SQRT_GATES[g], [q]))
        for a, b in grid_couplers(rows, cols, GRID_PATTERNS[t % 4]):
            gates.append((CZ, [a, b]))
    return gates


def seeded_bits(n, seed):
    """n seeded random bits (numpy PCG64)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return [int(b) for b in rng.integers(0, 2, size=n)]
