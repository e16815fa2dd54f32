"""OpenQASM 2.0 ingest without qiskit (src/layer3.jl:459-574).

The reference reads QASM through qiskit (`load_qasm_as_circuit`) and turns
the circuit into a network with qiskit's operator matrices
(`convert_qiskit_circ_to_network`).  This module parses the qelib1 subset
that circuits use -- one `qreg`, any `creg`, the standard one-, two- and
three-qubit gates with arithmetic parameters, `barrier` and `measure`
(skipped) -- into the same gate list the circuit generators produce:
[(U, [qubits, 0-based])] with qiskit's little-endian matrices (first target
qubit = least significant bit).  Custom `gate` definitions, `if`, `reset`
and `opaque` raise QasmError.
"""
from __future__ import annotations

import cmath
import math
import pathlib
import re
from dataclasses import dataclass, field

import numpy as np

from .network import network_from_gates


class QasmError(ValueError):
    pass


@dataclass
class QasmCircuit:
    """The parsed circuit: qubit count and gate list [(U, qubits)]."""
    n_qubits: int
    gates: list = field(default_factory=list)
    names: list = field(default_factory=list)


# --------------------------------------------------------------------------- #
# Gate matrices (qelib1 definitions; U[row = output bits, col = input bits])
# --------------------------------------------------------------------------- #

def _u3(theta, phi, lam):
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return np.array([[c, -cmath.exp(1j * lam) * s],
                     [cmath.exp(1j * phi) * s, cmath.exp(1j * (phi + lam)) * c]], dtype=np.complex128)


def _controlled(U):
    """Control on the first listed qubit (least significant bit), U on the
    second."""
    d = U.shape[0]
    out = np.eye(2 * d, dtype=np.complex128)
    # basis index = control + 2 * target_bits
    for r in range(d):
        for c in range(d):
            out[1 + 2 * r, 1 + 2 * c] = U[r, c]
    return out


_I = np.eye(2, dtype=np.complex128)
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.diag([1, -1]).astype(np.complex128)
_H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / math.sqrt(2)


def _rx(t):
    return np.array([[math.cos(t / 2), -1j * math.sin(t / 2)], [-1j * math.sin(t / 2), math.cos(t / 2)]])


def _ry(t):
    return np.array([[math.cos(t / 2), -math.sin(t / 2)], [math.sin(t / 2), math.cos(t / 2)]], dtype=np.complex128)


def _rz(t):
    return np.diag([cmath.exp(-0.5j * t), cmath.exp(0.5j * t)])


def _phase(lam):
    return np.diag([1, cmath.exp(1j * lam)])


def _swap():
    return np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=np.complex128)


def _ccx():
    U = np.eye(8, dtype=np.complex128)
    # controls = bits 0 and 1, target = bit 2: flip bit 2 when bits 0, 1 are set
    U[[3, 7], :] = U[[7, 3], :]
    return U


# name -> (number of parameters, number of qubits, matrix builder)
GATES = {
    "id": (0, 1, lambda: _I), "x": (0, 1, lambda: _X), "y": (0, 1, lambda: _Y), "z": (0, 1, lambda: _Z),
    "h": (0, 1, lambda: _H), "s": (0, 1, lambda: _phase(math.pi / 2)), "sdg": (0, 1, lambda: _phase(-math.pi / 2)),
    "t": (0, 1, lambda: _phase(math.pi / 4)), "tdg": (0, 1, lambda: _phase(-math.pi / 4)),
    "sx": (0, 1, lambda: 0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]])),
    "rx": (1, 1, _rx), "ry": (1, 1, _ry), "rz": (1, 1, _rz),
    "u1": (1, 1, _phase), "p": (1, 1, _phase),
    "u2": (2, 1, lambda phi, lam: _u3(math.pi / 2, phi, lam)),
    "u3": (3, 1, _u3), "u": (3, 1, _u3), "U": (3, 1, _u3),
    "cx": (0, 2, lambda: _controlled(_X)), "CX": (0, 2, lambda: _controlled(_X)),
    "cy": (0, 2, lambda: _controlled(_Y)), "cz": (0, 2, lambda: _controlled(_Z)),
    "ch": (0, 2, lambda: _controlled(_H)),
    "cu1": (1, 2, lambda lam: _controlled(_phase(lam))), "cp": (1, 2, lambda lam: _controlled(_phase(lam))),
    "crz": (1, 2, lambda t: _controlled(_rz(t))),
    "cu3": (3, 2, lambda t, p, l: _controlled(_u3(t, p, l))),
    "swap": (0, 2, _swap),
    "ccx": (0, 3, _ccx),
}


# --------------------------------------------------------------------------- #
# Parameter expressions: numbers, pi, + - * / ^, unary minus, parentheses and
# sin / cos / tan / exp / ln / sqrt -- evaluated by a small recursive-descent
# parser (no eval)
# --------------------------------------------------------------------------- #

_TOKEN = re.compile(r"\s*(?:(\d+\.\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|\d+(?:[eE][-+]?\d+)?)|([A-Za-z_]\w*)|(.))")
_FUNCS = {"sin": math.sin, "cos": math.cos, "tan": math.tan, "exp": math.exp, "ln": math.log, "sqrt": math.sqrt}


def eval_param(text):
    toks = []
    pos = 0
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            break
        pos = m.end()
        num, name, op = m.groups()
        if num is not None:
            toks.append(("n", float(num)))
        elif name is not None:
            toks.append(("id", name))
        elif op is not None and not op.isspace():
            toks.append(("op", op))
    if text[pos:].strip():
        raise QasmError(f"bad parameter expression: {text!r}")
    i = 0

    def peek():
        return toks[i] if i < len(toks) else (None, None)

    def take():
        nonlocal i
        if i >= len(toks):
            raise QasmError(f"bad parameter expression: {text!r}")
        i += 1
        return toks[i - 1]

    def atom():
        kind, val = take()
        if kind == "n":
            return val
        if kind == "id":
            if val == "pi":
                return math.pi
            if val in _FUNCS and peek() == ("op", "("):
                take()
                v = expr()
                if take() != ("op", ")"):
                    raise QasmError(f"missing ) in {text!r}")
                return _FUNCS[val](v)
            raise QasmError(f"unknown name {val!r} in {text!r}")
        if (kind, val) == ("op", "("):
            v = expr()
            if take() != ("op", ")"):
                raise QasmError(f"missing ) in {text!r}")
            return v
        if (kind, val) == ("op", "-"):
            return -power()
        if (kind, val) == ("op", "+"):
            return power()
        raise QasmError(f"bad parameter expression: {text!r}")

    def power():
        b = atom()
        if peek() == ("op", "^"):
            take()
            return b ** power()
        return b

    def term():
        v = power()
        while peek() in (("op", "*"), ("op", "/")):
            _, op = take()
            r = power()
            v = v * r if op == "*" else v / r
        return v

    def expr():
        v = term()
        while peek() in (("op", "+"), ("op", "-")):
            _, op = take()
            r = term()
            v = v + r if op == "+" else v - r
        return v

    v = expr()
    if i != len(toks):
        raise QasmError(f"bad parameter expression: {text!r}")
    return v


def _split_params(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
            continue
        depth += (ch == "(") - (ch == ")")
        cur += ch
    out.append(cur)
    return [x for x in (y.strip() for y in out) if x]


_STMT = re.compile(r"^([A-Za-z_]\w*)\s*(?:\((.*)\))?\s*(.*)$", re.S)
_QARG = re.compile(r"^([A-Za-z_]\w*)\s*(?:\[\s*(\d+)\s*\])?$")


def parse_qasm(text):
    """QasmCircuit from OpenQASM 2.0 source (qelib1 gates, one quantum register)."""
    text = re.sub(r"//[^\n]*", "", text)
    stmts = [s.strip() for s in text.split(";")]
    qregs = {}
    circ = None
    for st in stmts:
        if not st:
            continue
        if st.startswith("OPENQASM") or st.startswith("include"):
            continue
        head = st.split(None, 1)[0]
        if head in ("gate", "opaque", "if", "reset"):
            raise QasmError(f"unsupported statement: {head}")
        if head in ("creg", "barrier", "measure"):
            continue
        if head == "qreg":
            m = re.match(r"qreg\s+([A-Za-z_]\w*)\s*\[\s*(\d+)\s*\]$", st)
            if not m:
                raise QasmError(f"bad qreg: {st!r}")
            if qregs:
                raise QasmError("only one quantum register is supported")
            qregs[m.group(1)] = int(m.group(2))
            circ = QasmCircuit(int(m.group(2)))
            continue
        m = _STMT.match(st)
        if not m or circ is None:
            raise QasmError(f"bad statement: {st!r}")
        name, params, args = m.group(1), m.group(2), m.group(3)
        if name not in GATES:
            raise QasmError(f"unknown gate {name!r}")
        npar, nq, build = GATES[name]
        pvals = [eval_param(p) for p in _split_params(params)] if params else []
        if len(pvals) != npar:
            raise QasmError(f"{name} takes {npar} parameters, got {len(pvals)}")
        qargs = [a.strip() for a in args.split(",") if a.strip()]
        if len(qargs) != nq:
            raise QasmError(f"{name} acts on {nq} qubits, got {len(qargs)}")
        qubits = []
        for a in qargs:
            qm = _QARG.match(a)
            if not qm or qm.group(1) not in qregs or qm.group(2) is None:
                raise QasmError(f"bad qubit argument {a!r} (register broadcasting is not supported)")
            q = int(qm.group(2))
            if q >= qregs[qm.group(1)]:
                raise QasmError(f"qubit {a} out of range")
            qubits.append(q)
        if len(set(qubits)) != len(qubits):
            raise QasmError(f"repeated qubit in {st!r}")
        circ.gates.append((np.asarray(build(*pvals), dtype=np.complex128), qubits))
        circ.names.append(name)
    if circ is None:
        raise QasmError("no qreg declared")
    return circ


def load_qasm_as_circuit(qasm_str):
    """load_qasm_as_circuit (src/layer3.jl:481-484)."""
    return parse_qasm(qasm_str)


def load_qasm_as_circuit_from_file(qasm_path):
    """load_qasm_as_circuit_from_file (src/layer3.jl:467-474): the circuit, or
    False when the file does not exist."""
    p = pathlib.Path(qasm_path)
    if not p.is_file():
        return False
    return parse_qasm(p.read_text())


def convert_qiskit_circ_to_network(circ, backend=None, decompose=False, transpile=False, couplings=None):
    """convert_qiskit_circ_to_network (src/layer3.jl:544-574): one gate tensor
    per gate, qubit q of the circuit -> qubit q + 1 of the network.
    transpile=True (qiskit's BasicSwap pass) is not available without qiskit."""
    if transpile:
        raise NotImplementedError("transpile=True needs qiskit's BasicSwap pass, which is not available")
    return network_from_gates(circ.n_qubits, circ.gates, backend=backend, decompose=decompose)
