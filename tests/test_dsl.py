"""DSL record / replay (src/backends/dsl.jl, src/layer1.jl:211-315) on the
CPU: DSLBackend writes the reference's command text while a driver runs;
the recorded program, interpreted here with the oracle's contraction, gives
the reference test's amplitude (test/layer1_tests.jl:11-44) and a full QFT
wavefunction.  The GPU replay (execute_dsl_file) is tested in
test_gpu_drivers.py."""
import json
import math

import numpy as np
import pytest

from oracle import pq_oracle as O
import picoquant_hip as P
from picoquant_hip import circuits as C, dsl as D


def oracle_replay(dsl_path, store_path):
    """The recorded program executed with numpy (the oracle as checker)."""
    store = D.open_store(store_path)
    t, out = {}, {}
    for line in open(dsl_path):
        cmd = line.split()
        if cmd[0] == "tensor":
            t[cmd[1]] = store.read(cmd[2])
        elif cmd[0] == "ncon":
            t[cmd[1]] = O.contract_tensors_data(t[cmd[2]], [int(x) for x in cmd[3].split(",")],
                                                t[cmd[4]], [int(x) for x in cmd[5].split(",")])
        elif cmd[0] == "del":
            t.pop(cmd[1], None)
        elif cmd[0] == "permute":
            t[cmd[1]] = np.transpose(t[cmd[1]], [int(x) - 1 for x in cmd[2].split(",")])
        elif cmd[0] == "reshape":
            s = t[cmd[1]].shape
            dims = [math.prod(s[g - 1] for g in map(int, grp.split(","))) for grp in cmd[2].split(";")]
            t[cmd[1]] = np.reshape(t[cmd[1]], dims, order="F")
        elif cmd[0] == "save":
            out[cmd[3]] = t[cmd[1]]
        else:
            raise AssertionError(cmd)
    return out


def test_record_ghz_amplitude(tmp_path):
    tl, h5 = tmp_path / "contract_network.tl", tmp_path / "tensor_data.npz"
    be = D.DSLBackend(tl, h5, "", True)
    tn = P.network_from_gates(3, [(C.H, [0]), (C.CX, [0, 1]), (C.CX, [1, 2])], backend=be)
    P.add_input(tn, "000")
    P.add_output(tn, "000")
    plan = P.random_contraction_plan(tn, np.random.default_rng(0))
    P.contract_network(tn, plan)
    lines = tl.read_text().splitlines()
    ops = [l.split()[0] for l in lines]
    assert ops.count("tensor") == 9 and ops.count("ncon") == 8 and ops[-1] == "save"
    assert all(ops[i + 1] == "del" and ops[i + 2] == "del" for i, o in enumerate(ops) if o == "ncon")
    assert sorted(D.TensorStore(h5).labels()) == sorted(l.split()[1] for l in lines if l.startswith("tensor"))
    res = oracle_replay(tl, h5)["result"]
    assert abs(complex(np.asarray(res).reshape(-1)[0]) - 1 / math.sqrt(2)) < 1e-15
    assert be.metrics.flops > 0


@pytest.mark.parametrize("ext", ["npz", "h5"])
def test_record_qft_wavefunction(tmp_path, ext):
    tl, h5 = tmp_path / "qft.tl", tmp_path / f"qft.{ext}"
    be = D.DSLBackend(tl, h5, "", True)
    n, x = 5, 19
    tn = P.network_from_gates(n, C.qft_gates(n), backend=be)
    P.add_input(tn, "".join(str((x >> q) & 1) for q in range(n)))
    P.full_wavefunction_contraction(tn, "vector")
    text = tl.read_text()
    assert "\npermute " in text and "\nreshape " in text
    psi = oracle_replay(tl, h5)["result"]
    np.testing.assert_allclose(psi, O.qft_closed_form(n, x), atol=1e-14)


def test_decompose_and_view_commands(tmp_path):
    be = D.DSLBackend(tmp_path / "a.tl", tmp_path / "a.npz", "", True)
    assert be.decompose_tensor("node_1", [1, 3], [2, 4], left_label="node_2", right_label="node_3") == 0
    be.view_tensor("node_4", "node_2", 2, [1, 1])
    last = (tmp_path / "a.tl").read_text().splitlines()
    assert last[0].startswith("decompose node_1 node_2 1,3 node_3 2,4 ")
    assert json.loads(last[0].split(" ", 6)[6]) == {"threshold": 1e-13, "max_rank": 0}
    assert last[1] == "view node_4 node_2 2 1,1"
