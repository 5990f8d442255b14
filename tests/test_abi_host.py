"""CPU checks of the boundary and the host logic (no device calls).

- libtofe.so is built in-tree, loads, and exports every entry point include/tofe.h
  declares (and the binding's list matches the header);
- the ctypes structures match the C layout;
- SplitMix64 known answers, the synthetic workloads' boundary conditions, and the
  host mirror's mesh numbering and Set_BC selection (RecMesh.cpp:486-597,
  Inputs.cpp:105-173), checked against the oracle's numbering.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from topopt3d_amd import _lib, configs
from topopt3d_amd.topopt import BCType, TopOpt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "tofe.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(tofe_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        from topopt3d_amd import build

        build.build()
    L = C.CDLL(_lib.LIB_PATH)
    declared = header_functions()
    assert len(declared) >= 26
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTS) == declared
    _lib.lib()  # argtypes for every symbol bind


def _c_layout(struct, fields, tmp_path):
    """sizeof and offsetof of a C struct from include/tofe.h, compiled with gcc."""
    src = tmp_path / "layout.c"
    body = "".join(f'printf("%zu\\n", offsetof({struct}, {f}));' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "tofe.h"\n'
                   f'int main(void){{printf("%zu\\n", sizeof({struct}));{body}return 0;}}\n')
    exe = tmp_path / "layout"
    import subprocess

    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    return int(out[0]), [int(v) for v in out[1:]]


@pytest.mark.parametrize("pystruct,cname", [(_lib.TofeDesc, "tofe_desc"),
                                            (_lib.TofeStats, "tofe_stats")])
def test_struct_layout_matches_header(pystruct, cname, tmp_path):
    names = [f for f, _ in pystruct._fields_]
    size, offs = _c_layout(cname, names, tmp_path)
    assert C.sizeof(pystruct) == size
    assert [getattr(pystruct, f).offset for f in names] == offs


def test_splitmix64_known_answers():
    # first outputs of SplitMix64 from state 0 and state 1234567
    assert int(configs.splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF
    assert [int(v) for v in configs.splitmix64(1234567, 3)] == [
        6457827717110365317, 3203168211198807973, 9817491932198370423]
    u = configs.uniform(2, 100000, 0.01, 1.0)
    assert u.min() >= 0.01 and u.max() < 1.0
    assert abs(u.mean() - 0.505) < 0.005


def test_cantilever_boundary_conditions():
    pb = configs.cantilever_48x24x24()
    assert pb.ndof == 91875 and pb.nelem == 27648
    assert pb.fixed_dof.sum() == 25 * 25 * 3
    fixed_nodes = np.flatnonzero(pb.fixed_dof[0::3])
    assert np.all(fixed_nodes % 49 == 0)
    load = np.flatnonzero(pb.F)
    assert load.tolist() == [3 * (48 + 49 * 12) + 2] and pb.F[load[0]] == -1.0
    pb2 = configs.cantilever_192x96x96()
    assert pb2.ndof == 5447811 and pb2.nelem == 1769472 and pb2.fixed_dof.sum() == 97 * 97 * 3


def test_host_mirror_mesh_and_bcs(oracle):
    t = TopOpt()
    t.CreateMesh([0, 6, 0, 3, 0, 4], (6, 3, 4), 1.5)
    assert t.nElem == 72 and t.nNode == 7 * 4 * 5
    for el in [0, 5, 17, 71]:
        nodes = np.zeros(8, np.int64)
        oracle.lib().orc_element_nodes(6, 3, 4, el, nodes.ctypes.data_as(C.POINTER(C.c_int64)))
        assert t.element[el].tolist() == nodes.tolist()
    assert np.allclose(t.node[t.element[71, 6]], [6, 3, 4])
    # support on the x = 0 face (radius 1e-10 in x), load at the bottom-centre of x = 6
    sup = t.Set_BC([0, 1.5, 2], [1e-10, 10, 10], [-100, 100, -100, 100, -100, 100], [1, 1, 1],
                   BCType.SUPPORT)
    assert sup.size == 4 * 5 and np.all(t.node[sup, 0] == 0)
    ld = t.Set_BC([6, 1.5, 0], [1e-10, 0.6, 1e-10], [-100, 100, -100, 100, -100, 100], [0, 0, -1],
                  BCType.LOAD)
    assert ld.size == 2  # y = 1 and y = 2 both lie within 0.6 of 1.5
    # the limits box restricts the selection
    lim = t.Set_BC([0, 1.5, 2], [1e-10, 10, 10], [-100, 100, -100, 1, -100, 100], [1, 1, 1],
                   BCType.SUPPORT)
    assert np.all(t.node[lim, 1] <= 1) and lim.size == 2 * 5
