"""TopOpt::MeshOut / PrintVals equivalents (TopOpt.cpp:97-360, 442-480) from a context on
the GPU: every file parses with the independent PETSc-binary / raw parsers and matches the
oracle (filters, interpolations) and the host mirror (numbering, coordinates) exactly."""
import numpy as np
import pytest

from topopt3d_amd import configs
from topopt3d_amd.topopt import TopOpt

from test_io_formats import parse_aij, parse_vec

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ib", [4, 8])
def test_mesh_out_and_print_vals(oracle, tofe, tmp_path, ib):
    nel, L = (9, 6, 5), 3
    print(f"case mesh_out nel={nel} index_bytes={ib}", flush=True)
    pb = configs.cantilever(nel, seed=2, lo=0.05, hi=1.0)
    ctx = tofe.Context(nel, h=(0.5, 0.5, 0.5), mg_levels=L, fixed_dof=pb.fixed_dof, rmax=1.6)
    ctx.set_design(pb.x)
    ctx.mesh_out(tmp_path, origin=(1.0, -2.0, 0.5), index_bytes=ib)
    it = np.int32 if ib == 4 else np.int64
    ne, nn = pb.nelem, ctx.nnode
    assert np.array_equal(np.fromfile(tmp_path / "Element_Distribution.bin", it), [0, ne])
    assert np.array_equal(np.fromfile(tmp_path / "Node_Distribution.bin", it), [0, nn])
    mesh = TopOpt()
    mesh.CreateMesh([1.0, 5.5, -2.0, 1.0, 0.5, 3.0], nel, 0.75)
    assert np.array_equal(np.fromfile(tmp_path / "elements.bin", it).reshape(ne, 8), mesh.element)
    assert np.array_equal(np.fromfile(tmp_path / "nodes.bin", np.float64).reshape(nn, 3), mesh.node)
    assert (np.fromfile(tmp_path / "active.bin", np.uint8) == 1).all()
    P = oracle.filter_csr(*nel, 1.5, h=(0.5, 0.5, 0.5))
    m, n, rp, col, w = parse_aij(tmp_path / "Filter.bin", ib)
    orp, ocol, ow = P.arrays()
    assert (m, n) == (ne, ne)
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol) and np.array_equal(w, ow)
    R, REdge = oracle.rfilter_csr(*nel, 1.6, h=(0.5, 0.5, 0.5))
    m, n, rp, col, w = parse_aij(tmp_path / "Max_Filter.bin", ib)
    orp, ocol, ow = R.arrays()
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol) and np.array_equal(w, ow)
    assert np.array_equal(parse_vec(tmp_path / "Void_Edge_Volume.bin", ib), REdge)
    dims = [n_ + 1 for n_ in nel]
    for l in range(ctx.levels - 1):
        Po = oracle.build_P(*dims)
        m, n, rp, col, w = parse_aij(tmp_path / f"P{l}.bin", ib)
        orp, ocol, ow = Po.arrays()
        assert (m, n) == Po.shape[:2]
        assert np.array_equal(rp, orp) and np.array_equal(col, ocol) and np.array_equal(w, ow)
        dims = [(d + 1) // 2 for d in dims]
    U = np.random.default_rng(1).standard_normal(ctx.ndof)
    ctx.print_vals(tmp_path, "_pen3", U, index_bytes=ib)
    assert np.array_equal(parse_vec(tmp_path / "U_pen3.bin", ib), U)
    for name, fld in (("x", tofe.FIELD_X), ("V", tofe.FIELD_V), ("Es", tofe.FIELD_ES),
                      ("E", tofe.FIELD_E)):
        assert np.array_equal(parse_vec(tmp_path / f"{name}_pen3.bin", ib), ctx.field(fld)), name
    ctx.close()
