"""Data formats (host code, no GPU): the Paraview VTU/PVD output and the DoF permutation between a
DoFHandler's numbering (given by its support points) and the lexicographic node grid."""
import base64
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from oracle import oracle as O


def read_vtu(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    arrays = {}
    for da in root.iter("DataArray"):
        raw = base64.b64decode(da.text.strip())
        nbytes = struct.unpack("<Q", raw[:8])[0]
        dt = {"Float64": "<f8", "Int64": "<i8", "UInt8": "u1"}[da.get("type")]
        a = np.frombuffer(raw[8:8 + nbytes], dtype=dt)
        assert a.nbytes == nbytes
        arrays[da.get("Name")] = a.reshape(-1, int(da.get("NumberOfComponents")))
    return int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells")), arrays


@pytest.mark.parametrize("k", [1, 2, 3])
def test_vtu_patches_carry_the_fe_field(tmp_path, k):
    """DataOut::build_patches(k): each cell split into k^3 equal hexahedra; at their vertices the written
    values are the FE field there, exact for a polynomial of degree <= k given by its nodal values."""
    import meltpooldg_amd as M
    nc, L = (3, 2, 2), (0.6, 0.5, 0.4)
    X = O.Operator(k, nc, L).node_coords()
    poly = lambda x, y, z: 1.0 + x ** k - 2.0 * y * z ** (k - 1) + 0.5 * z
    u = poly(*X)
    M.write_vtu(tmp_path / "a.vtu", k, nc, L, {"temperature": u, "heat_source": 2 * u})
    npts, ncell, A = read_vtu(tmp_path / "a.vtu")
    nd = [k * n + 1 for n in nc]
    assert npts == nd[0] * nd[1] * nd[2] and ncell == (nd[0] - 1) * (nd[1] - 1) * (nd[2] - 1)
    P = A["Points"]
    h = [L[d] / (k * nc[d]) for d in range(3)]
    i = np.arange(npts)
    assert np.allclose(P[:, 0], (i % nd[0]) * h[0]) and np.allclose(P[:, 2], (i // (nd[0] * nd[1])) * h[2])
    assert np.allclose(A["temperature"][:, 0], poly(P[:, 0], P[:, 1], P[:, 2]), atol=1e-12)
    assert np.allclose(A["heat_source"][:, 0], 2 * A["temperature"][:, 0])
    # hexahedra: VTK vertex order, edges of one sub-cell length, offsets 8, 16, ...
    conn = A["connectivity"].reshape(-1, 8)
    assert (A["types"] == 12).all() and (A["offsets"][:, 0] == 8 * np.arange(1, ncell + 1)).all()
    e = P[conn[:, 1]] - P[conn[:, 0]]
    assert np.allclose(e, [h[0], 0, 0]) and np.allclose(P[conn[:, 3]] - P[conn[:, 0]], [0, h[1], 0])
    assert np.allclose(P[conn[:, 4]] - P[conn[:, 0]], [0, 0, h[2]])
    assert np.allclose(P[conn[:, 6]] - P[conn[:, 0]], h)


def test_pvd_and_paraview_output(tmp_path):
    import meltpooldg_amd as M

    class Op:  # the three attributes the writer reads
        degree, n_cells, extent = 1, (2, 2, 2), (1.0, 1.0, 1.0)

    out = M.ParaviewOutput(tmp_path, Op(), filename="solution", n_digits=3)
    u = np.zeros(27)
    names = [out.write(s, 0.1 * s, {"temperature": u + s}) for s in (0, 1, 2)]
    assert names == ["solution_000.vtu", "solution_001.vtu", "solution_002.vtu"]
    root = ET.parse(tmp_path / "solution.pvd").getroot()
    ds = root.findall("Collection/DataSet")
    assert [d.get("file") for d in ds] == names
    assert [float(d.get("timestep")) for d in ds] == [0.0, 0.1, 0.2]
    assert read_vtu(tmp_path / names[2])[2]["temperature"].max() == 2.0


@pytest.mark.parametrize("k", [1, 2, 4])
def test_dof_permutation_from_support_points(k):
    """Support points in any DoF order map back to their lexicographic nodes; off-grid or repeated points
    are refused."""
    import meltpooldg_amd as M
    nc, L = (3, 2, 4), (0.3, 0.2, 0.8)
    X = O.Operator(k, nc, L).node_coords().T.copy()  # lexicographic support points
    rng = np.random.default_rng(k)
    order = rng.permutation(len(X))  # DoF i sits on node order[i]
    lex = M.dof_permutation(k, nc, L, X[order])
    assert (lex == order).all()
    bad = X[order].copy()
    bad[5, 0] += 0.3 * L[0] / nc[0] / max(k, 2)  # between nodes
    with pytest.raises(M.MpdgError):
        M.dof_permutation(k, nc, L, bad)
    dup = X[order].copy()
    dup[1] = dup[0]
    with pytest.raises(M.MpdgError):
        M.dof_permutation(k, nc, L, dup)


@pytest.mark.gpu
def test_device_permutation_round_trip(cuda):
    torch = cuda
    import meltpooldg_amd as M
    nc, L, k = (4, 3, 2), (0.4, 0.3, 0.2), 3
    X = O.Operator(k, nc, L).node_coords().T.copy()
    order = np.random.default_rng(0).permutation(len(X))
    lex = torch.tensor(M.dof_permutation(k, nc, L, X[order]), device="cuda")
    dealii_vec = torch.tensor(np.random.default_rng(1).uniform(size=len(X)), device="cuda")
    lex_vec = torch.empty_like(dealii_vec)
    back = torch.empty_like(dealii_vec)
    M.permute(lex, dealii_vec, lex_vec, to_lex=True)
    M.permute(lex, lex_vec, back, to_lex=False)
    torch.cuda.synchronize()
    assert torch.equal(back, dealii_vec)
    assert torch.equal(lex_vec[lex], dealii_vec)
