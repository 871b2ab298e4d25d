"""mpdg_heat_op_vmult_noalloc, the apply that allocates nothing, against mpdg_heat_op_vmult.

Both are called through the C ABI (ctypes). The no-allocation variant runs the same kernels
with the apply's scratch taken from a fixed per-GPU device array, so its result must equal
vmult's bit for bit, and the oracle's to the 1e-12 bar of every FP64 apply. Two calls in a
row, the second on a smaller mesh, then the first mesh again; a mesh beyond the array's
maximum is refused before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import make_pair, rel_err

pytestmark = pytest.mark.gpu


def _noalloc(torch, gop, x):
    import meltpooldg_amd as M
    lib = M.load_library()
    y = torch.full_like(x, 7.0)
    rc = lib.mpdg_heat_op_vmult_noalloc(gop.h, C.c_void_p(y.data_ptr()), C.c_void_p(x.data_ptr()),
                                        C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, rc
    return y


def _vmult(torch, gop, x):
    import meltpooldg_amd as M
    lib = M.load_library()
    y = torch.full_like(x, 5.0)
    rc = lib.mpdg_heat_op_vmult(gop.h, C.c_void_p(y.data_ptr()), C.c_void_p(x.data_ptr()),
                                C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, rc
    return y


CASES = [
    # many columns in y and z (rows where four columns meet), per-qp conductivity
    dict(degree=3, nc=(20, 26, 14), mode="qp", dirichlet=0, mapped=None, seed=31),
    # smaller, with Dirichlet faces
    dict(degree=2, nc=(9, 10, 7), mode="const", dirichlet=0b010110, mapped=None, seed=32),
    # mapped mesh
    dict(degree=3, nc=(10, 12, 9), mode="cell", dirichlet=0, mapped=0.05, seed=33),
]


def test_noalloc_matches_vmult_on_consecutive_calls(cuda):
    torch = cuda
    pairs = [make_pair(torch, c["degree"], c["nc"], mode=c["mode"], dirichlet=c["dirichlet"], mapped=c["mapped"],
                       seed=c["seed"]) for c in CASES]
    xs = [torch.tensor(np.random.default_rng(40 + i).uniform(-1.0, 1.0, p[0].n_dofs), device="cuda")
          for i, p in enumerate(pairs)]
    # first mesh, then the smaller one (the array keeps the first call's contents), then the
    # first again, then the mapped mesh
    for i in (0, 1, 0, 2):
        orc, gop, _ = pairs[i]
        print("noalloc case", CASES[i], flush=True)
        y = _noalloc(torch, gop, xs[i])
        ref = _vmult(torch, gop, xs[i])
        torch.cuda.synchronize()
        assert torch.equal(y, ref)
        assert rel_err(y.cpu().numpy(), orc.vmult(xs[i].cpu().numpy())) < 1e-12


def test_noalloc_refuses_meshes_beyond_its_scratch(cuda):
    torch = cuda
    import meltpooldg_amd as M
    # Q3 on 200 x 128 x 128 cells needs 31 * 63 * 601 > 2^20 scratch entries; the call returns
    # before launching anything, so the small vectors are never touched
    gop = M.HeatDiffuseMultiPhaseOperator(3, (200, 128, 128))
    gop.reset_time_increment(1e-3)
    gop.set_conductivity_constant(1.0)
    x = torch.zeros(8, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    lib = M.load_library()
    rc = lib.mpdg_heat_op_vmult_noalloc(gop.h, C.c_void_p(y.data_ptr()), C.c_void_p(x.data_ptr()),
                                        C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == -7
    # FP32 operators are refused
    g32 = M.HeatDiffuseMultiPhaseOperator(2, (4, 4, 4), precision=M.FP32)
    rc = lib.mpdg_heat_op_vmult_noalloc(g32.h, C.c_void_p(y.data_ptr()), C.c_void_p(x.data_ptr()), None)
    assert rc == -1
