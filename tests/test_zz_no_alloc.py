"""wasl_hip_limit_noalloc against wasl_hip_limit.

The allocation-free limiter works on a caller's device buffer with its
working memory in the library's static per-GPU scratch array. Called twice
in a row -- a 2x12x12 mesh, then a smaller 2x8x8 one that reuses the same
scratch with its previous contents -- each on inputs where the positivity
limiter acts, its flags must equal wasl_hip_limit's bit for bit and its
state within 1e-12 (the kernels are the same, so in fact bit for bit), and
its device status must be OK. A third call hits the floor error and must
report it through the status word with the reference's partial state. The
device buffers come from torch (plumbing only).
"""

import ctypes as C
import os

import numpy as np
import pytest

import cases as K
from conftest import GOLDEN
from oracle import oracle as O
from wasl_amd.engine import ERROR, OK, SIMULATION

pytestmark = pytest.mark.gpu

F = K.POSITIVITY_FLOORS


def _bind(lib):
    f = lib.wasl_hip_limit_noalloc
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    f.restype = C.c_int
    return f


def _noalloc(gpu, g, U):
    import torch

    dev = torch.from_numpy(np.ascontiguousarray(U, np.float64)).cuda()
    flags = torch.full((g.E,), 7, dtype=torch.uint8, device="cuda")
    status = torch.full((1,), 12345, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    rc = _bind(gpu.library())(g.h, dev.data_ptr(), flags.data_ptr(), status.data_ptr(), stream)
    assert rc == 0, gpu.library().wasl_hip_last_error()
    torch.cuda.synchronize()
    return dev.cpu().numpy(), flags.cpu().numpy(), int(status.cpu().item())


@pytest.mark.parametrize("p", [1, 2, 3])
def test_noalloc_limiter_matches_limit(gpu, p):
    for n, seed in ((12, 50 + p), (8, 60 + p)):        # the second call is smaller
        m = K.rarefaction_mesh(n)
        c = K.oracle_tri(m, p, boundary=O.OUTFLOW)
        U = K.positivity_state(c, p, seed)
        g = K.engine_tri(gpu, m, p, boundary=gpu.OUTFLOW, limiting=True, floors=F)
        g.state = U
        ref_flags = g.limit()
        ref = g.state
        got, flags, status = _noalloc(gpu, g, U)
        err = K.rel_inf_per_variable(got, ref, 4)
        print(f"P{p} 2x{n}x{n}: {int(ref_flags.sum())} troubled, no-alloc vs limit "
              f"rel L-inf {err:.3e}, status {status}", flush=True)
        assert status == OK
        assert np.array_equal(flags, ref_flags)
        assert err <= K.STEP_TOL
        assert np.array_equal(got, ref)


def test_noalloc_limiter_floor_error(gpu):
    f = np.load(os.path.join(GOLDEN, "reference_vectors.npz"))
    m = K.rarefaction_mesh(8)
    g = K.engine_tri(gpu, m, 2, boundary=gpu.OUTFLOW, limiting=True, floors=F)
    got, _, status = _noalloc(gpu, g, f["floor_error_p2_state"])
    assert status == SIMULATION
    assert K.rel_inf_per_variable(got, f["floor_error_p2_partial"], 4) <= K.STEP_TOL
    # the next call on the shared scratch starts clean
    U = f["positivity_transmissive_p2_state"]
    got, _, status = _noalloc(gpu, g, U)
    assert status == OK
    assert K.rel_inf_per_variable(got, f["positivity_transmissive_p2_limited"], 4) <= K.STEP_TOL


def test_noalloc_limiter_rejects_oversized(gpu):
    """600,000 limited 1D Euler cells at P3 need 132 B each, past the 64 MiB
    scratch: the call fails with a message before it touches the device."""
    n = 600_000
    tags = np.arange(1, n + 2)
    x = np.arange(n + 1) / n
    segs = np.stack([tags[:-1], tags[1:]], 1)
    g = gpu.Engine(gpu.Mesh.lines(tags, x, segs), gpu.EULER, 3,
                   boundary=(gpu.OUTFLOW, gpu.OUTFLOW), limiting=True)
    rc = _bind(gpu.library())(g.h, 1, None, None, None)   # rejected before any use
    assert rc == ERROR
    assert b"more than its" in gpu.library().wasl_hip_last_error()
