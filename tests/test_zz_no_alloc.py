"""necgpu_zgetrf_noalloc (the FP64 LU whose workspace is a fixed device
array of the library) against necgpu_zgetrf (the same factorisation on a
caller-provided workspace), through the C ABI: two calls in a row, the
second one smaller, each with the pivots identical and the factors within
the LU bar of the parity tests; the size limit is refused with a message."""
import numpy as np
import pytest
import torch

from necpp_amd import necpp as api

pytestmark = pytest.mark.gpu

NMAX = 4096


def factor_both(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    # the existing entry point, workspace allocated by the caller
    a1 = torch.from_numpy(A.copy()).to("cuda")
    p1 = torch.zeros(2 * n, dtype=torch.int32, device="cuda")
    i1 = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    work = torch.empty(int(api.necgpu_zgetrf_workspace(n)), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    assert api.necgpu_zgetrf(a1.data_ptr(), n, n, p1.data_ptr(), i1.data_ptr(), work.data_ptr(), None) == 0, \
        api.nec_error_message()
    # the no-allocation entry point on the same input
    a2 = torch.from_numpy(A.copy()).to("cuda")
    p2 = torch.zeros(2 * n, dtype=torch.int32, device="cuda")
    i2 = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    print(f"necgpu_zgetrf_noalloc n={n} seed={seed}", flush=True)
    assert api.necgpu_zgetrf_noalloc(a2.data_ptr(), n, n, p2.data_ptr(), i2.data_ptr(), None) == 0, \
        api.nec_error_message()
    torch.cuda.synchronize()
    return A, (a1.cpu().numpy(), p1.cpu().numpy(), int(i1.item())), (a2.cpu().numpy(), p2.cpu().numpy(),
                                                                     int(i2.item()))


def check(n, seed):
    A, (lu1, p1, i1), (lu2, p2, i2) = factor_both(n, seed)
    assert i1 == 0 and i2 == 0
    assert np.array_equal(p1, p2)
    scale = np.max(np.abs(lu1))
    assert np.max(np.abs(lu2 - lu1)) <= 1e-12 * scale * max(1.0, n / 64.0)
    # and the factors are those of A: P A = L U to FP64 rounding
    perm = p2[n:]
    L = np.tril(lu2, -1) + np.eye(n)
    U = np.triu(lu2)
    assert np.max(np.abs(L @ U - A[perm])) <= 1e-13 * n * np.max(np.abs(A))


def test_noalloc_lu_matches_the_workspace_lu_twice():
    check(1500, 11)
    check(333, 12)  # second call, smaller: the fixed workspace is reused


def test_noalloc_lu_refuses_orders_beyond_its_workspace():
    n = NMAX + 1
    assert api.necgpu_zgetrf_noalloc(None, n, n, None, None, None) == 1
    assert "exceeds the largest order" in api.nec_error_message()
