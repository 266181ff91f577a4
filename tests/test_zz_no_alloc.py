"""The allocation-free multi-dot (gd_multidot_f64_noalloc) against the
multi-dot it is a variant of (gd_multidot_f64), through the C-ABI: the same
kernels, so the same results bit for bit, on two calls in a row, the second
smaller than the first (the fixed scratch array keeps the first call's
partials), and against an FP64 torch reference."""
import ctypes as C

import pytest
import torch

from gendil_amd import _lib as L

pytestmark = pytest.mark.gpu


def multidot(entry, V, w):
    k, n = V.shape
    res = torch.full((k,), float("nan"), dtype=torch.float64, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(L.lib(), entry)(n, k, V.data_ptr(), n, w.data_ptr(), res.data_ptr(), stream)
    L.check(rc, entry)
    torch.cuda.synchronize()
    return res


def test_noalloc_multidot_matches_multidot_on_two_calls():
    gen = torch.Generator(device="cuda").manual_seed(11)
    for n, k in ((3_000_017, 32), (4_099, 7)):
        print(f"multidot n={n} k={k}", flush=True)
        V = torch.rand((k, n), generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
        w = torch.rand(n, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1
        ref = multidot("gd_multidot_f64", V, w)
        got = multidot("gd_multidot_f64_noalloc", V, w)
        assert torch.equal(got, ref), (got - ref).abs().max().item()
        exact = (V @ w).cpu()
        assert torch.allclose(got.cpu(), exact, rtol=1e-12, atol=1e-9)


def test_noalloc_multidot_rejects_more_than_32_vectors():
    res = torch.empty(33, dtype=torch.float64, device="cuda")
    V = torch.zeros((33, 8), dtype=torch.float64, device="cuda")
    w = torch.zeros(8, dtype=torch.float64, device="cuda")
    rc = L.lib().gd_multidot_f64_noalloc(8, 33, V.data_ptr(), 8, w.data_ptr(), res.data_ptr(), None)
    assert rc == L.GD_ERR_INVALID
    assert b"at most 32" in L.lib().gd_last_error()
