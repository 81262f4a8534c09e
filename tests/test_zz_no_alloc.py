"""halma_pot_sum_f32_noalloc (no workspace argument, no allocation: a static device
scratch array) against halma_pot_sum_f32 through the C ABI: the same FP32 and FP64
results, bit for bit, on two calls in a row, the second with a smaller input (the scratch
keeps the first call's contents); sizes beyond the scratch are refused before any GPU
work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call_both(L, rng, n_src, n_test):
    import torch

    from pyhalma_amd.device import ptr, stream_ptr

    dev = torch.device("cuda", 0)
    src = [torch.tensor(rng.normal(scale=0.01, size=n_src), dtype=torch.float32, device=dev)
           for _ in range(3)]
    m = torch.tensor(rng.uniform(1e4, 1e6, n_src), dtype=torch.float32, device=dev)
    tst = [torch.tensor(rng.normal(scale=0.01, size=n_test), dtype=torch.float32, device=dev)
           for _ in range(3)]
    out_a = torch.full((n_test,), -1.0, dtype=torch.float32, device=dev)
    out_b = torch.full((n_test,), -2.0, dtype=torch.float32, device=dev)
    f64_a = torch.full((n_test,), -1.0, dtype=torch.float64, device=dev)
    f64_b = torch.full((n_test,), -2.0, dtype=torch.float64, device=dev)
    nbytes = L.halma_pot_sum_workspace_size(n_src, n_test)
    ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=dev)
    s = stream_ptr()
    rc = L.halma_pot_sum_f32(ptr(m), *(ptr(t) for t in src), n_src, *(ptr(t) for t in tst),
                             n_test, ptr(out_a), ptr(f64_a), ptr(ws), ws.numel(), s)
    assert rc == 0, L.halma_last_error()
    rc = L.halma_pot_sum_f32_noalloc(ptr(m), *(ptr(t) for t in src), n_src,
                                     *(ptr(t) for t in tst), n_test, ptr(out_b), ptr(f64_b), s)
    assert rc == 0, L.halma_last_error()
    torch.cuda.synchronize()
    return out_a, out_b, f64_a, f64_b


def test_noalloc_matches_workspace_entry_twice(gpu):
    import torch

    from pyhalma_amd import _lib

    L = _lib.load()
    rng = np.random.default_rng(61)
    for n_src, n_test in ((40000, 30000), (7000, 1500)):
        out_a, out_b, f64_a, f64_b = _call_both(L, rng, n_src, n_test)
        print(f"n_src={n_src} n_test={n_test}", flush=True)
        assert torch.equal(out_a, out_b)
        assert torch.equal(f64_a, f64_b)
        assert bool((f64_b > 0).all())


def test_noalloc_refuses_sizes_beyond_its_scratch(gpu):
    from pyhalma_amd import _lib

    L = _lib.load()
    assert L.halma_pot_sum_noalloc_max_bytes() == 64 << 20
    n = 2_000_000  # the plan's partial sums alone need far more than 64 MiB
    one = ctypes.c_void_p(16)  # never dereferenced: the call is refused while planning
    rc = L.halma_pot_sum_f32_noalloc(one, one, one, one, n, one, one, one, n, one, one, None)
    assert rc == _lib.E_LIMIT
    assert b"static scratch" in L.halma_last_error()
    # an empty test set is a no-op
    assert L.halma_pot_sum_f32_noalloc(None, None, None, None, 0, None, None, None, 0, None,
                                       None, None) == 0
