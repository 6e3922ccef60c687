"""trixi_hip_max_dt_noalloc (static device scratch, no allocation) against
trixi_hip_max_dt (the context's buffers), called through the C ABI: equal
results on two calls in a row, the second on a smaller mesh with a slower
state, so that partials left in the scratch by the first call would show."""
import ctypes

import pytest

from cases import Case, product_semi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _both(th, torch, semi, u):
    L = th._lib
    lib = L.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    a, b = ctypes.c_double(), ctypes.c_double()
    L.check(lib.trixi_hip_max_dt(semi.ctx, ctypes.c_void_p(u.data_ptr()), ctypes.byref(a), stream))
    L.check(lib.trixi_hip_max_dt_noalloc(semi.ctx, ctypes.c_void_p(u.data_ptr()), ctypes.byref(b), stream))
    return a.value, b.value


def test_max_dt_noalloc_matches_max_dt(th, torch):
    big = Case("euler3d_level4", 3, "euler", 4, "fd", "ranocha", ic="seeded_random")      # 4096 elements, 16 blocks
    small = Case("euler3d_level2", 3, "euler", 2, "fd", "ranocha", ic="constant")          # 64 elements, 1 block
    results = []
    for case, scale in ((big, 3.0), (small, 1.0)):
        semi = product_semi(th, case)
        try:
            u = th.compute_coefficients(0.0, semi)
            # a faster state on the first (larger) mesh: its partials exceed
            # every partial of the second call
            mom = u.view(semi.nvariables, -1)[1:4]
            mom.mul_(scale)
            torch.cuda.synchronize()
            ref, got = _both(th, torch, semi, u)
            print(f"{case.name}: max_dt {ref!r} noalloc {got!r}", flush=True)
            assert got == ref
            results.append(ref)
        finally:
            semi.close()
    assert results[1] > results[0]  # the second state is slower: a larger step
