"""o3d_statistics_noalloc (statistics_calc with its partial sums in the
library's fixed device array) against o3d_statistics on the same device
fields: bit-identical on two calls in a row, the second on a smaller grid,
so a stale record of the first call would show."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fields(torch, shape, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return [torch.rand(shape, generator=g, dtype=torch.float64).cuda().contiguous()
            for _ in range(3)]


def test_statistics_noalloc_matches_allocating_entry_point():
    import torch
    from osinco3d_amd import capi
    cases = [((36, 40, 48), (0.13, 0.11, 0.17), (0, 0, 0), 1600.0, 0.5),
             ((12, 16, 24), (0.21, 0.19, 0.23), (1, 1, 0), 3000.0, 1.25)]
    for shape, d, bc, re, t in cases:
        u = _fields(torch, shape, seed=shape[0])
        want = capi.statistics(*u, d, re, t, bc)
        got = capi.statistics(*u, d, re, t, bc, noalloc=True)
        torch.cuda.synchronize()
        print(f"noalloc statistics {shape[::-1]} bc {bc}: E_k {got[1]:.6e}", flush=True)
        assert np.array_equal(got, want), (shape, got - want)
        assert got[0] == t and np.isfinite(got).all()


def test_statistics_noalloc_rejects_host_arrays_and_oversized_grids():
    import torch
    from osinco3d_amd import capi
    host = [np.zeros((8, 8, 8)) for _ in range(3)]
    with pytest.raises(capi.O3DError, match="device arrays"):
        capi.statistics(*host, (0.1, 0.1, 0.1), 100.0, 0.0, (0, 0, 0), noalloc=True)
    # 8192 x 1040 x 6: 128 x 130 tiles in one z chunk = 16640 partial records.
    big = [torch.zeros((6, 1040, 8192), dtype=torch.float64, device="cuda") for _ in range(3)]
    with pytest.raises(capi.O3DError, match="partial records"):
        capi.statistics(*big, (0.1, 0.1, 0.1), 100.0, 0.0, (0, 0, 0), noalloc=True)
