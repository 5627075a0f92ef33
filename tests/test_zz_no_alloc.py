"""``sp_reduce_noalloc`` (the allocation-free reduction) against ``sp_reduce`` through the C ABI:
bit-identical results on two calls in a row, the second on a smaller input, for every kind,
including an input large enough for the clamped stage-1 grid; and the same call recorded in a
hipGraph, replayed with eager work in between."""
import gc

import pytest
import torch

from stenpy_amd import _lib as L
from tests._util import field

pytestmark = pytest.mark.gpu

KINDS = ["sum", "mean", "norm_l2", "variance", "variance_biased", "min_max", "integrate"]


def both(kind, x):
    wp = (0.5, 1.0, 2.0, 0.0) if kind == "integrate" else None
    ref = L.reduce(kind, x, wp).reshape(-1).clone()
    out = torch.full((2,), float("nan"), dtype=torch.float64, device=x.device)
    gc.collect()                      # tensors of earlier tests freed now, not during the call
    before = torch.cuda.memory_allocated(x.device)
    L.reduce_noalloc(kind, x, out, wp)
    assert torch.cuda.memory_allocated(x.device) <= before      # no caching-allocator growth
    return out[:ref.numel()], ref


@pytest.mark.parametrize("kind", KINDS)
def test_noalloc_matches_the_workspace_entry_point_on_two_calls(kind):
    big = field((160, 160, 200), 41)      # 5.1 M elements: the stage-1 grid is clamped
    small = field((10, 12, 14), 42)
    for x in (big, small):
        got, ref = both(kind, x)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (kind, tuple(x.shape), got.tolist(), ref.tolist())


def test_noalloc_scratch_size_and_bad_input():
    assert L.lib().sp_reduce_noalloc_scratch_bytes() == 2048 * 4 * 8
    x = field((8, 8, 8), 43)
    out = torch.zeros(2, dtype=torch.float64, device=x.device)
    with pytest.raises(ValueError):
        L.reduce_noalloc("sum", x.transpose(0, 2), out)   # not contiguous: rejected on the host


def test_noalloc_under_graph_capture():
    x = field((64, 96, 160), 44)
    eager = torch.zeros(1, dtype=torch.float64, device=x.device)
    L.reduce_noalloc("variance", x, eager)           # first call: the scratch address lookup
    torch.cuda.synchronize()
    out = torch.zeros(1, dtype=torch.float64, device=x.device)
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g):
            L.reduce_noalloc("variance", x, out)
    torch.cuda.current_stream().wait_stream(s)
    other, tmp = field((20, 20, 20), 45), torch.zeros(1, dtype=torch.float64, device=x.device)
    for rep in range(3):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), rep
        L.reduce_noalloc("sum", other, tmp)           # eager work on the same scratch array
        torch.cuda.synchronize()
