"""sseqs_topk_noalloc (the top-k entry point that allocates nothing) against
sseqs_topk with a caller workspace, through the C ABI: two calls in a row on
one device, the second on a smaller input, for every score type, sorted and
unsorted; and the size limit reported as an error."""
import numpy as np
import pytest
import torch

from conftest import announce

pytestmark = pytest.mark.gpu


def _with_ws(s, dt, k, sorted_):
    from sseqs_amd import _lib
    from sseqs_amd._lib import call, ptr, stream_of
    out = torch.empty(k, dtype=torch.int64, device=s.device)
    ws = _lib.workspace(_lib.load().sseqs_topk_ws_bytes(s.numel()), s.device)
    call("sseqs_topk", ptr(s), dt, s.numel(), k, int(sorted_), ptr(out), ptr(ws), ws.numel(),
         stream_of(s.device))
    return out


def _no_alloc(s, dt, k, sorted_):
    from sseqs_amd._lib import call, ptr, stream_of
    out = torch.full((k,), -7, dtype=torch.int64, device=s.device)
    call("sseqs_topk_noalloc", ptr(s), dt, s.numel(), k, int(sorted_), ptr(out),
         stream_of(s.device))
    return out


@pytest.mark.parametrize("dtype", [torch.float32, torch.float16, torch.int32])
@pytest.mark.parametrize("sorted_", [True, False])
def test_topk_noalloc_matches_workspace_entry(orc, dev, dtype, sorted_):
    from sseqs_amd import _lib
    dt = {torch.float32: _lib.F32, torch.float16: _lib.F16, torch.int32: _lib.I32}[dtype]
    rng = np.random.default_rng(5)
    for n, k in ((3_000_000, 4000), (50_000, 777)):  # second call smaller than the first
        announce(f"topk_noalloc {dtype} sorted={sorted_} n={n} k={k}")
        if dtype == torch.int32:
            v = rng.integers(0, 5000, n).astype(np.int32)  # heavy ties
        else:
            v = rng.normal(size=n).astype(np.float32)
            v[rng.choice(n, 100, replace=False)] = v[0]  # ties
        s = torch.from_numpy(v).to(dev).to(dtype)
        want = _with_ws(s, dt, k, sorted_)
        got = _no_alloc(s, dt, k, sorted_)
        torch.cuda.synchronize()
        assert torch.equal(got, want)
        ref = orc.topk_desc(s.float().cpu().numpy().astype(np.float64), k)
        assert np.array_equal(got.cpu().numpy(), ref if sorted_ else np.sort(ref))


def test_topk_noalloc_rejects_sizes_past_its_scratch(dev):
    from sseqs_amd import _lib
    from sseqs_amd._lib import SseqsError, call, ptr, stream_of
    announce("topk_noalloc size limit (host-side check)")
    lim = int(_lib.load().sseqs_topk_noalloc_max_n())
    assert lim >= 200_000_000
    s = torch.zeros(16, dtype=torch.float32, device=dev)
    out = torch.zeros(1, dtype=torch.int64, device=dev)
    # n past the limit is refused on the host before anything is launched
    with pytest.raises(SseqsError, match="scratch limit"):
        call("sseqs_topk_noalloc", ptr(s), _lib.F32, lim + 1, 1, 0, ptr(out), stream_of(dev))
