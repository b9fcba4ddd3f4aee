"""The compiled PyTorch extension and the ``sseqs`` import alias, on the CPU:
the modules, functions, argument names and defaults of the reference's
bindings (sseqs/xbit_bind.cpp:33-36, sseqs/filter_bind.cpp:67-82,
sseqs/sw_bind.cpp:111-132), and errors raised for host tensors before
anything is launched."""
import pytest
import torch


def test_extension_modules_and_signatures():
    from sseqs_amd import ext
    assert ext.sseqs_sw_ext is ext.sw
    doc = ext.sw.sw_cuda_blosum62.__doc__
    assert ("sw_cuda_blosum62(query_indices_tensor: torch.Tensor, good_idx: torch.Tensor, "
            "ascii_tensor: torch.Tensor, lengths_tensor: torch.Tensor, gap_penalty_param") in doc
    assert "= -10" in doc
    doc = ext.sw.sw_cuda_affine.__doc__
    assert "gap_open: typing.SupportsInt = 11" in doc and "gap_extend: typing.SupportsInt = 1" in doc
    assert "decompress(input: torch.Tensor, n: typing.SupportsInt, lo: typing.SupportsInt)" in \
        ext.xbit.decompress.__doc__
    for name in ("corr", "query_reversed_indices", "weight_LUTS_A", "W", "flat", "offsets",
                 "lengths", "k", "query_length", "maxlen", "n", "query_rfft_tensor"):
        assert f"{name}:" in ext.conv.warp_corr.__doc__, name


def test_extension_rejects_host_tensors():
    from sseqs_amd import ext
    q = torch.zeros(5, dtype=torch.uint8)
    gi = torch.zeros(2, dtype=torch.int32)
    with pytest.raises(RuntimeError, match="device tensor"):
        ext.sw.sw_cuda_affine(q, gi, q, gi)
    with pytest.raises(RuntimeError, match="device tensor"):
        ext.sw.sw_cuda_blosum62(q, gi, q, gi, gap_penalty_param=-1)


def test_sseqs_alias():
    import sseqs
    import sseqs.msa
    import sseqs.sw
    import sseqs.to_a3m
    import sseqs_amd
    assert sseqs.msa is sseqs_amd.msa and sseqs.sw is sseqs_amd.sw
    from sseqs.msa import conv, sw, xbit  # the reference's module-level names
    from sseqs_amd import ext
    assert (xbit, conv, sw) == (ext.xbit, ext.conv, ext.sw)
    assert sseqs.to_a3m.to_a3m is sseqs_amd.to_a3m.to_a3m
    assert sseqs.sw is sseqs_amd.sw
