"""hps_down_level_static (no workspace, library-owned static scratch) against
hps_down_level with a caller workspace: identical outputs, through the C ABI,
on two calls in a row with the second one smaller."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [torch.complex128, torch.float32])
def test_static_down_level_matches_workspace_version(gpu, dtype):
    from jaxhps_amd import _lib
    from jaxhps_amd._stages import down_level, down_level_static
    from jaxhps_amd._tables import quad_merge_tables

    lib = _lib.load_library()
    cap = int(lib.hps_static_scratch_bytes())
    assert cap == 32 << 20
    kind = "iti" if dtype.is_complex else "dtn"
    g = torch.Generator(device=gpu).manual_seed(7)

    def rand(*shape):
        if dtype.is_complex:
            return torch.complex(torch.randn(*shape, generator=g, device=gpu, dtype=torch.float64),
                                 torch.randn(*shape, generator=g, device=gpu, dtype=torch.float64)).to(dtype)
        return torch.randn(*shape, generator=g, device=gpu, dtype=torch.float64).to(dtype)

    # first call: 1024 merges of children with 56 Gauss nodes per side (7.3 MB of
    # interface data in complex128); second call: a smaller level on the same scratch
    for nm, m, nsrc in ((1024, 14, 4), (16, 28, 3)):
        tb = quad_merge_tables(m, kind)
        print(f"static down level {dtype} nm={nm} m={m} nsrc={nsrc}", flush=True)
        S = rand(nm, tb.ni, tb.ne).contiguous()
        gt = rand(nm, tb.ni, nsrc).contiguous()
        g_in = rand(nm, tb.ne, nsrc).contiguous()
        ref = down_level(S, gt, g_in, tb)
        out = down_level_static(S, gt, g_in, tb)
        torch.cuda.synchronize()
        assert out.shape == ref.shape
        assert torch.equal(out, ref)
        # the interface part depends on S g + gt: a nonzero result rules out an unwritten scratch
        assert float(out.abs().max()) > 0


def test_static_down_level_rejects_oversized_calls(gpu):
    from jaxhps_amd import _lib
    from jaxhps_amd._stages import down_level_static
    from jaxhps_amd._tables import quad_merge_tables

    tb = quad_merge_tables(14, "dtn")
    nm, nsrc = 64, (33 << 20) // (64 * tb.ni * 8) + 1
    S = torch.zeros(nm, tb.ni, tb.ne, dtype=torch.float64, device=gpu)
    gin = torch.zeros(nm, tb.ne, nsrc, dtype=torch.float64, device=gpu)
    assert nm * tb.ni * nsrc * 8 > int(_lib.load_library().hps_static_scratch_bytes())
    with pytest.raises(_lib.HPSLibraryError, match="static scratch"):
        down_level_static(S, None, gin, tb)
