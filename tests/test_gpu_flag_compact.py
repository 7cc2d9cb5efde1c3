"""The ordered compaction of the query flags (k_b_flag_count, a scan of the
tile counts, k_b_flag_write), called through its test hook and checked
against torch.nonzero: sizes around the 16-byte lane load and the
4,096-flag tile, no flag set, every flag set, and garbage in the padding
past the last flag (which must be ignored)."""

import pytest

from gpu_inscy_amd import _lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def compact(flags, nb):
    temp_bytes = _lib.debug_flag_compact_temp(nb)
    assert temp_bytes > 0
    temp = torch.empty(temp_bytes, dtype=torch.uint8, device="cuda")
    out = torch.full((nb,), -7, dtype=torch.int32, device="cuda")
    nsel = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    _lib.debug_flag_compact(flags.data_ptr(), nb, temp.data_ptr(), temp_bytes, out.data_ptr(), nsel.data_ptr(),
                            torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    k = int(nsel.item())
    return k, out[:k].cpu(), out[k:].cpu()


@pytest.mark.parametrize("nb", [1, 15, 16, 17, 4095, 4096, 4097, 3 * 4096 + 9, 1_000_003])
@pytest.mark.parametrize("density", [0.0, 0.01, 0.5, 1.0])
def test_flag_compaction_matches_nonzero(nb, density):
    print(f"flag compaction nb={nb} density={density}", flush=True)
    g = torch.Generator().manual_seed(nb * 7 + int(density * 100))
    padded = (nb + 15) // 16 * 16
    host = torch.zeros(padded, dtype=torch.uint8)
    host[:nb] = (torch.rand(nb, generator=g) < density).to(torch.uint8)
    host[nb:] = 1  # padding bytes past nb are not flags
    flags = host.cuda()
    k, got, rest = compact(flags, nb)
    want = torch.nonzero(host[:nb]).flatten().to(torch.int32)
    assert k == want.numel()
    assert torch.equal(got, want)
    assert bool((rest == -7).all())  # nothing written past the count
