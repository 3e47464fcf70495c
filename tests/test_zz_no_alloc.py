"""The allocation-free EG3IC1 entry point (neorx_eg3ic1_noalloc_dev) against
the allocating one (neorx_eg3ic1_dev), through the C ABI: the same kernels
run on a static device scratch array instead of hipMalloc'ed partials, so the
results must agree bit for bit, also when a smaller call follows a larger
one on the same scratch, and sizes beyond the array are rejected."""
import numpy as np
import pytest

import neorx
import oracle
from neorx import synth
from _util import check_close

pytestmark = pytest.mark.gpu

SCRATCH_DOUBLES = 1 << 21


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available()
    torch.cuda.init()
    arch = neorx.init(0)
    assert arch.startswith("gfx950"), arch
    return torch


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _call(torch, name, ne, npb, d_de, d_dp, d_g3, FE0):
    fe = _dev(torch, FE0.ravel(order="F"))
    fp = torch.zeros(npb * npb, dtype=torch.float64, device="cuda")
    e = torch.zeros(1, dtype=torch.float64, device="cuda")
    neorx.call_dev(name, ne, npb, d_de.data_ptr(), d_dp.data_ptr(), d_g3.data_ptr(), fe.data_ptr(), fp.data_ptr(),
                   e.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return fe.cpu().numpy(), fp.cpu().numpy(), e.cpu().numpy()


def test_noalloc_matches_allocating_entry(torch_dev):
    """Two calls in a row, the second smaller: FE, FP and E of the
    allocation-free entry point equal the allocating entry point's bit for
    bit, and the oracle's to the FP64 bar."""
    torch = torch_dev
    for ne, npb in ((10, 3), (6, 2)):
        need = npb * npb * ne * (ne + 1) + npb * npb
        assert need <= SCRATCH_DOUBLES
        print(f"case eg3 noalloc ne={ne} np={npb}", flush=True)
        rng = np.random.default_rng(500 + ne)
        DE, DP = synth.density(ne, 4.0, ne), synth.density(npb, 1.0, 3)
        g3 = rng.standard_normal(ne ** 6 * npb * npb)
        FE0 = rng.standard_normal((ne, ne))
        d_de, d_dp, d_g3 = _dev(torch, DE.ravel(order="F")), _dev(torch, DP.ravel(order="F")), _dev(torch, g3)
        ref = _call(torch, "neorx_eg3ic1_dev", ne, npb, d_de, d_dp, d_g3, FE0)
        got = _call(torch, "neorx_eg3ic1_noalloc_dev", ne, npb, d_de, d_dp, d_g3, FE0)
        for what, a, b in zip(("FE", "FP", "E"), got, ref):
            assert np.array_equal(a, b), f"ne={ne} np={npb}: {what} differs from neorx_eg3ic1_dev"
        FE, FP = np.asfortranarray(FE0.copy()), np.zeros((npb, npb), order="F")
        E = oracle.eg3ic1(ne, npb, DE, DP, g3, FE, FP)
        check_close(got[0], FE.ravel(order="F"), what="FE")
        check_close(got[1], FP.ravel(order="F"), what="FP")
        assert got[2][0] == pytest.approx(E, rel=1e-11)


def test_noalloc_rejects_sizes_beyond_scratch(torch_dev):
    """nebf=24, npbf=60 needs 60^2*24*25 + 60^2 doubles, more than the
    scratch array holds: the call fails on the host before any launch (the
    small buffers passed are never touched), and a fitting call still works
    afterwards."""
    torch = torch_dev
    ne, npb = 24, 60
    assert npb * npb * ne * (ne + 1) + npb * npb > SCRATCH_DOUBLES
    print(f"case eg3 noalloc oversize ne={ne} np={npb}", flush=True)
    small = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = small.data_ptr()
    with pytest.raises(neorx.NeorxError, match="scratch"):
        neorx.call_dev("neorx_eg3ic1_noalloc_dev", ne, npb, p, p, p, p, p, p, 0)
    ne, npb = 4, 2
    rng = np.random.default_rng(7)
    DE, DP = synth.density(ne, 2.0, 2), synth.density(npb, 1.0, 1)
    g3 = rng.standard_normal(ne ** 6 * npb * npb)
    FE0 = np.zeros((ne, ne))
    d_de, d_dp, d_g3 = _dev(torch, DE.ravel(order="F")), _dev(torch, DP.ravel(order="F")), _dev(torch, g3)
    ref = _call(torch, "neorx_eg3ic1_dev", ne, npb, d_de, d_dp, d_g3, FE0)
    got = _call(torch, "neorx_eg3ic1_noalloc_dev", ne, npb, d_de, d_dp, d_g3, FE0)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
