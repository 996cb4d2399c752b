"""The allocation-free entry point gimic_jtensor_batch_device_noalloc (work
buffers in the library's per-device scratch array) against the allocating
gimic_jtensor_batch_device, through the C ABI: two calls in a row, the
second with fewer points, in FP64 (1e-12 relative per point: only the
chunking, and with it the summation order of the block lists, differs) and
mixed precision (1e-5 per point, the mixed path's stated bound).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _device_points(rng, n, half):
    import torch
    return torch.tensor(rng.uniform(-half, half, (n, 3)), dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("system,sizes,half", [("ring12", (5000, 1000), 6.0),
                                               ("cage60", (40000, 3000), 7.5)])
def test_noalloc_matches_allocating_entry(system, sizes, half):
    import torch
    from gimic_amd import synth
    s = getattr(synth, system)()
    rng = np.random.default_rng(21)
    with s.context(screening=1e-8) as ctx:
        # one allocating call per precision makes the right operands resident
        warm = _device_points(rng, 1, half)
        for prec in ("fp64", "mixed"):
            ctx.jtensor_device(warm, precision=prec)
        for n in sizes:
            pts = _device_points(rng, n, half)
            for prec, tol in (("fp64", 1e-12), ("mixed", 1e-5)):
                print(f"case noalloc {system} n={n} {prec}", flush=True)
                ref = ctx.jtensor_device(pts, precision=prec)
                torch.cuda.synchronize()
                nref = ctx.fp64_refined()[1]
                out = ctx.jtensor_device_noalloc(pts, precision=prec)
                torch.cuda.synchronize()
                e = rel_err(out.cpu().numpy(), ref.cpu().numpy())
                assert e.max() <= tol, (system, n, prec, e.max())
                if prec == "mixed":
                    # the same points go through the FP64 pass near nuclei
                    assert ctx.fp64_refined()[1] == nref
                    J64 = ctx.jtensor_device(pts, precision="fp64").cpu().numpy()
                    assert rel_err(out.cpu().numpy(), J64).max() <= 1e-5


def test_noalloc_needs_resident_operands_and_takes_empty_input():
    import torch
    from gimic_amd import synth
    from gimic_amd.lib import GimicError
    s = synth.ring12()
    with s.context(screening=1e-8) as ctx:
        pts = _device_points(np.random.default_rng(2), 10, 5.0)
        with pytest.raises(GimicError, match="not resident"):
            ctx.jtensor_device_noalloc(pts, precision="fp64")
        empty = torch.zeros((0, 3), dtype=torch.float64, device="cuda")
        assert ctx.jtensor_device_noalloc(empty, precision="fp64").shape == (0, 9)
