"""wfs_observables_noalloc: the box moments without a plan and without allocation.

Its partial records live in a static device workspace; the result must equal what
wfs_observables on a plan of the same grid writes, call after call (the workspace keeps the
previous call's contents), and requests beyond the workspace are refused on the host.
"""

import numpy as np
import pytest
import torch

from wavefunctionsimpython_amd import _abi

pytestmark = pytest.mark.gpu


def _state(shape, seed, precision):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    dt = np.complex128 if precision == _abi.WFS_C128 else np.complex64
    return torch.from_numpy(a.astype(dt)).to("cuda:0")


def _compare(n, dx, precision, batch, boxes, seed):
    shape = n if batch == 1 else (batch,) + n
    psi = _state(shape, seed, precision)
    plan = _abi.Plan(n, dx, 0.01, 1.0, precision, _abi.WFS_METHOD_FOURIER, batch=batch)
    want = plan.observables(psi, boxes)
    out = torch.full((batch * len(boxes) * 7,), float("nan"), dtype=torch.float64, device="cuda:0")
    _abi.observables_noalloc(n, dx, precision, psi, boxes, out, batch=batch)
    torch.cuda.synchronize()
    got = out.reshape(want.shape)
    print(f"case: {batch} x {n}, {len(boxes)} boxes, precision {precision}: max |diff| "
          f"{float((got - want).abs().max()):.1e}", flush=True)
    assert torch.equal(got, want)


def test_noalloc_observables_match_the_plan_entry_point_twice(dev):
    # first call: 3 packets of 96 x 128 x 64, five boxes (two passes); second call smaller
    n = (96, 128, 64)
    boxes = [((1, 2, 3), (93, 125, 61)), ((0, 0, 0), (96, 128, 64)), ((10, 20, 5), (1, 3, 5)),
             ((50, 0, 1), (45, 127, 63)), ((5, 6, 7), (9, 11, 13))]
    _compare(n, (0.3, 0.25, 0.4), _abi.WFS_C64, 3, boxes, 1)
    _compare((32, 32, 32), (0.5, 0.5, 0.5), _abi.WFS_C128, 1, [((3, 4, 5), (21, 23, 25)), ((0, 0, 0), (1, 1, 1))], 2)


def test_noalloc_observables_refuse_requests_beyond_the_workspace(dev):
    n = (512, 512, 8)
    psi = torch.zeros((2,) + n, dtype=torch.complex64, device="cuda:0")
    out = torch.zeros(2 * 7, dtype=torch.float64, device="cuda:0")
    # 512 x 512 rows in 8192 blocks x 7 x 2 packets = 114688 records: fits
    _abi.observables_noalloc(n, (0.1, 0.1, 0.1), _abi.WFS_C64, psi, [((0, 0, 0), (512, 512, 8))], out, batch=2)
    # four boxes spanning the grid on 2 packets: 458752 records > 262144, refused before any launch
    boxes = [((0, 0, 0), (512, 512, 8))] * 4
    big = torch.zeros(2 * 4 * 7, dtype=torch.float64, device="cuda:0")
    with pytest.raises(_abi.WfsError, match="static"):
        _abi.observables_noalloc(n, (0.1, 0.1, 0.1), _abi.WFS_C64, psi, boxes, big, batch=2)
    torch.cuda.synchronize()
