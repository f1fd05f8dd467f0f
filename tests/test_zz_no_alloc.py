"""mimadv_batch_budget_device, the budget entry point that allocates
nothing (its partials live in a static device array of the library), called
through the C ABI and compared with mimadv_batch_budget, bit for bit, on two
calls in a row: a 20-field batch on one grid, then 3 fields on a smaller
grid with another context, which reuses the same device array."""
import ctypes as C

import numpy as np
import pytest
import torch

from fdd3260_amd import Advection, _lib, cases
from fdd3260_amd import grid as gm

pytestmark = pytest.mark.gpu


def _fields(g, cuda, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for f in range(n):
        q = cases.lognormal(g, seed=seed + f, mu=-6.0, sigma=1.0, zero_frac=0.2)
        dt = np.float64 if f % 3 == 0 else np.float32
        out.append(torch.from_numpy(q.astype(dt)).to(cuda))
    return out


def test_budget_device_matches_budget_two_calls(cuda):
    print("case no-alloc budget: 20 fields 48x40x30, then 3 fields 20x16x12", flush=True)
    for maxx, maxy, nz, nf, seed in ((48, 40, 30, 20, 1), (20, 16, 12, 3, 50)):
        g = gm.stretched(maxx, maxy, nz)
        with Advection(g) as a:
            fields = _fields(g, cuda, nf, seed)
            want = a.batch_budget(fields)
            out = torch.full((nf,), float("nan"), dtype=torch.float64, device=cuda)
            a.batch_budget_device(fields, out)
            a.stream_sync()
            got = out.cpu().numpy()
            for f in range(nf):
                assert got[f] == want[f], (maxx, f, got[f], want[f])


def test_budget_device_rejects_too_many_fields(cuda):
    g = gm.stretched(20, 16, 12)
    with Advection(g) as a:
        fields = _fields(g, cuda, 33, 7)
        out = torch.zeros(33, dtype=torch.float64, device=cuda)
        with pytest.raises(_lib.MimadvError, match="at most 32 fields"):
            a.batch_budget_device(fields, out)
