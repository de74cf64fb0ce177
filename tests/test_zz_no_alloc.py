"""fsg_chem_integrate_flagged_noalloc (static device scratch, no allocation)
against fsg_chem_integrate_flagged (caller scratch) through the C ABI: two
calls in a row, the second on a smaller batch, each bit-identical to the
entry point with caller scratch, including the untouched non-flagged cells;
sizes above the static maximum are rejected before any launch."""

import ctypes

import numpy as np
import pytest
import torch

from hybrid import _native as nat
from hybrid.chemistry import Air11ParkMechanism
from hybrid.chemistry.batch import make_batch

pytestmark = pytest.mark.gpu

MAX_CELLS = 1 << 21


def _run(fn, table, rho, T, flag, n, scratch=None):
    lib = nat.lib()
    out = torch.full_like(rho, -1.0)
    Tout = torch.full_like(T, -1.0)
    stream = nat.stream_handle(rho.device)
    args = [nat.ptr(table), nat.ptr(rho), nat.ptr(T), None, ctypes.c_double(1e-7), 12, 0,
            nat.ptr(flag)]
    if scratch is not None:
        args += [nat.ptr(scratch[0]), nat.ptr(scratch[1])]
    args += [nat.ptr(out), nat.ptr(Tout), ctypes.c_int64(n), stream]
    nat.check_chem(getattr(lib, fn)(*args))
    torch.cuda.synchronize()
    return out, Tout


def test_noalloc_matches_caller_scratch_twice():
    table = Air11ParkMechanism().table
    g = torch.Generator().manual_seed(4)
    for n, seed in ((3000, 1), (700, 2)):          # second call smaller than the first
        print(f"case noalloc n={n}", flush=True)
        rho, T = make_batch(n, seed=seed)
        rho = torch.as_tensor(rho, device="cuda")
        T = torch.as_tensor(T, device="cuda")
        flag = (torch.rand(n, generator=g) < 0.3).to(torch.int32).cuda()
        scratch = (torch.empty(n, dtype=torch.int32, device="cuda"),
                   torch.empty(1, dtype=torch.int32, device="cuda"))
        want, wT = _run("fsg_chem_integrate_flagged", table, rho, T, flag, n, scratch)
        got, gT = _run("fsg_chem_integrate_flagged_noalloc", table, rho, T, flag, n)
        assert int(scratch[1].item()) == int(flag.sum())
        assert torch.equal(got, want) and torch.equal(gT, wT)
        f = flag.bool()
        assert bool((got[~f] == -1.0).all()) and bool((got[f] >= 0).all())


def test_noalloc_rejects_oversized_batch():
    lib = nat.lib()
    one = ctypes.c_void_p(1)
    rc = lib.fsg_chem_integrate_flagged_noalloc(one, one, one, None, ctypes.c_double(1e-7), 12, 0,
                                                one, one, None, ctypes.c_int64(MAX_CELLS + 1),
                                                None)
    assert rc == -13 and b"FSG_NOALLOC_MAX_CELLS" in lib.fsg_chem_last_error()
