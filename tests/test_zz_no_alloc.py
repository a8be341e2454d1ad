"""chromo_energy_adjoint_noalloc, the allocation-free variant of
chromo_energy_adjoint (its scratch is a file-scope device array), against the
allocating entry point through the C ABI: E, dE/dr and dE/dtheta bit for bit
on two calls in a row, the second with a smaller chain and grid, and a clear
error for an evaluation larger than the array."""
import ctypes

import numpy as np
import pytest
import torch

from chromo_amd import _lib as L
from chromo_amd import workloads as W

pytestmark = pytest.mark.gpu


def _noalloc(model, r, th, grad=True):
    from chromo_amd.engine import _current_stream
    E = torch.zeros(4, dtype=torch.float64, device=model.device)
    gr = torch.full_like(r, float("nan")) if grad else None
    gt = torch.full_like(th, float("nan")) if grad else None
    rc = model.lib.chromo_energy_adjoint_noalloc(ctypes.byref(model.m), r.data_ptr(), th.data_ptr(),
                                                 E.data_ptr(), gr.data_ptr() if grad else None,
                                                 gt.data_ptr() if grad else None, _current_stream())
    torch.cuda.synchronize()
    return rc, E, gr, gt


@pytest.mark.parametrize("grad", [True, False])
def test_noalloc_equals_allocating_entry_point(grad):
    from chromo_amd.energy import EnergyModel, energy
    for N, B, nvox, seed in [(20000, 3, 48, 4), (3000, 2, 12, 5)]:
        print("noalloc", N, B, nvox, grad, flush=True)
        s = W.make_system(N, B, nvox, seed=seed, eps=[-4.0, -2.5, -3.0][:B], mu=[-0.4, 0.3, -1.0][:B],
                          cross_eps=0.6)
        s.chain.states = np.random.RandomState(seed).randint(0, 3, size=(N, B))
        model = EnergyModel(s.chain, s.field, s.binders)
        r = torch.tensor(s.chain.r, dtype=torch.float64, device="cuda")
        th = torch.tensor(model.theta0, dtype=torch.float64, device="cuda")
        E0, g0, t0 = energy(model, r, th, grad=grad)
        rc, E1, g1, t1 = _noalloc(model, r, th, grad=grad)
        L.check(rc, "chromo_energy_adjoint_noalloc")
        assert torch.equal(E0, E1)
        if grad:
            assert torch.equal(g0, g1) and torch.equal(t0, t1)


def test_noalloc_rejects_an_evaluation_larger_than_its_array():
    from chromo_amd.energy import EnergyModel
    s = W.make_system(200, 4, 128, seed=6, eps=-3.0, mu=-0.5)
    model = EnergyModel(s.chain, s.field, s.binders)
    r = torch.tensor(s.chain.r, dtype=torch.float64, device="cuda")
    th = torch.tensor(model.theta0, dtype=torch.float64, device="cuda")
    rc, _, _, _ = _noalloc(model, r, th)
    assert rc == -1
    buf = ctypes.create_string_buffer(512)
    model.lib.chromo_last_error(buf, 512)
    assert b"bytes of scratch" in buf.value
