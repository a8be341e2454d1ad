"""chromo_checkpoint / chromo_restore: k sweeps, a checkpoint, a restore into a
fresh replica table and m more sweeps give the chains of k + m sweeps bit for
bit -- positions, binder states, voxel densities, RNG words (glibc, MT19937,
Philox position), controllers and counters -- in both RNG modes and across
launch shapes; a checkpoint of other chains is refused."""
import copy

import numpy as np
import pytest

from chromo_amd import _lib as L
from chromo_amd import workloads as W

pytestmark = pytest.mark.gpu


def _dev(systems, **kw):
    from chromo_amd.engine import DeviceReplicas
    return DeviceReplicas([s.chain for s in systems], [s.field for s in systems],
                          [copy.deepcopy(s.ctrls) for s in systems], **kw)


def _same(a, b, n):
    for i in range(n):
        for x, y in zip(a.positions(i), b.positions(i)):
            assert np.array_equal(x, y)
        assert np.array_equal(a.binder_states(i), b.binder_states(i))
        assert np.array_equal(a.density_of(i), b.density_of(i))
        assert np.array_equal(a.rng_words(i), b.rng_words(i))
        assert a.controllers(i) == b.controllers(i)
    assert np.array_equal(a.counter_array(), b.counter_array())


@pytest.mark.parametrize("rng_mode,w_before,w_after", [(L.RNG_GLIBC, 1, 1), (L.RNG_PHILOX, 2, 2),
                                                       (L.RNG_GLIBC, 2, 1), (L.RNG_PHILOX, 1, 2)])
def test_checkpoint_restore_continues_bit_for_bit(tmp_path, rng_mode, w_before, w_after):
    print("checkpoint", rng_mode, w_before, w_after, flush=True)
    systems = [W.make_system(n, 2, g, seed=80 + i, chi=0.6 + 0.2 * i, mu=[-0.5, 0.4])
               for i, (n, g) in enumerate([(1500, 12), (700, 9), (3000, 16)])]
    kw = dict(rng_mode=rng_mode, seeds=[13, 14, 15], streams=[0, 1, 2])
    a = _dev(systems, **kw)
    if rng_mode == L.RNG_GLIBC:
        a.np_seed([7, 8, 9])
    a.run(6, waves=w_before)
    path = tmp_path / "ck.bin"
    a.checkpoint(path)
    a.run(5, waves=w_before)
    b = _dev(systems, **dict(kw, seeds=[99, 98, 97]))          # different streams until restored
    b.restore(path)
    b.run(5, waves=w_after)
    _same(a, b, 3)


def test_restore_refuses_other_chains(tmp_path):
    s = W.make_system(600, 2, 10, seed=90)
    a = _dev([s], rng_mode=L.RNG_PHILOX, seeds=[1], streams=[0])
    a.run(1)
    path = tmp_path / "ck.bin"
    a.checkpoint(path)
    other = _dev([W.make_system(601, 2, 10, seed=90)], rng_mode=L.RNG_PHILOX, seeds=[1], streams=[0])
    with pytest.raises(L.ChromoError, match="differ"):
        other.restore(path)
    glibc = _dev([s], rng_mode=L.RNG_GLIBC, seeds=[1])
    with pytest.raises(L.ChromoError, match="differ"):
        glibc.restore(path)
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"not a checkpoint")
    with pytest.raises(L.ChromoError, match="not a chromo checkpoint"):
        a.restore(bad)
