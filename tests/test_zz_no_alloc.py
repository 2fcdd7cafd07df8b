"""plum_cbmc_trials_noalloc (scratch in a file-scope device array, no
allocation) against plum_cbmc_trials on the same context and inputs, through
the C ABI: bit-equal energies on two calls in a row, the second smaller, and
the capacity limits rejected with an error."""
import ctypes as C

import numpy as np
import pytest

from plum_amd import engine as E
from plum_amd import systems

pytestmark = pytest.mark.gpu


def _noalloc(eng, rq, tri, grown, K):
    out = np.zeros(K)
    rc = E.lib().plum_cbmc_trials_noalloc(eng.h, C.byref(rq), E._dp(tri), E._dp(grown), E._dp(out))
    return rc, out


def _inputs(sd, K, cur, seed):
    rng = np.random.default_rng(seed)
    L = np.asarray(sd["box"])
    grown = rng.uniform(0, 1, (cur, 3)) * L
    gions = rng.uniform(0, 1, (cur, 3)) * L
    tri = rng.uniform(0, 1, (K, 3)) * L
    ions = rng.uniform(0, 1, (K, 3)) * L
    return tri, ions, grown, gions


@pytest.mark.parametrize("name", ["small", "c0", "slab_small"])
def test_noalloc_trials_equal_the_allocating_entry(name):
    sd = systems.config(name)
    eng = E.Engine(sd)
    for K, cur, seed in ((16, 7, 1), (5, 2, 2)):
        tri, ions, grown, gions = _inputs(sd, K, cur, seed)
        ref = eng.cbmc_trials(tri, 0, -1.0, grown=grown, ions=ions, type_ion=1, q_ion=1.0,
                              grown_ions=gions, exclude=(3, 9))
        rq = E.CbmcRequest(K, cur, 3, 9, 0, 1, 1, -1.0, 1.0)
        t = np.ascontiguousarray(np.concatenate([tri, ions]))
        g = np.ascontiguousarray(np.concatenate([grown, gions]))
        rc, got = _noalloc(eng, rq, t, g, K)
        print(name, K, cur, rc, got[:4], ref[:4], flush=True)
        assert rc == 0, E.lib().plum_last_error()
        assert np.array_equal(got, ref), (got, ref)
    # beyond the scratch limits: an error, nothing launched
    rq = E.CbmcRequest(65, 0, 0, -1, 0, 1, 0, -1.0, 0.0)
    t = np.zeros(3 * 65)
    rc, _ = _noalloc(eng, rq, t, np.zeros(3), 65)
    assert rc == -5
    assert "without allocation" in E.lib().plum_last_error().decode()
    eng.close()
