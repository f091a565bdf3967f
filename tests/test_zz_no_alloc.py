"""vb_logconf_noalloc: the log-conformation update without any device allocation (its kick table in
the library's fixed device scratch array), against vb_logconf (which allocates the same table with
hipMalloc).  Two calls in a row on one GPU, the second with fewer particles, so the second reads a
scratch array that still holds the first call's larger table: both must equal vb_logconf bit for bit
(same kernels, same inputs).  Also: sizes beyond the scratch are refused, and kick = 0 needs no
scratch."""
import numpy as np
import pytest

from tests.hipmem import DeviceArray as D

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    tau = np.zeros((n, 4))
    tau[:, 0] = rng.uniform(-0.2, 0.6, n); tau[:, 3] = rng.uniform(-0.2, 0.6, n)
    tau[:, 1] = tau[:, 2] = rng.uniform(-0.3, 0.3, n)
    tau[: n // 16] = 0.0                              # the A = I start
    trc = 2.0 + tau[:, 0] + tau[:, 3]
    gv = rng.uniform(-2, 2, (n, 4))
    return tau, trc, gv


def _run(fn, tau, trc, gv, kick):
    n = tau.shape[0]
    a = [D(tau), D(trc), D(gv)]
    to, tro = D(np.zeros((n, 4))), D(np.zeros(n))
    st = fn(n, *[x.ptr for x in a], 1e-3, 0.8, 0.1, kick, to.ptr, tro.ptr)
    return st, to.numpy(), tro.numpy()


def test_logconf_noalloc_matches_allocating_entry(gpu_lib):
    L = gpu_lib
    for n, seed in ((200_000, 1), (5_003, 2)):           # second call smaller than the first
        tau, trc, gv = _inputs(n, seed)
        for kick in (1, 0):
            st_a, to_a, tr_a = _run(L.vb_logconf, tau, trc, gv, kick)
            st_b, to_b, tr_b = _run(L.vb_logconf_noalloc, tau, trc, gv, kick)
            print(f"n {n} kick {kick}: status {st_a}/{st_b}", flush=True)
            assert st_a == 0 and st_b == 0
            assert np.array_equal(to_a, to_b) and np.array_equal(tr_a, tr_b)
    # the scratch holds 1,048,576 particles: more is refused (VB_ERR_CAPACITY) before any launch
    assert L.vb_logconf_noalloc(1_048_577, None, None, None, 1e-3, 0.8, 0.0, 1, None, None) == 4
    assert L.vb_logconf_noalloc(-1, None, None, None, 1e-3, 0.8, 0.0, 1, None, None) == 8
