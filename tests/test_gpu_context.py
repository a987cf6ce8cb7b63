"""The context half of the boundary (device and stream, no counterpart in the reference): cdft_set_device /
cdft_get_device and cdft_system_set_stream.  A system whose work runs on another system's stream gives the same
results, bit for bit, as on its own, and returns to its own stream."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_device_selection(cdft):
    n = cdft.device_count()
    assert n >= 1
    cdft.set_device(0)
    assert cdft.get_device() == 0
    with pytest.raises(cdft.CdftError):
        cdft.set_device(n)
    with pytest.raises(cdft.CdftError):
        cdft.set_device(-1)
    assert cdft.get_device() == 0


def test_work_on_a_caller_stream(cdft):
    shape, dx = (32, 32, 32), 0.1
    rho = 0.5 * (1 + 0.1 * (2 * oracle.splitmix64_uniform(int(np.prod(shape)), 0x5EED8001).reshape(shape) - 1))
    a = cdft.System(shape, dx, 1.0, model=cdft.WHITEBEAR_I, ideal=True, mu=0.3)
    b = cdft.System(shape, dx, 1.0, model=cdft.WHITEBEAR_I, ideal=True, mu=0.3)
    p0, d0 = b.eval(rho)
    own = b.stream
    b.set_stream(a.stream)
    assert b.stream == a.stream and b.stream != own
    p1, d1 = b.eval(rho)
    f = cdft.Fire(b, rho)
    rc, rep = f.run(20, 1e-12)
    f.close()
    b.set_stream(None)
    assert b.stream == own
    p2, d2 = b.eval(rho)
    assert np.array_equal(p0, p1) and np.array_equal(d0, d1)
    assert np.array_equal(p0, p2) and np.array_equal(d0, d2)
    assert rc in (0, 1) and rep["F"] < p0[0]
    b.close()
    a.close()


def test_mixture_shares_one_stream_and_releases_it(cdft):
    mix = cdft.Mixture((16, 16, 16), 0.1, [1.0, 0.6])
    F, _ = mix.fmt_eval(np.full((2, 16, 16, 16), 0.2))
    assert np.isfinite(F)
    mix.close()
