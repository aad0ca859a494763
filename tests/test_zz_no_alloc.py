"""sogrand_decode_cubic_noalloc (no allocation, static device scratch) through
the C ABI against sogrand_decode_cubic on the same device buffers: bits,
statistics and soft outputs identical, on two calls in a row (the second with
fewer codewords, so stale scratch from the first must not leak into it), its
argument checks and non-finite inputs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _decode_both(decoder, n, seed):
    import sogrand_amd as sg
    from sogrand_amd import hip

    ebn0 = np.random.default_rng(seed).uniform(0.8, 3.5, n)
    llr = hip.DeviceArray((n, 4096))
    decoder.gen_llr_device(seed, 0, n, [sg.sigma_for_ebn0(float(e)) for e in ebn0], llr)
    out = []
    for noalloc in (False, True):
        bits = hip.DeviceArray((n, 64), np.uint8)
        st = hip.DeviceArray((n,), sg.CW_STATS_DTYPE)
        soft = hip.DeviceArray((n, 4096))
        bits.zero()
        st.zero()
        if noalloc:
            decoder.decode_noalloc(llr, bits, st, soft)
        else:
            decoder.decode(llr, out_bits=bits, out_stats=st, out_soft=soft)
        out.append((bits.to_host(), st.to_host(), soft.to_host()))
        soft.free()
    llr.free()
    return out


@pytest.mark.parametrize("backend", [0, 1], ids=["plain", "persistent"])
def test_noalloc_matches_decode_cubic(decoder, backend):
    decoder.set_params(backend=backend)
    for n, seed in ((1500, 41), (300, 42)):
        _say("noalloc: backend", backend, "n", n)
        (b0, s0, f0), (b1, s1, f1) = _decode_both(decoder, n, seed)
        assert (b0 == b1).all()
        assert (s0 == s1).all()
        assert (f0 == f1).all()
        assert s0["half_iters"].min() >= 1


def test_noalloc_argument_checks(decoder):
    import sogrand_amd as sg
    from sogrand_amd import hip

    lib = decoder._lib
    llr = np.zeros((2, 4096))
    bits = np.zeros((2, 64), np.uint8)
    # host buffers are refused (they would need staging memory)
    assert lib.sogrand_decode_cubic_noalloc(decoder._h, llr.ctypes.data, 2, bits.ctypes.data,
                                            None, None, None) == sg.EINVAL
    # non-finite inputs decode as sogrand_decode_cubic decodes them
    x = sg.gen_llr(3, 2, ebn0_db=2.0)
    x[0, 5] = np.nan
    x[1, 70] = -np.inf
    d2 = hip.DeviceArray.from_host(x)
    db = hip.DeviceArray((2, 64), np.uint8)
    ds = hip.DeviceArray((2,), sg.CW_STATS_DTYPE)
    decoder.decode_noalloc(d2, db, ds)
    ref = decoder.decode(x)
    assert (db.to_host() == ref.bits).all() and (ds.to_host() == ref.stats).all()
    assert lib.sogrand_decode_cubic_noalloc(decoder._h, None, 0, None, None, None, None) == sg.OK
