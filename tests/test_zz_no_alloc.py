"""sogrand_decode_cubic_noalloc (no allocation, static device scratch) through
the C ABI against sogrand_decode_cubic on the same device buffers: bits,
statistics and soft outputs identical, on two calls in a row (the second with
fewer codewords, so stale scratch from the first must not leak into it), and
its argument checks."""
import ctypes as C

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
    # more codewords than the static scratch takes: refused before any GPU work
    d = hip.DeviceArray((1, 4096))
    db = hip.DeviceArray((1, 64), np.uint8)
    assert lib.sogrand_decode_cubic_noalloc(decoder._h, C.c_void_p(d.data_ptr()),
                                            (1 << 20) + 1, C.c_void_p(db.data_ptr()),
                                            None, None, None) == sg.ETOOBIG
    # a non-finite input is reported as by sogrand_decode_cubic
    x = sg.gen_llr(3, 1, ebn0_db=2.0)
    x[0, 5] = np.nan
    d2 = hip.DeviceArray.from_host(x)
    with pytest.raises(sg.SograndError) as e:
        decoder.decode_noalloc(d2, db)
    assert e.value.code == sg.ENONFINITE
    assert lib.sogrand_decode_cubic_noalloc(decoder._h, None, 0, None, None, None, None) == sg.OK
