"""The continuous-batching serving loop (sogrand_server_*) on the GPU.

The identity bar: every frame decodes to exactly the bits, statistics and soft
output it gets when decoded alone through sogrand_decode_cubic, whatever else
is in flight (the decoder is batch independent, so this is bit-exact).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sg():
    import sogrand_amd as sg

    return sg


def _say(*a):
    print(*a, flush=True)


def _frames(seed, n_frames, max_cw):
    sg = _sg()
    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, max_cw + 1, n_frames)
    snrs = rng.uniform(1.0, 3.0, n_frames)
    out = []
    off = 0
    for k in range(n_frames):
        llr = sg.gen_llr(seed, int(sizes[k]), ebn0_db=float(snrs[k]), cw_offset=off)
        off += int(sizes[k])
        out.append(llr)
    return out


@pytest.mark.parametrize("precision,lanegroup_max", [(0, 0), (1, 0), (0, 1 << 20), (1, 1 << 20)],
                         ids=["fp64", "fp32", "fp64-lanegroup", "fp32-lanegroup"])
def test_frames_identical_alone_and_in_the_mix(decoder, precision, lanegroup_max):
    # lanegroup_max: batches run the lane-group kernel whenever the in-flight
    # codewords allow (FP64; the FP32 server ignores it and runs its kernels)
    sg = _sg()
    from sogrand_amd import hip

    decoder.set_params(backend=sg.BACKEND_PERSISTENT, precision=precision)
    # frames of 1-256 codewords, as the serving configuration draws them
    frames = _frames(2024 + precision, 24, 256)
    alone = [decoder.decode(f, soft=True) for f in frames]
    _say("frames:", len(frames), "codewords:", sum(len(f) for f in frames),
         "largest:", max(len(f) for f in frames))
    # batches of at most 384 codewords and four slots: several in flight at once
    with sg.Server(decoder, hbm_budget=2 << 30, max_inflight=4, max_batch_cw=384,
                   with_soft=True, lanegroup_max_cw=lanegroup_max) as srv:
        outs = []
        for k, f in enumerate(frames):
            n = len(f)
            bits = np.zeros((n, 64), np.uint8)
            st = np.zeros(n, sg.CW_STATS_DTYPE)
            soft = np.zeros((n, 4096))
            src = hip.DeviceArray.from_host(f) if k % 3 == 0 else f  # device and host inputs
            outs.append((bits, st, soft, src))
            assert srv.submit(k, src, bits, st, soft) == sg.OK
        srv.drain()
        done = srv.poll(max_n=len(frames))
    assert sorted(c.id for c in done) == list(range(len(frames)))
    assert all(c.status == sg.OK for c in done)
    overlapped = sum(1 for c in done if c.t_start < max(d.t_done for d in done if d.id != c.id))
    _say("completions:", len(done), "launched while others were in flight:", overlapped)
    for k, (bits, st, soft, _) in enumerate(outs):
        assert (bits == alone[k].bits).all(), k
        assert (st == alone[k].stats).all(), k
        assert (soft == alone[k].soft).all(), k
    decoder.set_params(precision=sg.FP64)


def test_admission_and_errors(decoder):
    sg = _sg()
    decoder.set_params(backend=sg.BACKEND_PERSISTENT, precision=sg.FP64)
    good = sg.gen_llr(5, 8, ebn0_db=2.5)
    bad = good.copy()
    bad[3, 100] = np.nan
    with sg.Server(decoder, hbm_budget=256 << 20, max_inflight=2, max_batch_cw=16,
                   max_pending_cw=64) as srv:
        assert srv.batch_cw == 16
        big = sg.gen_llr(6, 17, ebn0_db=2.5)
        with pytest.raises(sg.SograndError) as e:
            srv.submit(99, big, np.zeros((17, 64), np.uint8))
        assert e.value.code == sg.EINVAL
        bits_g = np.zeros((8, 64), np.uint8)
        bits_b = np.zeros((8, 64), np.uint8)
        assert srv.submit(1, good, bits_g) == sg.OK
        assert srv.submit(2, bad, bits_b) == sg.OK
        srv.drain()
        res = {c.id: c for c in srv.poll()}
        # a frame with a NaN decodes as the reference's arithmetic decodes it
        assert res[1].status == sg.OK and res[2].status == sg.OK
        assert (bits_g == decoder.decode(good).bits).all()
        assert (bits_b == decoder.decode(bad).bits).all()
        # admission control: more waiting codewords than max_pending_cw is refused
        frames = [(sg.gen_llr(7, 16, ebn0_db=1.0, cw_offset=16 * k), np.zeros((16, 64), np.uint8))
                  for k in range(40)]
        refused = 0
        for k, (f, b) in enumerate(frames):
            refused += srv.submit(100 + k, f, b) == sg.EBUSY
        srv.drain()
        _say("refused while the queue was full:", refused)
        assert refused > 0
