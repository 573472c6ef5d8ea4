"""GPU: the asynchronous device entry points on caller streams.

The engine keeps one scratch set per direction.  Encapsulations and decapsulations issued on two
streams with events between them (the headline's two-stream step) must give the same bytes as the
same calls serialised on one stream, and two calls of one direction on different streams (which
share that direction's scratch) must each give their own correct result.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SET = "348864"


@pytest.fixture(scope="module")
def tm():
    import tensormceliece_amd as tm
    tm.init(0)
    return tm


@pytest.fixture(scope="module")
def keys(orc):
    kseed, st = orc.xorshift_seeds(1)
    pk, sk = orc.keypair_seeded(SET, kseed[0].tobytes())
    return pk, sk, st


def test_two_stream_steps_match_serial(tm, orc, keys):
    pk, sk, st = keys
    p = tm.PARAMS[SET]
    n, steps = 4096, 5
    PK, SK = tm.PublicKey(SET, pk), tm.SecretKey(SET, sk)
    rng = np.random.default_rng(7)
    seeds = [rng.integers(0, 256, (n, 32), dtype=np.uint8) for _ in range(steps)]  # a batch per step
    d_seeds = []
    for sd in seeds:
        b = tm.DeviceBuffer(n * 32)
        b.upload(sd)
        d_seeds.append(b)
    # serial reference on the engine stream
    ref = []
    d_ct, d_k, d_k2 = tm.DeviceBuffer(n * p["ct"]), tm.DeviceBuffer(n * 32), tm.DeviceBuffer(n * 32)
    for i in range(steps):
        tm.enc_device(PK, d_ct, d_k, None, d_seeds[i], n)
        tm.dec_device(SK, d_k2, None, d_ct, n)
        tm.sync()
        ref.append((d_ct.download((n, p["ct"])), d_k.download((n, 32)), d_k2.download((n, 32))))
    # two streams, double-buffered, encapsulation i+1 concurrent with decapsulation i
    se, sdec = tm.Stream(), tm.Stream()
    ct = [tm.DeviceBuffer(n * p["ct"]) for _ in range(2)]
    k = [tm.DeviceBuffer(n * 32) for _ in range(2)]
    k2 = [tm.DeviceBuffer(n * 32) for _ in range(2)]
    ev_e, ev_d = [tm.Event(), tm.Event()], [tm.Event(), tm.Event()]
    got = []
    print(f"two-stream steps {SET} n={n} x {steps}", flush=True)
    for i in range(steps):
        b = i & 1
        if i >= 2:
            se.wait(ev_d[b])
        tm.enc_device(PK, ct[b], k[b], None, d_seeds[i], n, stream=se)
        ev_e[b].record(se)
        sdec.wait(ev_e[b])
        tm.dec_device(SK, k2[b], None, ct[b], n, stream=sdec)
        ev_d[b].record(sdec)
        if i >= 1:  # step i-1's buffers are final once its decapsulation is done
            sdec.sync()
            c = (i - 1) & 1
            got.append((ct[c].download((n, p["ct"])), k[c].download((n, 32)), k2[c].download((n, 32))))
    se.sync()
    sdec.sync()
    c = (steps - 1) & 1
    got.append((ct[c].download((n, p["ct"])), k[c].download((n, 32)), k2[c].download((n, 32))))
    for i in range(steps):
        for a, r in zip(got[i], ref[i]):
            assert np.array_equal(a, r), f"step {i}"
        assert np.array_equal(got[i][1], got[i][2])
    ct_o, k_o, _ = orc.enc_batch(SET, pk, seeds[0][:64])
    assert np.array_equal(got[0][0][:64], ct_o) and np.array_equal(got[0][1][:64], k_o)
    for x in ct + k + k2 + [d_ct, d_k, d_k2] + d_seeds:
        x.free()
    for e in ev_e + ev_d:
        e.free()
    se.free()
    sdec.free()


@pytest.mark.parametrize("direction", ["enc", "dec"])
def test_same_direction_on_two_streams(tm, orc, keys, direction):
    pk, sk, st = keys
    p = tm.PARAMS[SET]
    PK, SK = tm.PublicKey(SET, pk), tm.SecretKey(SET, sk)
    n1, n2 = 8192, 3000
    rng = np.random.default_rng(11 if direction == "enc" else 12)
    s1 = rng.integers(0, 256, (n1, 32), dtype=np.uint8)
    s2 = rng.integers(0, 256, (n2, 32), dtype=np.uint8)
    ct1, k1, _ = PK.enc(s1)
    ct2, k2, _ = PK.enc(s2)
    sa, sb = tm.Stream(), tm.Stream()
    print(f"{direction} on two streams {SET} n={n1},{n2}", flush=True)
    if direction == "enc":
        bufs = []
        for sd, nn, stream in ((s1, n1, sa), (s2, n2, sb)):
            d_s, d_c, d_k = tm.DeviceBuffer(nn * 32), tm.DeviceBuffer(nn * p["ct"]), tm.DeviceBuffer(nn * 32)
            d_s.upload(sd)
            tm.enc_device(PK, d_c, d_k, None, d_s, nn, stream=stream)
            bufs.append((d_s, d_c, d_k, nn))
        sa.sync()
        sb.sync()
        for (d_s, d_c, d_k, nn), (c_ref, k_ref) in zip(bufs, ((ct1, k1), (ct2, k2))):
            assert np.array_equal(d_c.download((nn, p["ct"])), c_ref)
            assert np.array_equal(d_k.download((nn, 32)), k_ref)
            for x in (d_s, d_c, d_k):
                x.free()
    else:
        bufs = []
        for c, nn, stream in ((ct1, n1, sa), (ct2, n2, sb)):
            d_c, d_k = tm.DeviceBuffer(nn * p["ct"]), tm.DeviceBuffer(nn * 32)
            d_c.upload(c)
            tm.dec_device(SK, d_k, None, d_c, nn, stream=stream)
            bufs.append((d_c, d_k, nn))
        sa.sync()
        sb.sync()
        for (d_c, d_k, nn), k_ref in zip(bufs, (k1, k2)):
            assert np.array_equal(d_k.download((nn, 32)), k_ref)
            for x in (d_c, d_k):
                x.free()
    sa.free()
    sb.free()
