"""Serving loop (tmce_server_*): continuous batching of mixed enc/dec requests under several keys.

Bar: every request's outputs are byte-identical to the same request run alone through the plain
batched calls (which the other GPU tests pin to the oracle), whether the requests arrive all at once
or spread out (Poisson), and the HBM-footprint admission keeps the bytes in flight within budget.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_KP = {}


@pytest.fixture(scope="module")
def tm():
    import tensormceliece_amd as tm
    tm.init(0)
    return tm


def keypairs(orc, s, n):
    if (s, n) not in _KP:
        seeds, _ = orc.xorshift_seeds(n, 0x1234567)
        _KP[(s, n)] = [orc.keypair_seeded(s, seeds[i].tobytes()) for i in range(n)]
    return _KP[(s, n)]


def check_alone(tm, wl, reqs, outputs):
    for (d, k, inp, _), (ct, key, st) in zip(reqs, outputs):
        if d == "enc":
            ct_a, k_a, st_a = wl.pk[k].enc(inp)
            assert np.array_equal(ct, ct_a) and np.array_equal(key, k_a) and np.array_equal(st, st_a)
        else:
            k_a, st_a = wl.sk[k].dec(inp)
            assert np.array_equal(key, k_a) and np.array_equal(st, st_a)


@pytest.mark.parametrize("rate", [0.0, 4000.0])
def test_same_alone_and_in_mix(tm, orc, rate):
    from tensormceliece_amd import serving
    s = "348864"
    wl = serving.Workload(s, keypairs(orc, s, 4), pool_per_key=1024)
    reqs = serving.make_requests(wl, 160, seed=11, hi=1024)
    server = tm.Server(s, slot_capacity=4096, slots_per_direction=2)
    try:
        print(f"serving mix rate={rate}", flush=True)
        res = serving.run_poisson(server, wl, reqs, rate, keep_outputs=True)
        st = server.stats()
    finally:
        server.close()
    assert st["requests"] == len(reqs) and st["batches"] < len(reqs)  # requests were batched together
    check_alone(tm, wl, reqs, res["outputs"])
    # decapsulated keys equal the keys of the pool's encapsulations; a sample against the oracle
    for (d, k, inp, off), (_, key, stt) in zip(reqs, res["outputs"]):
        if d == "dec":
            assert not stt.any() and np.array_equal(key, wl.pool_k[k][off:off + inp.shape[0]])
    d0 = next(i for i, r in enumerate(reqs) if r[0] == "enc")
    ct_o, k_o, _ = orc.enc_batch(s, keypairs(orc, s, 4)[reqs[d0][1]][0], reqs[d0][2][:16])
    assert np.array_equal(res["outputs"][d0][0][:16], ct_o) and np.array_equal(res["outputs"][d0][1][:16], k_o)


def test_hbm_budget_admission(tm, orc):
    """A budget of about one slot's worth of ops: batches are admitted one at a time, bytes in
    flight stay within the budget, and every result is still exact."""
    from tensormceliece_amd import serving
    s = "8192128"
    wl = serving.Workload(s, keypairs(orc, s, 2), pool_per_key=2048)
    reqs = serving.make_requests(wl, 40, seed=12, hi=1500)
    probe = tm.Server(s, slot_capacity=4096, slots_per_direction=2)
    per_op = max(probe.stats()["enc_bytes_per_op"], probe.stats()["dec_bytes_per_op"])
    probe.close()
    budget = 1600 * per_op
    server = tm.Server(s, slot_capacity=4096, slots_per_direction=2, hbm_budget=budget)
    try:
        res = serving.run_poisson(server, wl, reqs, 0.0, keep_outputs=True)
        st = server.stats()
    finally:
        server.close()
    assert st["peak_in_flight_bytes"] <= budget
    check_alone(tm, wl, reqs, res["outputs"])


def test_request_errors(tm, orc):
    s = "348864"
    pk, sk = keypairs(orc, s, 1)[0]
    server = tm.Server(s, slot_capacity=1024, slots_per_direction=1)
    try:
        with pytest.raises(tm.TmceError):  # larger than a slot can take
            server.submit("enc", tm.PublicKey(s, pk), np.zeros((1000, 32), np.uint8))
        with pytest.raises(tm.TmceError):  # secret key for encapsulation
            server.submit("enc", tm.SecretKey(s, sk), np.zeros((4, 32), np.uint8))
        rid, (ct, k, st) = server.submit("enc", tm.PublicKey(s, pk), np.zeros((4, 32), np.uint8))
        assert server.wait(rid) > 0 and not st.any()
    finally:
        server.close()


@pytest.mark.parametrize("s", ["348864", "6960119pc"])
def test_zero_copy_requests(tm, orc, s):
    """Requests whose buffers lie in pinned slabs (tmce_host_alloc) are served in place -- the GPU
    reads the inputs and writes the outputs across the host link -- with the same bytes as staged
    requests; a slab request with a misaligned output buffer falls back to staging."""
    from tensormceliece_amd import serving
    wl = serving.Workload(s, keypairs(orc, s, 3), pool_per_key=1024)
    reqs = serving.make_requests(wl, 90, seed=13, hi=1024)
    zreqs, inb, outb = serving.to_slab(reqs, tm.PARAMS[s])
    server = tm.Server(s, slot_capacity=4096, slots_per_direction=2)
    try:
        print(f"zero-copy serving {s}", flush=True)
        staged = serving.run_poisson(server, wl, reqs, 0.0, keep_outputs=True)
        zc = serving.run_poisson(server, wl, zreqs, 0.0, keep_outputs=True, slab=outb)
        st = server.stats()
        # a slab request with k 8 bytes off the required 16-byte alignment: staged, still exact
        d, k, inp, _ = zreqs[0]
        kbuf = outb.array(inp.shape[0] * 32 + 8, np.uint8)[8:].reshape(-1, 32)
        key = wl.pk[k] if d == "enc" else wl.sk[k]
        rid, outs = server.submit(d, key, inp, k=kbuf, slab=outb)
        server.wait(rid)
        st2 = server.stats()
    finally:
        server.close()
    assert st["zero_copy_requests"] == len(reqs)
    assert st2["zero_copy_requests"] == len(reqs)  # the misaligned one was staged
    for a, b in zip(staged["outputs"], zc["outputs"]):
        for x, y in zip(a, b):
            assert (x is None and y is None) or np.array_equal(x, y)
    check_alone(tm, wl, zreqs, zc["outputs"])
    assert np.array_equal(outs[1], zc["outputs"][0][1])
    inb.free()
    outb.free()
