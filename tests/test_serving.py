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


def test_wait_and_poll_on_two_threads(tm, orc):
    """One thread waits for every other request while another polls all completions: each request
    is collected exactly once.  A request the waiter has claimed is left alone by poll; one that
    poll collected before the waiter reached it makes that wait fail cleanly ("already
    collected"); nothing is freed twice, and every result is exact."""
    import threading
    s = "348864"
    pk, _ = keypairs(orc, s, 1)[0]
    PK = tm.PublicKey(s, pk)
    server = tm.Server(s, slot_capacity=2048, slots_per_direction=2)
    rng = np.random.default_rng(31)
    try:
        subs = []
        for i in range(120):
            sd = rng.integers(0, 256, (int(rng.integers(1, 200)), 32), dtype=np.uint8)
            rid, outs = server.submit("enc", PK, sd)
            subs.append((rid, sd, outs))
        all_ids = {rid for rid, _, _ in subs}
        waited, polled, refused, errors = set(), set(), set(), []
        done = threading.Event()

        def waiter():
            try:
                for rid, _, _ in subs[::2]:
                    try:
                        server.wait(rid)
                        waited.add(rid)
                    except tm.TmceError as e:
                        if "already collected" not in str(e):
                            raise
                        refused.add(rid)
            except Exception as e:  # surfaced below
                errors.append(e)
            finally:
                done.set()

        th = threading.Thread(target=waiter)
        th.start()
        while not errors and not (done.is_set() and all_ids <= waited | polled):
            ids, _ = server.poll()
            polled.update(ids)
        th.join()
    finally:
        server.close()
    assert not errors
    assert not (waited & polled) and waited | polled == all_ids
    assert refused <= polled and not (refused & waited)
    assert {rid for rid, _, _ in subs[1::2]} <= polled
    for _, sd, (ct, k, st) in subs:
        ct_a, k_a, st_a = PK.enc(sd)
        assert np.array_equal(ct, ct_a) and np.array_equal(k, k_a) and not st.any()


def _gpu_keypairs(tm, s, n, state):
    from oracle import oracle as orc
    seeds, _ = orc.xorshift_seeds(n, state)
    pks, sks = tm.keypair_batch(s, seeds)
    return list(zip(pks, sks))


def test_serving_8192128_16_keys_poisson(tm, orc):
    """The serving configuration at its real shape: mceliece8192128, 16 device-resident keypairs
    (the multi-batch plan's 16 key groups can fill), Poisson arrivals, request sizes log-uniform in
    [1, 4096], 50/50 enc/dec, keys chosen uniformly.  Every request equals the same request alone,
    decapsulations return the pool's session keys, and enc and dec samples equal the oracle."""
    from tensormceliece_amd import serving
    s = "8192128"
    kps = _gpu_keypairs(tm, s, 16, 0xC5)  # batched keypairs equal the host generator (test_keygen)
    wl = serving.Workload(s, kps, pool_per_key=4096)
    reqs = serving.make_requests(wl, 160, seed=51)
    assert max(r[2].shape[0] for r in reqs) > 2048 and len({r[1] for r in reqs}) == 16
    server = tm.Server(s, slot_capacity=32768, slots_per_direction=2, hbm_budget=8 << 30)
    try:
        print("serving 8192128, 16 keys, Poisson", flush=True)
        res = serving.run_poisson(server, wl, reqs, 800.0, keep_outputs=True)
        st = server.stats()
    finally:
        server.close()
    assert st["requests"] == len(reqs) and st["batches"] < len(reqs)
    assert st["peak_in_flight_bytes"] <= 8 << 30
    check_alone(tm, wl, reqs, res["outputs"])
    for (d, k, inp, off), (_, key, stt) in zip(reqs, res["outputs"]):
        if d == "dec":
            assert not stt.any() and np.array_equal(key, wl.pool_k[k][off:off + inp.shape[0]])
    ie = next(i for i, r in enumerate(reqs) if r[0] == "enc")
    ct_o, k_o, _ = orc.enc_batch(s, kps[reqs[ie][1]][0], reqs[ie][2][:4])
    assert np.array_equal(res["outputs"][ie][0][:4], ct_o) and np.array_equal(res["outputs"][ie][1][:4], k_o)
    idn = next(i for i, r in enumerate(reqs) if r[0] == "dec")
    k_o, _ = orc.dec_batch(s, kps[reqs[idn][1]][1], reqs[idn][2][:4])
    assert np.array_equal(res["outputs"][idn][1][:4], k_o)


def test_serving_17_keys_splits_batches(tm, orc):
    """17 resident keys with every request submitted at once: a batch holds at most 16 key groups,
    so a request under a 17th key ends the batch being formed and goes to the next one; results
    stay exact."""
    from tensormceliece_amd import serving
    s = "8192128"
    kps = _gpu_keypairs(tm, s, 17, 0x17)
    wl = serving.Workload(s, kps, pool_per_key=512)
    rng = np.random.default_rng(52)
    # three slot-filling requests per direction first: while they run one after the other (one slot
    # per direction), the small requests behind them queue up, so the batch of each direction that
    # follows them meets all 17 keys.  (With a single large request the small ones could still be
    # arriving when its batch completed: the worker holds the server's lock while it stages a
    # batch, which delays the submitting thread.)
    reqs = []
    for _ in range(3):
        reqs += [("enc", 0, rng.integers(0, 256, (32000, 32), dtype=np.uint8), None),
                 ("dec", 0, rng.integers(0, 256, (32000, tm.PARAMS[s]["ct"]), dtype=np.uint8), None)]
    for i in range(68):  # keys 0..16 round-robin, so every batch of 16 groups meets a 17th key
        k = i % 17
        n = int(rng.integers(1, 64))
        if i % 2 == 0:
            reqs.append(("enc", k, rng.integers(0, 256, (n, 32), dtype=np.uint8), None))
        else:
            off = int(rng.integers(0, 512 - n))
            reqs.append(("dec", k, wl.pool_ct[k][off:off + n], off))
    server = tm.Server(s, slot_capacity=32768, slots_per_direction=1)
    try:
        print("serving 8192128, 17 keys, all at once", flush=True)
        res = serving.run_poisson(server, wl, reqs, 0.0, keep_outputs=True)
        st = server.stats()
    finally:
        server.close()
    # 37 requests per direction under 17 keys: one slot per direction needs at least two batches each
    assert st["batches"] >= 4 and st["requests"] == len(reqs)
    assert st["max_batch_keys"] == 16  # a batch filled all 16 key groups and the 17th key waited
    check_alone(tm, wl, reqs, res["outputs"])
