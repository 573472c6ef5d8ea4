"""Writes tests/golden/kem_vectors.json: small seeded KEM vectors from the CPU oracle.

The oracle is pinned by libmceliece's KAT checksums (tests/test_oracle.py), so these vectors are
libmceliece-20250507's outputs for: the keypair from randombytes = seeded chain on the first 32
xorshift64 seed bytes of the reference harness (mceliece-gpu-params/main.cu:75-76), the next four
seeded encapsulations, and decapsulations of two random ciphertexts.  pk/sk are stored as SHA-256.
Run: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as orc  # noqa: E402


def main():
    out = {}
    for s in orc.SETS:
        kseed, st = orc.xorshift_seeds(1)
        pk, sk = orc.keypair_seeded(s, kseed[0].tobytes())
        seeds, _ = orc.xorshift_seeds(4, st)
        ct, k, ret = orc.enc_batch(s, pk, seeds)
        rng = np.random.default_rng(2024)
        rct = rng.integers(0, 256, size=(2, orc.params(s).synd_bytes), dtype=np.uint8)
        if s == "6960119":
            rct[:, -1] &= 0x07  # keep the padding bits zero
        rk, rret = orc.dec_batch(s, sk, rct)
        out[s] = {
            "keypair_seed": kseed[0].tobytes().hex(),
            "pk_sha256": hashlib.sha256(pk.tobytes()).hexdigest(),
            "sk_sha256": hashlib.sha256(sk.tobytes()).hexdigest(),
            "enc": [{"seed": seeds[i].tobytes().hex(), "ct": ct[i].tobytes().hex(), "k": k[i].tobytes().hex()}
                    for i in range(len(seeds))],
            "dec_random": [{"ct": rct[i].tobytes().hex(), "k": rk[i].tobytes().hex(), "ret": int(rret[i])}
                           for i in range(len(rct))],
        }
    with open(os.path.join(HERE, "kem_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
