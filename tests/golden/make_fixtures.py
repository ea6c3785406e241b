"""Writes the small RNG fixtures the oracle is pinned against.

philox4x32_10_kat.json: the published Random123 known-answer vectors for
  philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11; kat_vectors), typed in
  from the publication -- not computed here.
xorshift64star.json: the reference's Rng stream (crates/omg-core/src/rng.rs:
  9-21: shifts 12/25/27, multiplier 0x2545F4914F6CDD1D, top 24 bits / 2^24)
  computed with Python integers, independently of the C oracle.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = (1 << 64) - 1

PHILOX_KAT = [
    dict(ctr=[0, 0, 0, 0], key=[0, 0],
         out=[0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    dict(ctr=[0xffffffff] * 4, key=[0xffffffff] * 2,
         out=[0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    dict(ctr=[0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], key=[0xa4093822, 0x299f31d0],
         out=[0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


def xorshift(seed, n):
    x = max(seed, 1)
    u64, f32 = [], []
    for _ in range(n):
        x ^= x >> 12
        x ^= (x << 25) & M64
        x ^= x >> 27
        v = (x * 0x2545F4914F6CDD1D) & M64
        u64.append(v)
        f32.append(float(np.float32(v >> 40) / np.float32(1 << 24)))
    return u64, f32


def main():
    with open(os.path.join(HERE, "philox4x32_10_kat.json"), "w") as f:
        json.dump(PHILOX_KAT, f, indent=1)
    cases = []
    for seed in (0xC0FFEE, 7, 0x5EED5EED, 0):
        u, fl = xorshift(seed, 16)
        cases.append(dict(seed=seed, u64=u, f32=fl))
    with open(os.path.join(HERE, "xorshift64star.json"), "w") as f:
        json.dump(cases, f, indent=1)


if __name__ == "__main__":
    main()
