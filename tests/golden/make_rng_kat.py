"""Writes tests/golden/rng_kat.json: known-answer vectors of the reference RNG.

Evaluates shader/HLSL/random.hlsl:1-20 (InitRandomSeed, NextRandom,
RandomFloat) with Python integers reduced mod 2^32 after each operation, i.e.
independently of the oracle's C++ uint32 arithmetic.  The reference ships no
vectors of its own.

    python tests/golden/make_rng_kat.py
"""
import json
from fractions import Fraction
from pathlib import Path

M = 1 << 32


def init_seed(x, y, w, frame):
    s = (x * 1 + y * w) % M
    s = (s ^ 61) ^ (s >> 16)
    s = (s * 9) % M
    s = s ^ (s >> 4)
    s = (s * 0x27D4EB2D) % M
    s = s ^ (s >> 15)
    return (s * (frame + 1)) % M


def next_random(state):
    state = (state * 747796405 + 2891336453) % M
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) % M
    return state, (word >> 22) ^ word


def main():
    seeds = []
    for (x, y, w, frame) in [(0, 0, 1, 0), (1, 0, 640, 1), (639, 359, 640, 16), (17, 5, 1920, 1),
                             (1919, 1079, 1920, 9999), (960, 540, 1920, 128), (3, 4, 7, 4294967294)]:
        seeds.append(dict(x=x, y=y, w=w, frame=frame, seed=init_seed(x, y, w, frame)))
    streams = []
    for st in [0, 1, 123456789, 0xDEADBEEF, init_seed(100, 200, 1920, 3)]:
        s = st
        words, floats = [], []
        for _ in range(8):
            s, word = next_random(s)
            words.append(word)
            # RandomFloat: (word & 0xFFFFFF) / 2^24, exact in binary32
            floats.append(float(Fraction(word & 0xFFFFFF, 1 << 24)))
        streams.append(dict(state=st, words=words, floats=floats))
    out = Path(__file__).with_name("rng_kat.json")
    out.write_text(json.dumps(dict(source="shader/HLSL/random.hlsl:1-20", seeds=seeds, streams=streams), indent=1))
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
