"""Writes the golden vectors under tests/golden/ from independent pure-Python
restatements (no oracle, no reference code is executed).

pcg4d.json: inputs and outputs of the PCG4D hash of common/wltcRS.glsl:22-27
(Jarzynski and Olano, "Hash Functions for GPU Rendering", JCGT 9(3) 2020),
computed with explicit modulo-2^32 arithmetic, plus the seeds of
rng_initialize (wltcRS.glsl:9-19) for a few pixels and frames.

Run: python tests/golden/make_golden.py
"""
import json
import os

M = 0xFFFFFFFF


def pcg4d(v):
    x, y, z, w = [(c * 1664525 + 1013904223) & M for c in v]
    x = (x + y * w) & M
    y = (y + z * x) & M
    z = (z + x * y) & M
    w = (w + y * z) & M
    x ^= x >> 16
    y ^= y >> 16
    z ^= z >> 16
    w ^= w >> 16
    x = (x + y * w) & M
    y = (y + z * x) & M
    z = (z + x * y) & M
    w = (w + y * z) & M
    return [x, y, z, w]


def seeds(x, y, frame):
    s0 = [x, y, frame, (x + y) & M]
    s1 = [frame, (frame * 15843) & M, (frame * 31 + 4566) & M, (frame * 2345 + 58585) & M]
    return s0, s1


def main():
    inputs = [[0, 0, 0, 0], [1, 2, 3, 4], [M, M, M, M], [0x12345678, 0x9ABCDEF0, 0x0BADF00D, 0xDEADBEEF],
              [511, 511, 63, 1022], [1919, 1079, 255, 2998]]
    chain = [[7, 11, 13, 17]]
    for _ in range(8):
        chain.append(pcg4d(chain[-1]))
    data = {
        "pcg4d": [{"in": v, "out": pcg4d(v)} for v in inputs],
        "pcg4d_chain": chain,
        "rng_initialize": [{"pixel": [x, y], "frame": f, "s0": seeds(x, y, f)[0], "s1": seeds(x, y, f)[1]}
                           for x, y, f in [(0, 0, 0), (3, 5, 1), (1919, 1079, 255), (7, 9, 300000)]],
    }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pcg4d.json")
    with open(out, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main()
