"""Writes the small committed fixtures of tests/golden/ from first principles.

- tiny.gr: a 3-node weighted graph in the .gr v1 layout of
  LONESTAR/rt/src/csr_graph.cu:187-199, packed by hand with struct (edges
  0->1 w5, 0->2 w7, 1->2 w1; nedges odd, so 4 bytes of padding precede the
  weights).
- splitmix64.json: the SplitMix64 sequence for state 1234567 as published with
  Vigna's splitmix64.c (known-answer vector), and the first outputs of the
  counter-based uniform generator for two seeds.

Run: python tests/golden/make_fixtures.py
"""
import json
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))
M = (1 << 64) - 1
G = 0x9E3779B97F4A7C15


def mix(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def main():
    # .gr v1: version, edge type size, nnodes, nedges, row ends, dst, pad, weights
    b = struct.pack("<4Q", 1, 4, 3, 3)
    b += struct.pack("<3Q", 2, 3, 3)
    b += struct.pack("<3I", 1, 2, 2)
    b += struct.pack("<I", 0)
    b += struct.pack("<3I", 5, 7, 1)
    with open(os.path.join(HERE, "tiny.gr"), "wb") as f:
        f.write(b)

    seq, x = [], 1234567
    for _ in range(5):
        x = (x + G) & M
        seq.append(mix(x))
    gen = {}
    for seed, lo, hi in ((0, 0, 255), (20240601, 0, 15)):
        st = mix(seed)
        vals = []
        for i in range(16):
            r = mix((st + (i + 1) * G) & M) >> 32
            vals.append(lo + ((r * (hi - lo + 1)) >> 32))
        gen[f"{seed}:{lo}:{hi}"] = vals
    with open(os.path.join(HERE, "splitmix64.json"), "w") as f:
        json.dump({"splitmix64_state_1234567": seq, "uniform_u32": gen}, f, indent=1)


if __name__ == "__main__":
    main()
