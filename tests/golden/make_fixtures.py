"""Writes the hand-made .gr byte fixtures in this directory.

The bytes are laid out here with struct.pack, field by field, following the
Galois .gr v1 layout read by the reference's graph_loader::ReadGraphGR
(include/graph_loader.cuh:71-154): u64 header {1, sizeEdgeTy, N, M}, u64[N]
end offsets, u32[M] destinations, a 4-byte pad when M is odd and the file is
weighted, u32[M] weights.  Neither the product's writer nor the oracle is
used, so the fixtures check both readers independently.

Run: python tests/golden/make_fixtures.py   (rewrites the .gr files)
"""
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))

# (name, V, rows as lists of (dst, weight), weighted)
FIXTURES = [
    # M = 5 (odd) weighted: exercises the pad (graph_loader.cuh:139-141)
    ("odd_weighted", 4, [[(1, 7), (2, 3)], [(2, 1)], [(3, 9)], [(0, 4)]], True),
    # M = 4 (even) weighted: no pad
    ("even_weighted", 3, [[(1, 2), (2, 5)], [(2, 8)], [(0, 6)]], True),
    # M = 3 unweighted: no weight section at all
    ("odd_unweighted", 4, [[(1, 0)], [(2, 0)], [], [(0, 0)]], False),
    # isolated vertices and an empty last row
    ("isolated", 5, [[], [(3, 11)], [], [(1, 11)], []], True),
]


def write(name, V, rows, weighted):
    M = sum(len(r) for r in rows)
    b = struct.pack("<4Q", 1, 4 if weighted else 0, V, M)
    end = 0
    for r in rows:
        end += len(r)
        b += struct.pack("<Q", end)
    for r in rows:
        for d, _ in r:
            b += struct.pack("<I", d)
    if weighted:
        if M % 2:
            b += struct.pack("<I", 0)
        for r in rows:
            for _, w in r:
                b += struct.pack("<I", w)
    with open(os.path.join(HERE, name + ".gr"), "wb") as f:
        f.write(b)


if __name__ == "__main__":
    for fx in FIXTURES:
        write(*fx)
