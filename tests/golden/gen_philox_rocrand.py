"""Writes philox_rocrand.json: Philox4x32-10 blocks computed by rocRAND's
host-callable engine (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h), an
independent implementation used as a known-answer check of the oracle's and
the kernels' Philox.

rocRAND's engine(seed, subsequence, offset) keys with seed and starts at
counter (offset/4 as 64 bits in words 0-1, subsequence in words 2-3), so its
first four outputs are the raw block for that counter.

Run:  python tests/golden/gen_philox_rocrand.py   (needs hipcc; host code only)
"""
import json
import os
import random
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = r'''
#include <rocrand/rocrand_kernel.h>
#include <cstdio>
#include <cstdlib>
int main(int argc, char** argv) {
  unsigned c0 = strtoul(argv[1], 0, 0), c1 = strtoul(argv[2], 0, 0), c2 = strtoul(argv[3], 0, 0),
           c3 = strtoul(argv[4], 0, 0), k0 = strtoul(argv[5], 0, 0), k1 = strtoul(argv[6], 0, 0);
  unsigned long long seed = ((unsigned long long)k1 << 32) | k0;
  unsigned long long sub = ((unsigned long long)c3 << 32) | c2;
  unsigned long long off = 4ull * (((unsigned long long)c1 << 32) | c0);
  rocrand_device::philox4x32_10_engine e(seed, sub, off);
  unsigned a = e(), b = e(), c = e(), d = e();
  printf("%u %u %u %u\n", a, b, c, d);
  return 0;
}
'''


def main():
    rng = random.Random(20261203)
    # rocRAND takes the counter's low 64 bits as a number-offset of 4*x, so
    # word 1 must stay below 2^30 here.
    cases = [((0, 0, 0, 0), (0, 0)),
             ((0xFFFFFFFF, 0x3FFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF)),
             ((0x243F6A88, 0x05A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0)),
             ((7, 3, 0, 0x4D43), (2, 0)),
             ((16777215, 999, 5, 0x4943), (1, 0))]
    for _ in range(20):
        c = [rng.getrandbits(32) for _ in range(4)]
        c[1] &= 0x3FFFFFFF
        cases.append((tuple(c), tuple(rng.getrandbits(32) for _ in range(2))))
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "p.cpp")
        exe = os.path.join(td, "p")
        open(src, "w").write(SRC)
        subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-o", exe, src],
                       check=True)
        out = []
        for ctr, key in cases:
            r = subprocess.run([exe, *map(str, ctr), *map(str, key)], check=True, capture_output=True, text=True)
            out.append({"ctr": list(ctr), "key": list(key), "out": [int(x) for x in r.stdout.split()]})
    with open(os.path.join(HERE, "philox_rocrand.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {len(out)} vectors")


if __name__ == "__main__":
    main()
