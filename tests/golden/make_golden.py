#!/usr/bin/env python3
"""Regenerates the committed golden fixtures in tests/golden/ (run in the build container).

* philox_torch.json -- Philox-4x32-10 outputs from torch's open-source
  ATen/core/PhiloxRNGEngine.h (a third-party implementation, compiled here with g++),
  for random (counter, key) pairs.  philox_engine(seed, subsequence, offset) uses
  counter = (offset_lo, offset_hi, subseq_lo, subseq_hi) and key = (seed_lo, seed_hi).
* randnum_ref.json -- RandNum::lRand and RandNum::inlineRandPair outputs of the
  reference's own util/util.hpp, produced by oracle/_ref/randnum_ref (built by
  `make -C oracle ref` from the reference sources; nothing of them is stored here).
"""
import json
import os
import random
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))

TORCH_SRC = r"""
#include <ATen/core/PhiloxRNGEngine.h>
#include <cstdio>
#include <cstdlib>
int main(int argc, char** argv) {
  for (int i = 1; i + 5 < argc + 1; i += 6) {
    unsigned c[4], k[2];
    for (int j = 0; j < 4; j++) c[j] = (unsigned)strtoul(argv[i + j], 0, 10);
    for (int j = 0; j < 2; j++) k[j] = (unsigned)strtoul(argv[i + 4 + j], 0, 10);
    at::philox_engine e((uint64_t)k[0] | ((uint64_t)k[1] << 32),
                        (uint64_t)c[2] | ((uint64_t)c[3] << 32),
                        (uint64_t)c[0] | ((uint64_t)c[1] << 32));
    unsigned o0 = e(), o1 = e(), o2 = e(), o3 = e();
    printf("%u %u %u %u\n", o0, o1, o2, o3);
  }
  return 0;
}
"""


def torch_philox():
    import torch
    inc = os.path.join(os.path.dirname(torch.__file__), "include")
    rng = random.Random(20240601)
    cases = [([0, 0, 0, 0], [0, 0]),
             ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2),
             ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])]
    for _ in range(61):
        cases.append(([rng.getrandbits(32) for _ in range(4)], [rng.getrandbits(32) for _ in range(2)]))
    # counters the walk kernels use: (walker, hop, k, tag) with seed 42
    for w, h, k, t in [(0, 0, 0, 0), (1, 0, 0, 1), (123456, 79, 3, 0), (2**31 + 5, 29, 0, 1)]:
        cases.append(([w, h, k, t], [42, 0]))
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t.cpp")
        exe = os.path.join(d, "t")
        open(src, "w").write(TORCH_SRC)
        subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", inc, "-o", exe, src])
        args = []
        for c, k in cases:
            args += [str(x) for x in c + k]
        out = subprocess.check_output([exe] + args, text=True).split("\n")
    res = []
    for (c, k), line in zip(cases, out):
        res.append({"ctr": c, "key": k, "out": [int(x) for x in line.split()]})
    return res


def randnum_ref():
    exe = os.path.join(REPO, "oracle", "_ref", "randnum_ref")
    if not os.path.exists(exe):
        subprocess.check_call(["make", "-C", os.path.join(REPO, "oracle"), "ref"])
    res = []
    for seed in (42, 9898676785859, 0, 2**64 - 1):
        for max_val, bound in ((2.0, 1), (2.0, 2), (0.5, 7), (4.0, 55), (2.0, 3000000),
                               (1.2e-6, 3000000)):
            out = subprocess.check_output([exe, str(seed), "64", repr(max_val), str(bound), "64"], text=True)
            res.append(json.loads(out))
    return res


def main():
    json.dump(torch_philox(), open(os.path.join(HERE, "philox_torch.json"), "w"))
    json.dump(randnum_ref(), open(os.path.join(HERE, "randnum_ref.json"), "w"))
    print("wrote philox_torch.json, randnum_ref.json")


if __name__ == "__main__":
    sys.exit(main())
