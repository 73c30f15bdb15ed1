"""Write tests/golden/rand_kat.json: known answers of Rand(seed) (base/header/Random.h:18-27).

Computed with numpy uint32 arithmetic (wrap-around multiply, shifts, xor), independent of the
C++ oracle and of the HIP library; the final division is float32(i) / float32(UINT_MAX).
"""
import json
import os

import numpy as np


def rand_np(seeds):
    with np.errstate(over="ignore"):
        i = (np.asarray(seeds, np.uint32) ^ np.uint32(12345391)) * np.uint32(2654435769)
        i = i ^ ((i << np.uint32(6)) ^ (i >> np.uint32(26)))
        i = i * np.uint32(2654435769)
        i = i + ((i << np.uint32(5)) ^ (i >> np.uint32(12)))
    return i.astype(np.float32) / np.float32(4294967295.0)


if __name__ == "__main__":
    seeds = [0, 1, 2, 3, 7, 42, 1000, 12345391, 12345390, 65535, 1 << 20, 2 ** 31 - 1, 2 ** 31, 2 ** 32 - 1,
             3 * 134217727 + 2, 987654321]
    vals = rand_np(seeds)
    out = {"seeds": seeds, "values": [float(v) for v in vals]}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rand_kat.json")
    json.dump(out, open(path, "w"), indent=1)
    print(path)
