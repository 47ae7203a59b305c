"""Writes tests/golden/config1_small.npz: a small seeded config-1 instance
with expected bitmaps and group sums computed by the numpy definitions in
tests/npref.py (not by the oracle), so the fixture pins the oracle and the
GPU path alike.  Run from the repo root: python tests/golden/make_golden.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from merle_amd import configs  # noqa: E402
from merle_amd.spec import Schema  # noqa: E402
import npref  # noqa: E402

N = 5003   # not a multiple of 32 or 64: exercises the tail word


def main():
    cols = configs.config1_host(N, seed=12345)
    bounds = np.array(configs.C1_BOUNDS, dtype=np.uint64)
    W = (N + 31) // 32
    stride = ((W + 63) // 64) * 64
    p0 = cols["p0"]
    bins = np.stack([npref.pack((p0 >= bounds[j]) & (p0 < bounds[j + 1]), stride)
                     for j in range(len(bounds) - 1)])
    s = Schema(N)
    for k, v in cols.items():
        s.add(k, v)
    sums, nsel = npref.group_sum(configs.config1_spec(), s)
    np.savez_compressed(os.path.join(HERE, "config1_small.npz"),
                        bounds=bounds, bins_p0=bins, sums=sums,
                        nsel=np.int64(nsel), **cols)
    print("wrote config1_small.npz", N, "rows,", nsel, "selected")


if __name__ == "__main__":
    main()
