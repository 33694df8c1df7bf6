"""Writes the golden vectors under tests/golden/.

rng_vectors.npz: Lehmer RNG of the reference (rng.cl:37-45 with the seeding
of integratefields.cl:93-99) evaluated with exact Python integers, which is
independent of the C oracle and the HIP kernels:
  idx0, rng0[i] = seed state of seed index idx0+i, draws[i, k] = k-th draw;
  seed_idx / seed_rng0: seeding of chosen edge indices.
seeds_small_256.npz: the first 1000 shuffled seeds of the 256x256 workload
(create_seeds, useful.py:470-517), from numpy's legacy MT19937 permutation.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
M = 0xfffffffb


def lehmer(s):
    return ((s + 1) & 0xffffffff) * 279470273 % M


def brev(v):
    return int.from_bytes(int(v).to_bytes(4, "little"), "big")


def rng_seed(i):
    return brev(lehmer(brev(i)))


def main():
    idx0, n, nd = 4000000000, 256, 24
    rng0 = np.zeros(n, dtype=np.uint32)
    draws = np.zeros((n, nd), dtype=np.uint32)
    for i in range(n):
        s = rng_seed(idx0 + i)
        rng0[i] = s
        for k in range(nd):
            s = lehmer(s)
            draws[i, k] = s
    seed_idx = np.array([0, 1, 255, 256, 65535, 65536, 0x7fffffff, 0x80000000, 0xfffffffa, 0xffffffff],
                        dtype=np.uint64)
    seed_rng0 = np.array([rng_seed(int(i)) for i in seed_idx], dtype=np.uint32)
    np.savez(os.path.join(HERE, "rng_vectors.npz"), idx0=np.uint64(idx0), rng0=rng0, draws=draws,
             seed_idx=seed_idx, seed_rng0=seed_rng0)

    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from slm_amd import workloads
    job = workloads.small_256()
    full = np.argwhere(~job.mask).astype(np.float32) - 2
    perm = np.random.RandomState(1).permutation(full.shape[0])
    np.savez(os.path.join(HERE, "seeds_small_256.npz"), perm_head=perm[:1000].astype(np.int64),
             seeds_head=full[perm[:1000]], n_seeds=np.int64(full.shape[0]))


if __name__ == "__main__":
    main()
