"""Writes tests/golden/deviates.json: deviate vectors of RunContractionLP's
draw (lib/components/exponential_contraction.h:467-471) computed through the
standard library objects the reference uses (std::minstd_rand0 as
default_random_engine, std::exponential_distribution<double>), from the
oracle's std_deviate.cpp built with and without FMA contraction.

Run from the repository root after `make -C oracle`:  python tests/golden/make_golden.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

CASES = [
    # (gid0, count, seed, iteration, rng_offset, beta)
    (0, 256, 1001, 1, 1 << 20, 0.1),
    (1 << 20, 256, 1001, 2, 1 << 20, 0.1),
    (4294967040, 256, 7, 5, 123456789, 0.1),
    (12345, 256, 1003, 1, 16777216, 0.05),
    (999, 256, 1, 3, 1000, 0.5),
    (0, 256, 2000000000, 9, (1 << 33) + 5, 1.0),
    (77, 256, 1002, 40, 3, 0.01),
]


def main():
    L = oracle.lib()
    out = []
    for gid0, count, seed, it, off, beta in CASES:
        fma = [int(L.or_std_deviate_fma(gid0 + i, seed, it, off, beta)) for i in range(count)]
        nofma = [int(L.or_std_deviate_nofma(gid0 + i, seed, it, off, beta)) for i in range(count)]
        out.append({"gid0": gid0, "count": count, "seed": seed, "iteration": it, "rng_offset": off, "beta": beta,
                    "fma": fma, "nofma": nofma})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "deviates.json")
    with open(path, "w") as f:
        json.dump({"source": "std::exponential_distribution over std::default_random_engine (libstdc++)",
                   "cases": out}, f, separators=(",", ":"))
    print("wrote", path)


if __name__ == "__main__":
    main()
