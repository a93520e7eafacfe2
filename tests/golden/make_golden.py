"""Writes tests/golden/oracle_golden.npz: oracle (FP64, reference schedule)
outputs for reduced versions of the benchmark configurations, on seeded
synthetic inputs that tests regenerate from the same seeds.

Run from the repository root:  python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'tests'))

import golden_cases as G  # noqa: E402


def main():
    import oracle
    be = oracle.backend(schedule=0)
    out = {}
    for name in G.CASES:
        res = G.run_case(be, name)
        for k, v in res.items():
            if k.startswith('_'):
                continue
            out[f'{name}/{k}'] = np.asarray(v)
        print(name, 'done', flush=True)
    np.savez_compressed(os.path.join(HERE, 'oracle_golden.npz'), **out)


if __name__ == '__main__':
    main()
