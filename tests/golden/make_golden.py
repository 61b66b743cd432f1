"""Writes tests/golden/*.npz: inputs and expected sorted result ids.

The expected sets come from the FP64 oracle (oracle/), which tests/test_oracle.py
pins against closed-form invariants, a KD-tree k-NN, planted mutants and the
reference's own rknn-naive binary.  The reference ships no fixtures of its own
(it has no tests), so these files are this project's regression vectors.

  config1.npz  10,000 users / 1,000 facilities, uniform fp32 (the generator's
               config 1), 10 queries, k = 5  -- exactly the brute-force workload
  ties.npz     20,000 users / 500 facilities, clustered and snapped to a 2^-7
               grid with users placed exactly on query bisectors or duplicated
               (the generator's config 3 at reduced size), 12 queries x
               k in {1, 5, 10, 25}
Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "rt-rknn_amd")]

import oracle as O  # noqa: E402
import rknn_amd as R  # noqa: E402


def flatten(results):
    offs = np.zeros(len(results) + 1, np.int64)
    offs[1:] = np.cumsum([len(r) for r in results])
    ids = np.concatenate(results) if results else np.zeros(0, np.int32)
    return ids.astype(np.int32), offs


def make(path, config, nu, nf, nq, ks):
    seed = R.default_seed(config)
    ux, uy, uid, fx, fy, fid = R.generate(config, nu, nf, seed=seed, nq_ties=nq, seed_ties=seed)
    qs = R.generate_queries(nf, nq, seed)
    O.set_threads(8)
    qk, res = [], []
    for q in qs:
        for k in ks:
            qk.append((int(q), int(k)))
            res.append(O.rknn(ux, uy, uid, fx, fy, int(q), int(k)))
    ids, offs = flatten(res)
    np.savez_compressed(path, ux=ux, uy=uy, uid=uid, fx=fx, fy=fy, fid=fid, qk=np.array(qk, np.int64), ids=ids,
                        offs=offs)
    print(path, "queries", len(qk), "result ids", len(ids))


if __name__ == "__main__":
    make(os.path.join(HERE, "config1.npz"), 1, 10_000, 1_000, 10, [5])
    make(os.path.join(HERE, "ties.npz"), 3, 20_000, 500, 12, [1, 5, 10, 25])
