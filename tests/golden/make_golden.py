#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/.

The reference holds no vectors for this path (it has no path tracer), so the
fixtures pin this build against itself across changes:
  pcg32.json            PCG32 outputs from the Python-int restatement
                        (tests/test_oracle_pcg.py::py_pcg32), not from the C code.
  cornell_hits.npz      4096 rays on the Cornell + 10k-soup scene and their
                        closest-hit records from the oracle's brute force
                        (every triangle tested; no tree involved).
  generator_hashes.json SHA-256 of generated vertex and index buffers.
Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from foundation_amd import _abi as A  # noqa: E402
from foundation_amd import pt  # noqa: E402
from foundation_amd import rays as R  # noqa: E402
from oracle import pt_oracle as O  # noqa: E402
from tests.test_oracle_pcg import py_pcg32  # noqa: E402

GEN_CASES = {
    "cornell_soup": (A.GEN_CORNELL_SOUP, {}),
    "sphere_n32_soup5k": (A.GEN_DISPLACED_SPHERE, {"subdivision": 32, "soup_triangles": 5000}),
    "instanced_8x_n10": (A.GEN_INSTANCED_SPHERES, {"instances": 8, "subdivision": 10}),
    "furnace_n8": (A.GEN_FURNACE, {"subdivision": 8}),
}


def scene_hash(gen):
    h = hashlib.sha256()
    for m in gen.meshes():
        for k in ("positions", "indices", "material", "light"):
            h.update(np.ascontiguousarray(m[k]).tobytes())
    for inst in gen.instances():
        h.update(inst["o2w"].tobytes())
        h.update(inst["w2o"].tobytes())
    return h.hexdigest()


def golden_rays(gen, oscene):
    prim = R.primary_rays(gen.camera, 64, 32)
    ph = oscene.trace(prim, brute=True)
    bounce, _ = R.bounce_rays(prim, ph, gen.triangle_lookup(), seed=1)
    rays = np.concatenate([prim, bounce], axis=0)[:4096]
    return rays


def main():
    cases = []
    for seed, seq in [(42, 54), (0, 0), (1, 1), (2**64 - 1, 2**63), (0xDEADBEEF, 7)]:
        cases.append({"initstate": seed, "initseq": seq, "out": py_pcg32(seed, seq, 16)})
    json.dump({"algorithm": "pcg32 XSH-RR, pcg32_srandom_r(initstate, initseq)", "cases": cases},
              open(os.path.join(HERE, "pcg32.json"), "w"), indent=1)

    gen = pt.GeneratedScene(A.GEN_CORNELL_SOUP)
    osc = O.scene_from_generated(gen)
    rays = golden_rays(gen, osc)
    hits = osc.trace(rays, brute=True)
    np.savez_compressed(os.path.join(HERE, "cornell_hits.npz"), rays=rays, hits=hits.view(np.uint32))

    hashes = {name: scene_hash(pt.GeneratedScene(kind, **kw)) for name, (kind, kw) in GEN_CASES.items()}
    json.dump(hashes, open(os.path.join(HERE, "generator_hashes.json"), "w"), indent=1)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
