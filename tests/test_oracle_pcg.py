"""PCG32 (O'Neill 2014) in the oracle, pinned three ways: the published demo
vector of the PCG C library (pcg32-demo, seed 42, sequence 54), an
independent restatement with Python ints, and the vectorised numpy version
the ray generators use."""
import json
import os

import numpy as np

from foundation_amd import rays as R

HERE = os.path.dirname(os.path.abspath(__file__))

# pcg32_srandom_r(&rng, 42u, 54u); first six pcg32_random_r outputs.
DEMO = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]
M64 = (1 << 64) - 1


def py_pcg32(initstate, initseq, n):
    state = 0
    inc = ((initseq << 1) | 1) & M64

    def step():
        nonlocal state
        old = state
        state = (old * 6364136223846793005 + inc) & M64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    step()
    state = (state + initstate) & M64
    step()
    return [step() for _ in range(n)]


def test_demo_vector(oracle):
    assert list(oracle.pcg32_seq(42, 54, 6)) == DEMO
    assert py_pcg32(42, 54, 6) == DEMO


def test_oracle_equals_python_int_restatement(oracle):
    for seed, seq in [(0, 0), (1, 1), (2**63 + 5, 2**64 - 1), (123456789, 987654321)]:
        assert list(oracle.pcg32_seq(seed, seq, 1000)) == py_pcg32(seed, seq, 1000)


def test_numpy_vectorised_equals_oracle(oracle):
    seqs = np.arange(257, dtype=np.uint64) * np.uint64(7919)
    state, inc = R.pcg32_srandom(np.full(seqs.shape, 1, np.uint64), seqs)
    outs = []
    for _ in range(4):
        o, state = R.pcg32_next(state, inc)
        outs.append(o)
    outs = np.stack(outs, axis=1)
    for i in (0, 1, 100, 256):
        assert list(outs[i]) == list(oracle.pcg32_seq(1, int(seqs[i]), 4))


def test_unit_float_is_exact():
    u = np.array([0, 0xFF, 0x100, 0xFFFFFFFF], np.uint32)
    f = R.u32_to_unit(u)
    assert f[0] == 0.0 and f[1] == 0.0 and f[2] == 2.0 ** -24
    assert f[3] == 1.0 - 2.0 ** -24 and f[3] < 1.0


def test_golden_pcg_vectors(oracle):
    g = json.load(open(os.path.join(HERE, "golden", "pcg32.json")))
    for case in g["cases"]:
        assert list(oracle.pcg32_seq(case["initstate"], case["initseq"], len(case["out"]))) == case["out"]
