"""Committed fixtures (tests/golden/, written by tests/golden/make_golden.py): the oracle must reproduce them
bit for bit on the CPU, and the HIP solver must reproduce the non-solver part bit for bit and the MGPCG pressure
within the 1e-5 relative-L2 bar on the GPU."""
import json
import os

import numpy as np
import pytest

import oracle_py as O
from slosh_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def _make_golden():
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLD, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_pcg_hash_vectors():
    d = json.load(open(os.path.join(GOLD, "pcg_hash.json")))
    for x, y in zip(d["input"], d["output"]):
        assert O.pcg_hash(x) == y


def test_oracle_reproduces_scene12():
    mg = _make_golden()
    gold = np.load(os.path.join(GOLD, "scene12.npz"))
    out, sim = mg.run_scene(O.OracleSim)
    for k, v in out.items():
        assert np.array_equal(np.asarray(v).view(np.uint8), np.asarray(gold[k]).view(np.uint8)), k
    ref, _, _ = sim.solve_fp64(1e-13, 400)
    assert np.allclose(ref, gold["pressure_fp64_reference"], rtol=0, atol=1e-12)


@pytest.mark.gpu
def test_gpu_reproduces_scene12():
    from slosh_amd import Sim

    mg = _make_golden()
    gold = np.load(os.path.join(GOLD, "scene12.npz"))
    out, _ = mg.run_scene(Sim)
    for k in ("particles", "brick_table", "cell_flags", "vel", "face_flags"):
        assert np.array_equal(np.asarray(out[k]).view(np.uint8), np.asarray(gold[k]).view(np.uint8)), k
    ref = gold["pressure_fp64_reference"]
    p = out["pressure_mgpcg"]
    assert np.linalg.norm(p - ref) / np.linalg.norm(ref) <= 1e-5
