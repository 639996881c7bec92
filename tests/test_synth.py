"""Seeded generators of the synthetic workloads: known answers of the
SplitMix64 stream, committed digests of generated arrays and files
(tests/golden/synth_checksums.json, written by make_synth_checksums.py), and
the shape/range contracts the benchmark relies on."""
import json
import os

import numpy as np

from rift_amd import synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_splitmix64_known_answers():
    # SplitMix64 (Steele, Lea, Flood 2014) from state 0: the published
    # first outputs
    z = synth.splitmix64(0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_uniform_and_normal_moments():
    u = synth.uniform(11, 200000)
    assert 0.0 < u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 3e-3 and abs(u.var() - 1.0 / 12.0) < 2e-3
    g = synth.normal(11, 200000)
    assert abs(g.mean()) < 1e-2 and abs(g.std() - 1.0) < 1e-2


def test_generated_inputs_match_committed_digests():
    import sys
    sys.path.insert(0, GOLDEN)
    import make_synth_checksums as M
    want = json.load(open(os.path.join(GOLDEN, "synth_checksums.json")))
    got = M.compute()
    assert got == want


def test_workload_contracts():
    rb = synth.rain_basin(64)
    assert rb["dem"].shape == (65, 65) and rb["K"].shape == (65, 65)
    assert 0.0 <= rb["dem"].max() <= 60.0 and rb["dem"].min() >= -3.0
    cs = synth.coastal_surge(64, nmaps=3)
    assert cs["surge_maps"].shape == (3, 65, 65)
    assert set(np.unique(cs["n"])) == {0.02, 0.06}
    db = synth.dam_break(64)
    assert db["h0"][:, :32].min() == 1.0 and db["h0"][:, 32:].max() == 0.1
