"""Writes tests/golden/synth_checksums.json: SHA-256 digests of the seeded
generator outputs (rift_amd/synth.py) and of the files write_case produces,
so that a seed names the same input on every machine.  Run from the
repository root after an intended generator change:

    python tests/golden/make_synth_checksums.py
"""
import hashlib
import json
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

from rift_amd import synth  # noqa: E402


def digest_arrays(d):
    out = {}
    for k in sorted(d):
        v = d[k]
        if isinstance(v, np.ndarray):
            out[k] = hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest()
        else:
            out[k] = repr(v)
    return out


def digest_files(case, **kw):
    with tempfile.TemporaryDirectory() as tmp:
        synth.write_case(case, tmp, **kw)
        return {f: hashlib.sha256(open(os.path.join(tmp, f), "rb").read()).hexdigest()
                for f in sorted(os.listdir(tmp))}


def compute():
    surge = synth.coastal_surge(64, nmaps=4)
    return {
        "dam_break_256": digest_arrays(synth.dam_break(256)),
        "rain_basin_256": digest_arrays(synth.rain_basin(256)),
        "rain_basin_2048_dem": hashlib.sha256(synth.rain_basin(2048)["dem"].tobytes()).hexdigest(),
        "coastal_surge_128": digest_arrays(synth.coastal_surge(128, nmaps=5)),
        "splitmix64_seed7": hashlib.sha256(synth.splitmix64(7, 1000).tobytes()).hexdigest(),
        "files_rain_basin_32": digest_files(synth.rain_basin(32), save_max=1),
        "files_coastal_surge_64": digest_files(surge, h_print=1, dt_print=60),
    }


if __name__ == "__main__":
    with open(os.path.join(HERE, "synth_checksums.json"), "w") as f:
        json.dump(compute(), f, indent=1, sort_keys=True)
        f.write("\n")
