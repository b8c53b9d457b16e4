"""The synthetic inputs (heightfield, ray sets, scenes) reproduce the committed sha256 values."""
import json
import os

from tests.golden import make_input_hashes

HERE = os.path.dirname(os.path.abspath(__file__))


def test_generated_inputs_match_committed_hashes():
    want = json.load(open(os.path.join(HERE, "golden", "inputs_sha256.json")))
    got = make_input_hashes.compute()
    assert got == want
