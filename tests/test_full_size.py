"""The oracle at the benchmark configurations' full sizes against the
reference's digests (tests/golden/full_sizes.json): config 1 (800x600, 10k
triangles over 366M fragments) and config 2 (1920x1080, 200k textured,
partly blended triangles)."""
import json
import os

import pytest

import scenes
from conftest import GOLDEN, render_oracle

with open(os.path.join(GOLDEN, "full_sizes.json")) as f:
    FULL = json.load(f)


@pytest.mark.parametrize("name,kw", scenes.FULL_SIZE_SCENES, ids=[n for n, _ in scenes.FULL_SIZE_SCENES])
def test_oracle_full_size_matches_reference(name, kw):
    g = FULL[name]
    assert g["params"] == json.loads(json.dumps(kw))
    c, d, e, frags = render_oracle(scenes.params(kw))
    assert e == 0
    assert scenes.digest(c, d) == (g["color_sha256"], g["depth_sha256"])
