"""Candidate sets of the autotuner (no device needed)."""
from hcgpu import autotune


def test_layout_is_a_candidate_up_to_24_variables():
    small = autotune.candidates_for(10)
    assert {c["layout"] for c in small} == {"lane", "group"}
    assert all(c["waves_per_eu"] == 0 for c in small if c["layout"] == "group")
    assert len([c for c in small if c["layout"] == "lane"]) == len(autotune.DEFAULT_CANDIDATES)
    assert {c["layout"] for c in autotune.candidates_for(24)} == {"lane", "group"}
    assert {c["layout"] for c in autotune.candidates_for(25)} == {"lane"}


def test_pairs_are_per_lane_candidates():
    assert autotune._as_candidate((2, 8)) == dict(layout="lane", waves_per_eu=2, blocks_per_cu=8)
    c = dict(layout="group", blocks_per_cu=2)
    assert autotune._as_candidate(c) == dict(layout="group", waves_per_eu=0, blocks_per_cu=2)
    assert autotune._normalized({"waves_per_eu": 2, "blocks_per_cu": 8})["layout"] == "lane"
