"""The M2L split autotuner's cache and key logic (CPU) and, on the GPU, that an
autotuned pass equals the default-rule pass (the split count only changes the
order of the split additions) and stores its choice on disk."""

import json

import numpy as np
import pytest

from sumpy_amd import autotune


def _entries(*shapes):
    return [{"sched": {"ntiles": t, "nslots": s, "npairs": p}} for t, s, p in shapes]


def test_cache_roundtrip(tmp_path):
    c = autotune.TuneCache(str(tmp_path))
    assert c.get("k") is None
    c.put("k", {"policy": [2, 6]})
    assert autotune.TuneCache(str(tmp_path)).get("k") == {"policy": [2, 6]}
    assert json.load(open(tmp_path/"autotune.json"))["k"]["policy"] == [2, 6]


def test_digest_groups_similar_schedules():
    a = autotune.schedule_digest(_entries((1200, 56000, 3_500_000), (60, 3000, 90_000)))
    b = autotune.schedule_digest(_entries((60, 3001, 90_100), (1204, 56100, 3_510_000)))
    c = autotune.schedule_digest(_entries((2400, 56000, 3_500_000), (60, 3000, 90_000)))
    assert a == b and a != c


def test_policy_read_from_disk_without_timing(tmp_path):
    class Actx:
        device = "cpu"
    ents = _entries((10, 100, 1000))
    key = autotune.make_key("LapKnl3D", 10, "fp32", "cpu", autotune.schedule_digest(ents))
    cache = autotune.TuneCache(str(tmp_path))
    cache.put(key, {"policy": [8, 24]})
    tuner = autotune.M2LSplitTuner(cache=autotune.TuneCache(str(tmp_path)))
    assert tuner.policy(Actx(), ents, "LapKnl3D", 10, "fp32") == (8, 24)


def test_shape_choice_read_from_disk_without_timing(tmp_path):
    """A stored (tile width, split policy) choice is returned without building
    the 32-column entries or timing anything."""
    class Actx:
        device = "cpu"
    ents = _entries((10, 100, 1000))
    key = autotune.make_shape_key("LapKnl3D", 10, "bf16x3", "cpu",
                                  autotune.schedule_digest(ents))
    cache = autotune.TuneCache(str(tmp_path))
    cache.put(key, {"tile_width": 32, "policy": [2, 24]})
    tuner = autotune.M2LSplitTuner(cache=autotune.TuneCache(str(tmp_path)))
    asked = []

    def make_entries(tw):
        asked.append(tw)
        return ents
    assert tuner.choose(Actx(), make_entries, "LapKnl3D", 10, "bf16x3") == (32, (2, 24))
    assert asked == [64]


def test_key_carries_library_digest(tmp_path):
    """A choice stored under another build of the library is not read back."""
    class Actx:
        device = "cpu"
    ents = _entries((10, 100, 1000))
    dig = autotune.schedule_digest(ents)
    assert autotune.make_key("K", 10, "fp32", "cpu", dig) == \
        autotune.make_key("K", 10, "fp32", "cpu", dig, autotune.code_version())
    cache = autotune.TuneCache(str(tmp_path))
    cache.put(autotune.make_key("K", 10, "fp32", "cpu", dig, "0123456789ab"),
              {"policy": [8, 24]})
    tuner = autotune.M2LSplitTuner(cache=autotune.TuneCache(str(tmp_path)))
    tuner._memo = {}
    key = autotune.make_key("K", 10, "fp32", "cpu", dig)
    assert tuner.cache.get(key) is None


@pytest.mark.gpu
def test_autotuned_pass_matches_default(actx, tmp_path, monkeypatch):
    from functools import partial

    from sumpy_amd.drive import drive_fmm
    from sumpy_amd.expansion.local import LinearPDEConformingVolumeTaylorLocalExpansion
    from sumpy_amd.expansion.m2l import VolumeTaylorM2LWithDenseClassMatrices
    from sumpy_amd.expansion.multipole import LinearPDEConformingVolumeTaylorMultipoleExpansion
    from sumpy_amd.fmm import SumpyExpansionWrangler, SumpyTreeIndependentDataForWrangler
    from sumpy_amd.kernel import LaplaceKernel
    from sumpy_amd.tree import build_tree_and_traversal
    monkeypatch.setenv("SUMPY_AMD_CACHE_DIR", str(tmp_path))
    rng = np.random.default_rng(3)
    src, q = rng.random((3, 20000)), rng.standard_normal(20000)
    tree, trav = build_tree_and_traversal(actx, src, None, 32)
    knl = LaplaceKernel(3)
    pots = []
    for tune in (False, True):
        tr = VolumeTaylorM2LWithDenseClassMatrices("fp32")
        ti = SumpyTreeIndependentDataForWrangler(
            actx, partial(LinearPDEConformingVolumeTaylorMultipoleExpansion, knl),
            partial(LinearPDEConformingVolumeTaylorLocalExpansion, knl,
                    m2l_translation_override=tr), [knl], exclude_self=True,
            autotune=tune)
        wr = SumpyExpansionWrangler(
            ti, trav, np.float64, fmm_level_to_order=lambda k, a, t, lev: 8,
            self_extra_kwargs={"target_to_source": actx.from_numpy(
                np.arange(src.shape[1], dtype=np.int32))})
        pot, = drive_fmm(actx, wr, [actx.from_numpy(q)])
        pots.append(actx.to_numpy(pot))
    assert np.linalg.norm(pots[1] - pots[0])/np.linalg.norm(pots[0]) < 1e-6
    data = json.load(open(tmp_path/"autotune.json"))
    (key, val), = data.items()
    assert key.startswith(f"m2l_apply|{autotune.code_version()}|LapKnl3D|8|fp32|")
    assert key.endswith("|tile_width")
    assert tuple(val["policy"]) in autotune.SPLIT_POLICIES
    assert val["tile_width"] in (64, 32)
    assert val["ms"] <= val["default_ms"]
