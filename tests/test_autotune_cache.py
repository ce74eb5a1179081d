"""CPU: the autotuner's cache keys and file handling (no GPU calls)."""

from __future__ import annotations

import numpy as np


def test_cache_roundtrip_and_keys(tmp_path):
    from hipfdtd import autotune as at

    p = tmp_path / "sub" / "tune.json"
    k1 = at.cache_key("fused", [(256, 256, 256)], np.float32, 12, False, "gfx950")
    k2 = at.cache_key("fused", [(256, 256, 256)], np.float64, 12, False, "gfx950")
    kb = at.cache_key("batch", [(48, 50, 60), (70, 80, 90)], np.float32, 8, False, "gfx950")
    kb2 = at.cache_key("batch", [(70, 80, 90), (48, 50, 60)], np.float32, 8, False, "gfx950")
    assert len({k1, k2, kb, kb2}) == 4
    assert "256x256x256" in k1 and "float32" in k1
    assert at.save_cache({k1: {"rows": 8, "planes": 64}}, p)
    assert at.save_cache({kb: {"blocks": 0, "rows": 4, "planes": 16}}, p)
    data = at.load_cache(p)
    assert data[k1]["planes"] == 64 and data[kb]["rows"] == 4
    assert at.load_cache(tmp_path / "missing.json") == {}
    (tmp_path / "bad.json").write_text("{not json")
    assert at.load_cache(tmp_path / "bad.json") == {}
    # an unwritable location only disables the cache
    blocker = tmp_path / "file"
    blocker.write_text("x")
    assert not at.save_cache({k1: {}}, blocker / "tune.json")
