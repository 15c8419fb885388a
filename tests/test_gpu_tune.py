"""The autotuner on the device: every knob is measured, the choice is one of
the candidates, and the measured timings are positive."""
import pytest

pytestmark = pytest.mark.gpu


def test_tune_measures_all_candidates(nla, tmp_path, monkeypatch):
    from nextla_amd import tune
    monkeypatch.setenv("NEXTLA_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    cfg = tune.tune(force=True, dtypes=("f32",), save=True, apply_result=False)
    p = cfg["panel"]["f32"]
    assert p["rows"] in tune.PANEL_ROWS and p["strip"] in tune.PANEL_STRIPS
    assert len(p["ms"]) == len(tune.PANEL_ROWS) * len(tune.PANEL_STRIPS)
    assert all(t > 0 for t in p["ms"].values())
    assert cfg["batch_classes"] in tune.BATCH_MASKS
    # a second call reads the cache instead of measuring again
    again = tune.tune(dtypes=("f32",), apply_result=False)
    assert again == cfg
