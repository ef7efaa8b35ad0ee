"""CPU checks of the autotuner's width table (host-only entry points: no GPU
call): unknown boxes have no entry, valid widths are kept, widths that do not
fit a box are refused."""
import pytest


def test_tune_table_roundtrip_and_checks():
    import ifdda
    from ifdda import capi
    assert capi.tune_get(11, 13, 17, 64) is None
    capi.tune_set(11, 13, 17, 64, [1, 2, 1, 2, 1])
    assert capi.tune_get(11, 13, 17, 64) == [1, 2, 1, 2, 1]
    assert capi.tune_get(11, 13, 17, 32) is None  # per precision
    capi.tune_set(11, 13, 17, 64, [2, 1, 2, 1, 2])  # overwrite
    assert capi.tune_get(11, 13, 17, 64) == [2, 1, 2, 1, 2]
    with pytest.raises(ifdda.IfddaError):
        capi.tune_set(11, 13, 17, 64, [0, 1, 1, 1, 1])  # no empty tile
    with pytest.raises(ifdda.IfddaError):
        capi.tune_set(11, 13, 17, 64, [1, 1, 1, 1, 500])  # more lines than P5 has (13*17)
    with pytest.raises(ifdda.IfddaError):
        capi.tune_set(100, 100, 100, 64, [64, 1, 1, 1, 1])  # 64 x-lines of 200 points: > 64 KB LDS
    with pytest.raises(ifdda.IfddaError):
        capi.tune_set(11, 13, 17, 16, [1, 1, 1, 1, 1])
    assert capi.tune_get(100, 100, 100, 64) is None  # refused entries are not kept
