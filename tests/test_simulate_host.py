"""Host parts of the simulate.py counterpart: the 0.1 mm coarse binning of the stored fields (simulate.py:199-205)."""
import numpy as np

from cerebrosim_amd import simulate


def test_coarse_map_bins_fine_grids():
    coords = np.array([[0, 0, 0], [1, 1, 1], [2, 0, 0], [3, 1, 0], [0, 2, 0]], np.int32)
    uk, inv, cnt, f = simulate.coarse_map({"coords": coords}, 0.05)
    assert f == 2 and uk.dtype == np.int16
    assert uk.tolist() == [[0, 0, 0], [0, 1, 0], [1, 0, 0]]
    assert inv.tolist() == [0, 0, 2, 2, 1] and cnt.tolist() == [2, 1, 2]
    uk1, inv1, cnt1, f1 = simulate.coarse_map({"coords": coords}, 0.1)
    assert f1 == 1 and len(uk1) == len(coords) and cnt1.tolist() == [1] * 5


def test_command_line_needs_a_calibration_for_clot_cases():
    """simulate.py:323-332: only the healthy case calibrates; the others take the healthy run's circuit."""
    import pytest

    from cerebrosim_amd import simulate
    with pytest.raises(SystemExit, match="--cal"):
        simulate.main(["narrow70", "--segment", "unused.json"])
    with pytest.raises(SystemExit):
        simulate.main(["no_such_case", "--segment", "unused.json"])
