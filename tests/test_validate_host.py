"""Host parts of the validate.py counterpart: the Young & Tsai pressure drop against the reference's published
values (geometry/solver.json:114-177, young_tsai_Pa), and the published table the comparison reads."""
from cerebrosim_amd import validate
from cerebrosim_amd.validation import NU_BLOOD


def test_young_tsai_matches_published():
    D0, sev, Ls = 2e-3, 0.7, 4e-3
    for Re0, ref in ((100, 4701.098008385743), (200, 13891.83671092476)):
        U0 = Re0 * NU_BLOOD / D0
        assert abs(validate.young_tsai(U0, D0, D0 * (1 - sev), Ls) / ref - 1) < 1e-12, Re0


def test_published_table_complete():
    assert set(validate.PUBLISHED["poiseuille"]) == {(10, True), (20, True), (40, True), (20, False)}
    assert set(validate.PUBLISHED["womersley"]) == {1.4, 2.8, 5.6}
    assert len(validate.PUBLISHED["open_pipe"]) == 2 and len(validate.PUBLISHED["stenosis"]) == 3
