"""The oracle's SAO encoder decision (src/sao_enc.rs:385-462) against
closed forms worked by hand (tests/sao_decide_cases.py) and the reference's
own tests (src/sao_enc.rs:767-840)."""
import pytest

import sao_decide_cases as cases


@pytest.mark.parametrize("name", sorted(cases.ALL_CASES))
def test_case(orc, vv, name):
    cases.ALL_CASES[name](vv, orc.oracle_sao_decide_picture)


def test_picture_geometry(orc, vv):
    # src/sao_enc.rs:798-806
    src, rec = cases.pic(vv, 32, 32, 8, 100, 100), cases.pic(vv, 32, 32, 8, 128, 128)
    rc, out = cases.run(vv, orc.oracle_sao_decide_picture, src, rec, 5, 8)
    assert rc == 0 and len(out) == 1


@pytest.mark.parametrize("name", sorted(cases.ALL_MERGE_CASES))
def test_merge_case(orc, vv, name):
    cases.ALL_MERGE_CASES[name](vv, orc.oracle_sao_chroma_merge, orc.oracle_sao_decide_picture)
