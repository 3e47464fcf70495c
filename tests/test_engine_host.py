"""The device integral engine (nrx_engine.h) compiled for the host, checked
against the oracle before any GPU run.  The same source is what the HIP
kernels execute."""
import numpy as np
import pytest

import oracle
from neorx import synth
from _util import check_close, engine_prim_list, random_tuples


@pytest.mark.parametrize("cls,n", [(1, 400), (2, 200), (3, 60)])
def test_engine_tuples_vs_oracle(cls, n):
    s = synth.small(seed=40 + cls, e_shells=(0, 1, 2), p_shells=(0, 1, 2))
    idx = random_tuples(s, cls, n, seed=7 * cls)
    ee, es = engine_prim_list(s, cls, idx)
    oe, os_ = oracle.prim_list(s, cls, idx)
    check_close(ee, oe, what=f"engine vs oracle cls{cls} E")
    if cls < 3:
        check_close(es, os_, what=f"engine vs oracle cls{cls} S")


def test_engine_config2_sample():
    s = synth.config2()
    for cls, n in [(1, 200), (2, 100), (3, 30)]:
        idx = random_tuples(s, cls, n, seed=cls)
        ee, _ = engine_prim_list(s, cls, idx)
        oe, _ = oracle.prim_list(s, cls, idx)
        check_close(ee, oe, what=f"config2 cls{cls}")


def test_engine_large_T_branch():
    """Far-apart centres push the Boys argument past 20 (asymptotic branch)."""
    s = synth.small(seed=9, e_shells=(0, 1), p_shells=(0, 1))
    s["elcbfc"] = np.asfortranarray(s["elcbfc"] * 4.0)
    s["elcex"] = s["elcex"] * 3.0
    idx = random_tuples(s, 2, 120, seed=3)
    ee, _ = engine_prim_list(s, 2, idx)
    oe, _ = oracle.prim_list(s, 2, idx)
    check_close(ee, oe, what="large-T")
