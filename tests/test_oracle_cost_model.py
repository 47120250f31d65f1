"""CPU pins of the oracle's cost-model restatement (oracle/cost_model.py):
values evaluated by hand from the reference formulas and constants
(MY_sparse/my_cost_model.py:786-816, 526-577, 820-860, 1219-1248)."""
import pytest

from oracle import cost_model as oc


# --- known answers, evaluated by hand from my_cost_model.py:786-816, 526-577, 1219-1248

def test_ell_cost_known_answers():
    # non-atomic, fp32 coefficients: 8.23071690e-08 * 1000 + 3.42962663e-03
    assert oc.ell_cost_model(1000.0, None, False, 32) == pytest.approx(3.51193380e-03, rel=1e-8)
    # (8, 32, 32) tile, 20 distinct columns, fp16: memory 16 + 16 + 40 = 72 units
    # -> x = 72 * 32 = 2304 -> 8.80425616e-08 * 2304 + 4.17697078e-04
    assert oc.ell_tile_cost(8, 32, 32, 20, False, 16) == pytest.approx(6.20547140e-04, rel=1e-8)
    # atomic: x_atomic = 88 * 32 = 2816 -> * (8.80688635 * 2816 / 2304 - 7.9094305)
    assert oc.ell_tile_cost(8, 32, 32, 20, True, 16) == pytest.approx(1.77137764e-03, rel=1e-7)


def test_sddmm_1d_and_lower_bound_known_answers():
    # rows 4, cols 30, K 32, bucket 32, fp32: mem = ((4+30)*32 + 32) * 4 = 4480 bytes,
    # x = 4480 -> 5.90185816e-08 * 4480 + 1.83606878e-04
    assert oc.sddmm_1d_cost(4, 30, 32, 32, 32) == pytest.approx(4.480101236e-04, rel=1e-8)
    # mbs 32, fp16: best shape (8,32,32) / (16,32,16): memory 66 units -> x = 2112
    # -> (8.80425616e-08 * 2112 + 4.17697078e-04) / 8192
    shapes = [(256, 32, 1), (128, 32, 2), (64, 32, 4), (32, 32, 8), (16, 32, 16), (8, 32, 32)]
    assert oc.lower_bound_ell(shapes, 16) == pytest.approx(6.03642968e-04 / 8192, rel=1e-7)
