"""Instruction decode (reference src/main.cu:317-337) and interpreter (src/main.cu:260-291), oracle
restatement, against known answers built from README.md:60-91."""
import math

import numpy as np
import pytest

from controllerengine_amd.problem import encode
from oracle import oracle

ADD, SUB, MUL, DIV, SAT, SGN = 1, 2, 3, 4, 5, 6
STATE, PREV = 0, 35
K_CONST, K_STATE, K_PREV = 0, 1, 2
BASE = np.array([-2.0, 2.0, 1.0, 0.5, 3.0], dtype=np.float32)


def dec(rows, dim):
    rows = np.array(rows, dtype=np.int32).reshape(-1, dim)
    return oracle.decode(rows, dim, rows.shape[0], BASE, len(BASE))


def test_decode_digits_columns_and_constants():
    iops, vops = dec([[0, encode(SAT, 1, 2)], [encode(MUL, STATE, 3), 0], [encode(ADD, PREV, 5), 0, 0][:2]], 2)
    assert list(iops[0]) == [SAT, 1, K_CONST, K_CONST] and list(vops[0]) == [-2.0, 2.0]
    assert list(iops[1]) == [MUL, 0, K_STATE, K_CONST] and vops[1][1] == 1.0
    assert list(iops[2]) == [ADD, 0, K_PREV, K_CONST] and vops[2][1] == 3.0


def test_decode_last_nonzero_column_wins_and_skip_rows():
    iops, _ = dec([[encode(ADD, 0, 3), 0, encode(MUL, 0, 3)], [0, 0, 0], [0, 38, 0]], 3)
    assert iops[0][0] == MUL and iops[0][1] == 2        # last nonzero entry gives op and column
    assert iops[1][0] == 0                               # all-zero row is SKIP
    assert iops[2][0] == 0 and iops[2][1] == 1           # 38 < 1296: op digit 0 (SKIP) at column 1


def test_decode_rejects_missing_params_and_negative_entries():
    with pytest.raises(ValueError, match="instruction digit 6 needs BASE_PARAMS\\[5\\] but n_params is 5"):
        dec([[encode(ADD, STATE, 6)]], 1)
    with pytest.raises(ValueError, match="instruction digit 20 needs BASE_PARAMS"):
        dec([[200]], 1)                                 # digits are checked even for a SKIP op digit
    with pytest.raises(ValueError, match="negative"):
        dec([[-5]], 1)


def run(rows, dim, src, seed=None):
    iops, vops = dec(rows, dim)
    return oracle.interp(iops, vops, np.array(src, dtype=np.float64), seed)


@pytest.mark.parametrize("op,fn", [(ADD, lambda x, c: x + c), (SUB, lambda x, c: x - c), (MUL, lambda x, c: x * c),
                                   (DIV, lambda x, c: x / c)])
def test_binary_ops_state_const(op, fn):
    for x in [0.75, -3.5, 1e-8, 12345.0]:
        assert run([[0, encode(op, STATE, 5)]], 2, [9.0, x]) == fn(x, 3.0)


def test_sat_clamps_previous_value_and_sign():
    for x, want in [(5.0, 2.0), (-7.0, -2.0), (0.3, 0.3)]:
        assert run([[encode(SAT, 1, 2), 0]], 2, [x, 0.0]) == want   # first instruction: prev = state
    for x, want in [(4.0, 1.0), (-0.1, -1.0), (0.0, 0.0), (float("nan"), 0.0)]:
        assert run([[encode(SGN, STATE, STATE)]], 1, [x]) == want


def test_chain_prev_seed_and_nan_seed():
    rows = [[encode(MUL, STATE, 4), 0], [encode(ADD, PREV, 3), 0]]       # x0 * 0.5 + 1
    assert run(rows, 2, [6.0, 0.0]) == 4.0
    sat = [[encode(SAT, 1, 2), 0]]                                        # clamp(seed, -2, 2)
    assert run(sat, 2, [0.5, 0.0], seed=10.0) == 2.0
    assert run(sat, 2, [0.5, 0.0], seed=float("nan")) == 0.5             # NaN seed: state instead
    assert run([[0], [0]], 1, [3.0]) == 0.0                               # all-SKIP chain is 0
    assert run([[encode(7, 0, 0)], [encode(ADD, PREV, 3)]], 1, [3.0]) == 1.0  # op>=7 marks started
    div0 = run([[encode(DIV, STATE, PREV)]], 1, [0.0])
    assert math.isnan(div0)                                               # 0/0 with prev = state


def test_y_slot_is_last_source_entry():
    # NL chains see [x_plant..., y]; column nd addresses y (reference src/main.cu:377)
    rows = [[0, 0, encode(MUL, STATE, 3)]]
    assert run(rows, 3, [1.0, 2.0, 7.5]) == 7.5
