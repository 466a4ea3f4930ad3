"""Automatic configuration (jc_config_from_stats, the rule behind
jc_select_config): host only, no device. The reference leaves this to a TODO
(main.cpp:203)."""
import ctypes

import pytest

from jaccard_amd import _native, api


def stats_of(degrees):
    """jc_graph_stats of a degree sequence, computed on the host."""
    st = api.JcGraphStats()
    for d in degrees:
        if d == 0:
            st.isolated += 1
            continue
        b = d.bit_length() - 1
        st.rows[b] += 1
        st.entries[b] += d
        st.max_degree = max(st.max_degree, d)
        if d > 128:
            st.long_rows += 1
    return st


@pytest.mark.parametrize("weighted", [False, True])
def test_rule_below_quotient_limit(weighted):
    st = stats_of([0, 3, 17, 100, 128, 129, 5000])
    ec, t = api.config_from_stats(st, 1 << 20, weighted)
    assert ec is True
    assert t == {"small_max": 128, "small_blocks_per_cu": 32, "wide_index": 0, "bitmap_blocks_per_cu": 1,
                 "hash_table": 0, "quot_max_disp": 0}


def test_rule_from_quotient_limit_on():
    st = stats_of([200] * 4)
    _, t = api.config_from_stats(st, (1 << 24) - 1, True)
    assert t["hash_table"] == 0
    _, t = api.config_from_stats(st, 1 << 24, True)
    assert t["hash_table"] == 1


def test_rule_argument_errors():
    lib = _native.load_jaccard_lib()
    st = stats_of([1, 2])
    ec, t = ctypes.c_int(), api.JcTuning()
    assert lib.jc_config_from_stats(None, 4, 0, ctypes.byref(ec), ctypes.byref(t)) == api.ERR_INVALID_ARG
    assert lib.jc_config_from_stats(ctypes.byref(st), -1, 0, ctypes.byref(ec), ctypes.byref(t)) == api.ERR_INVALID_ARG
    assert lib.jc_config_from_stats(ctypes.byref(st), 4, 0, None, ctypes.byref(t)) == api.ERR_INVALID_ARG
    assert lib.jc_select_config(None, 0, ctypes.byref(ec), ctypes.byref(t), None) == api.ERR_INVALID_ARG
    assert lib.jc_degree_stats(None, ctypes.byref(st), None) == api.ERR_INVALID_ARG
    # an empty graph needs no device: all-zero statistics
    empty = api.JcCsr(0, 0, None, None, None)
    z = api.JcGraphStats()
    assert lib.jc_degree_stats(ctypes.byref(empty), ctypes.byref(z), None) == 0
    assert z.max_degree == 0 and sum(z.rows) == 0 and z.isolated == 0
    # the table budgets are reported without a graph: the weighted layout
    # leaves less room (its staging area holds weights too)
    assert 0 < z.bitmap_table_bytes[1] < z.bitmap_table_bytes[0] <= 160 * 1024
