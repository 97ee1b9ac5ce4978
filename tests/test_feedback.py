"""runtime.Feedback on the host: what it learns from a run's status block
and the capacities it sets for the next run (no GPU)."""
import ctypes as C

import numpy as np

from dogqc_amd import runtime as rt


class _Module(object):
    def __init__(self):
        self.path = "/x/plan_abc.so"
        self.desc = {"hts": [{"index": 0, "kind": "unique"}, {"index": 1, "kind": "multi"},
                             {"index": 2, "kind": "agg"}],
                     "temps": [{"name": "t"}], "result": [{"name": "r"}]}


class _Inputs(object):
    def __init__(self, rows):
        self.rows = np.array(rows, dtype=np.int64)


class _Exec(object):
    def __init__(self):
        self.res = rt.QueryResult()
        st = self.res.status
        st.ht_fill[0], st.ht_fill[2] = 3000, 120
        st.mm_total[1] = 50000
        st.temp_count[0] = 777
        st.result_count = 99
        st.split_count[1] = 12
        self.res.layout.split_capacity[1] = 64


def test_learn_then_apply_at_same_and_double_size():
    fb = rt.Feedback()
    mod = _Module()
    fb.learn(mod, _Inputs([1000, 40000]), _Exec())
    t = rt.newTuning()
    assert fb.apply(mod, _Inputs([1000, 40000]), t)
    assert t.ht_capacity[0] == 6000 and t.ht_capacity[2] == 240
    assert t.mm_entries[1] == 50000 and t.ht_capacity[1] == 50000 and t.split_capacity[1] == 12
    assert t.temp_capacity[0] == 777 and t.result_capacity == 99
    t = rt.newTuning()
    fb.apply(mod, _Inputs([1000, 80000]), t)          # the largest input doubled
    assert t.ht_capacity[0] == 12000 and t.mm_entries[1] == 100000 and t.result_capacity == 198


def test_unknown_module_and_roundtrip(tmp_path):
    fb = rt.Feedback()
    t = rt.newTuning()
    assert not fb.apply(_Module(), _Inputs([5]), t)
    assert t.ht_capacity[0] == 0
    fb.learn(_Module(), _Inputs([10]), _Exec())
    p = str(tmp_path / "fb.json")
    fb.save(p)
    fb2 = rt.Feedback()
    fb2.load(p)
    assert fb2.records == fb.records


def test_census_flag_and_fill_fields():
    t = rt.newTuning()
    assert hasattr(t, "census") and C.sizeof(t) == 1640
    assert C.sizeof(rt.Status) == 856
