"""ctypes binding of the C oracle (oracle/cpu_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests and by the CPU-baseline leg of the
benchmark, never by the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(LIB)
        L.dqc_oracle_c1.restype = C.c_int64
        _lib = L
    return _lib


def _p(a):
    return np.ascontiguousarray(a).ctypes.data_as(C.c_void_p)


def c0(cols, cutoff=19981110, threads=8):
    """dict (g1, g2) -> (sum_qty, sum_price, sum_disc_price, sum_charge, sum_disc, count)."""
    arrs = [np.ascontiguousarray(cols[k]) for k in ("g1", "g2", "qty", "price", "disc", "tax", "date")]
    sums = np.zeros(65536 * 5, dtype=np.float64)
    cnt = np.zeros(65536, dtype=np.int64)
    rc = lib().dqc_oracle_c0(*[_p(a) for a in arrs], C.c_int64(len(arrs[0])), C.c_uint32(cutoff),
                             C.c_int(threads), _p(sums), _p(cnt))
    assert rc == 0
    out = {}
    for k in np.nonzero(cnt)[0]:
        g1, g2 = np.int8(np.uint8(k & 0xFF)), np.int8(np.uint8(k >> 8))
        out[(int(g1), int(g2))] = tuple(sums[k * 5:(k + 1) * 5]) + (int(cnt[k]),)
    return out


def c1(dim, fact, cat=ord("2"), lo=19950101, hi=19960101, threads=8):
    """(keys int32[], revenue float64[]) of the c1 plan."""
    arrs_d = [np.ascontiguousarray(dim[k]) for k in ("d_key", "d_cat")]
    arrs_f = [np.ascontiguousarray(fact[k]) for k in ("f_fk", "f_price", "f_disc", "f_date")]
    cap = len(arrs_d[0]) + 1
    keys = np.zeros(cap, dtype=np.int32)
    sums = np.zeros(cap, dtype=np.float64)
    n = lib().dqc_oracle_c1(_p(arrs_d[0]), _p(arrs_d[1]), C.c_int64(len(arrs_d[0])), C.c_int8(cat),
                            *[_p(a) for a in arrs_f], C.c_int64(len(arrs_f[0])), C.c_uint32(lo), C.c_uint32(hi),
                            C.c_int(threads), _p(keys), _p(sums), C.c_int64(cap))
    assert 0 <= n <= cap
    return keys[:n], sums[:n]


def c2(build, probe, fmax=10, threads=8):
    """(count, sum) of the c2 plan."""
    bk, bp = np.ascontiguousarray(build["b_key"]), np.ascontiguousarray(build["b_pay"])
    pk, pf, pp = (np.ascontiguousarray(probe[k]) for k in ("p_key", "p_f", "p_pay"))
    cnt, s = C.c_int64(), C.c_int64()
    rc = lib().dqc_oracle_c2(_p(bk), _p(bp), C.c_int64(len(bk)), _p(pk), _p(pf), _p(pp), C.c_int64(len(pk)),
                             C.c_int32(fmax), C.c_int(threads), C.byref(cnt), C.byref(s))
    assert rc == 0
    return cnt.value, s.value


def c3(tables, word, threads=8):
    """(by_item, by_nation_year) of the c3 plan, each an OrderedDict of
    columns in the plan's output order."""
    from collections import OrderedDict
    item, cust, nation, sales = tables["item"], tables["cust"], tables["nation"], tables["sales"]
    ik = np.ascontiguousarray(item["i_key"])
    offs, chars = item["i_text"]
    offs, chars = np.ascontiguousarray(offs, dtype=np.uint64), np.ascontiguousarray(chars, dtype=np.uint8)
    ck, cy = np.ascontiguousarray(cust["c_key"]), np.ascontiguousarray(cust["c_year"])
    nk = np.ascontiguousarray(nation["n_key"])
    s = [np.ascontiguousarray(sales[k]) for k in ("s_item", "s_cust", "s_nation", "s_qty", "s_price")]
    ci = len(ik) + 1
    cn = max(1, len(nk)) * 64 + 1
    out_i = [np.zeros(ci, dtype=t) for t in (np.int32, np.float64, np.float64, np.int32)]
    out_n = [np.zeros(cn, dtype=t) for t in (np.int32, np.int32, np.float64, np.float64, np.int32)]
    nny = C.c_int64()
    L = lib()
    L.dqc_oracle_c3.restype = C.c_int64
    n = L.dqc_oracle_c3(_p(ik), _p(offs), _p(chars), C.c_int64(len(ik)), C.c_char_p(word.encode()),
                        _p(ck), _p(cy), C.c_int64(len(ck)), _p(nk), C.c_int64(len(nk)),
                        *[_p(a) for a in s], C.c_int64(len(s[0])), C.c_int(threads),
                        *[_p(a) for a in out_i], C.c_int64(ci), *[_p(a) for a in out_n], C.c_int64(cn),
                        C.byref(nny))
    assert 0 <= n <= ci and 0 <= nny.value <= cn
    byItem = OrderedDict(zip(["i_key", "sum_qty", "revenue", "cnt"], [a[:n] for a in out_i]))
    byNy = OrderedDict(zip(["n_key", "c_year", "sum_qty", "revenue", "cnt"], [a[:nny.value] for a in out_n]))
    return byItem, byNy
