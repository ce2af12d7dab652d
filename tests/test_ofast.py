"""The reference binary's -Ofast float rewrites (F1..F6 in oracle/erq_oracle.c) against the
literal source semantics that the oracle and the GPU follow, on a small workload shaped like
the headline one (unit-norm power-law mixture, D = 1536, B = 4).  Bar: every re-estimated
distance within 1e-5 of |o - c|^2 + y^2 (the tolerance for FP32 estimated distances), and
the rewrites change almost no top-k id.  The bench-shaped figures come from
oracle/ofast_report.py."""
import os
import sys

import numpy as np

from conftest import REPO


def _report():
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import ofast_report
    return ofast_report


def test_fast_math_rewrites_stay_within_tolerance():
    rep = _report()
    idx, qr = rep.build(20000, 1536, 64, 4, 16, 1.15, seed=9)
    out = rep.measure(idx, qr, 100, 8)
    print(out)
    assert out["max_rel_ex_dist_delta"] < 1e-5
    assert out["topk_ids_changed"] <= 0.01 * out["topk_total"]


def test_fast_math_mode_changes_the_float_results():
    """The switch really reroutes the formulas: some estimate differs in its last bits."""
    from oracle import oracle as orc
    rep = _report()
    idx, qr = rep.build(4000, 1536, 8, 4, 2, 1.15, seed=10)
    oi = orc.OracleIndex(idx)
    pid, pd = oi.centroid_topn(qr[0], 1)
    lo_l, ex_l = oi.pair_estimates(qr[0], pid[0], pd[0])
    with orc.fast_math():
        lo_f, ex_f = oi.pair_estimates(qr[0], pid[0], pd[0])
    assert orc.lib().orc_fast_math() == 0
    assert np.any(ex_l != ex_f) and np.any(lo_l != lo_f)
