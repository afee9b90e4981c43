"""Shared parity check of a device search result against the FP64 CPU oracle (imported by
the GPU parity tests; not a test module itself).

Bar: in FP64 mode every period's best (width, position) index equals the oracle's and chi2
agrees to 1e-9 relative; in FP32-compensated mode chi2 agrees to 2e-6 relative and every
period whose best index differs is a near tie -- the oracle's own statistic at the device's
index lies within 2e-6 relative of the oracle's minimum -- and such periods are at most 1 %
(at least 2 allowed).
"""
import numpy as np

from oracle import oracle

FP64_RTOL = 1e-9
FP32_RTOL = 2e-6


def check(plan, mode, skip, periods, k_lo, k_hi, c_g, k_g, s_g, label=""):
    """Asserts the bar above for one curve; returns (max relative chi2 error, number of
    periods with the oracle's best index, number of non-empty periods)."""
    c_o, k_o, s_o = oracle.search_plan(plan, skip_point=skip, periods=periods, k_lo=k_lo,
                                       k_hi=k_hi)
    empty = np.asarray(k_lo) > np.asarray(k_hi)
    assert np.all(np.isnan(c_g[empty])) and np.all(k_g[empty] == -1), label
    ok = ~empty
    rel = np.abs(c_g[ok] - c_o[ok]) / np.abs(c_o[ok])
    worst = float(rel.max(initial=0.0))
    same = (k_g == k_o) & (s_g == s_o)
    n_same = int(np.sum(same[ok]))
    if mode == "fp64":
        assert worst < FP64_RTOL, (label, worst)
        bad = np.nonzero(ok & ~same)[0]
        assert bad.size == 0, (label, bad[:10], k_g[bad[:10]], k_o[bad[:10]], s_g[bad[:10]],
                               s_o[bad[:10]])
    else:
        assert worst < FP32_RTOL, (label, worst)
        bad = np.nonzero(ok & ~same)[0]
        tb = plan.table
        for i in bad:
            table, cmin, *_ = oracle.period_table(plan.t, plan.y, plan.dy, periods[i],
                                                  tb.widths, tb.S, tb.offsets, tb.overshoot,
                                                  k_lo[i], k_hi[i], skip_point=skip)
            at = table[k_g[i] - k_lo[i], s_g[i]]
            assert (at - cmin) <= FP32_RTOL * abs(cmin), (label, i, at, cmin)
        assert bad.size <= max(2, ok.sum() // 100), (label, bad.size, ok.sum())
    return worst, n_same, int(ok.sum())


def report(config, mode, worst, n_same, n_ok, extra=""):
    """One parseable line per configuration (collected for the error table)."""
    print("PARITY config=%s mode=%s periods=%d max_rel_chi2=%.3e best_index_match=%d/%d%s"
          % (config, mode, n_ok, worst, n_same, n_ok, (" " + extra) if extra else ""),
          flush=True)
