"""Comparison of two sets of INFO rows of the same problem (mixed vs FP64 chain).

Column layout of an INFO row (reference src/config.cu:51-56): [input | score, behaviour, margin |
wn, zeta, overshoot, yss | GM_dB, PM_deg, DM_ms, stable, b_PC | (wn, zeta) x nt, energy x nt, unique].
"""
import numpy as np


def columns(problem):
    d = problem.derived()
    return dict(sc=d.info_scoring, fr=d.info_freqresp, am=d.info_allmargin, md=d.info_modal,
                ue=float(problem.config["unstable_err"]))


def pass_flags(problem, info):
    c = columns(problem)
    sc0 = info[:, c["sc"]]
    margin = info[:, c["sc"] + 2]
    swept = info[:, c["am"] + 3] == 1.0
    return (sc0 < c["ue"]) & swept & (margin == 0.0)


def compare(problem, ref, got):
    c = columns(problem)
    am, sc = c["am"], c["sc"]
    both = (ref[:, am + 3] == 1.0) & (got[:, am + 3] == 1.0) & (ref[:, sc + 2] < c["ue"]) & (got[:, sc + 2] < c["ue"])
    both &= (ref[:, sc + 2] < 1e17) & (got[:, sc + 2] < 1e17)

    def dmax(col):
        a, b = ref[both, col], got[both, col]
        sentinel = (a > 1e38) & (b > 1e38)
        d = np.abs(a - b)
        d[sentinel] = 0.0
        return float(d.max()) if d.size else 0.0

    stable_ref = ref[:, sc] < c["ue"]
    stable_got = got[:, sc] < c["ue"]
    pf_ref, pf_got = pass_flags(problem, ref), pass_flags(problem, got)
    s_ref, s_got = ref[:, sc], got[:, sc]
    fin = np.isfinite(s_ref) & np.isfinite(s_got)
    rank_ref = np.argsort(np.argsort(s_ref[fin]))
    rank_got = np.argsort(np.argsort(s_got[fin]))
    nfin = int(fin.sum())
    spear = 1.0 - 6.0 * float(np.sum((rank_ref - rank_got).astype(np.float64) ** 2)) / (nfin * (nfin ** 2 - 1)) if nfin > 2 else 1.0
    beh = (ref[:, sc + 1] < c["ue"]) & (got[:, sc + 1] < c["ue"]) & (ref[:, sc + 1] > 0)
    rel_beh = np.abs(got[beh, sc + 1] - ref[beh, sc + 1]) / np.maximum(np.abs(ref[beh, sc + 1]), 1e-12)
    top = max(1, nfin // 100)
    order_ref = np.argsort(s_ref)[:top]
    order_got = np.argsort(s_got)[:top]
    return dict(
        n=int(ref.shape[0]), n_margin_rows=int(both.sum()),
        max_dgm_db=dmax(am + 0), max_dpm_deg=dmax(am + 1), max_db=dmax(am + 4), max_ddm_ms=dmax(am + 2),
        passfail_mismatch=int(np.sum(pf_ref != pf_got)), n_pass=int(pf_ref.sum()),
        stable_mismatch=int(np.sum(stable_ref != stable_got)),
        behaviour_rel_err_median=float(np.median(rel_beh)) if rel_beh.size else 0.0,
        behaviour_rel_err_p99=float(np.quantile(rel_beh, 0.99)) if rel_beh.size else 0.0,
        score_rank_spearman=spear,
        top1pct_overlap=float(len(set(order_ref) & set(order_got)) / top),
        rank_identical=bool(np.array_equal(np.argsort(s_ref, kind="stable"), np.argsort(s_got, kind="stable"))),
    )
