"""Comparison of two sets of INFO rows of the same problem (mixed vs FP64 chain), and the pass rule.

Column layout of an INFO row (reference src/config.cu:51-56): [input | score, behaviour, margin |
wn, zeta, overshoot, yss | GM_dB, PM_deg, DM_ms, stable, b_PC | (wn, zeta) x nt, energy x nt, unique].

A candidate passes (reference README.md:22-35: overshoot, steady-state error, gain / phase /
Vinnicombe margins) when its loop is linearly and nominally stable and its sweep completed
(stable = 1, score below unstable_err, margin cell not a penalty), the margin error is zero (GM,
PM and b meet their targets), the overshoot is within its spec and |yss / ref - 1| within the
steady-state spec.  Damping and bandwidth enter the score as squared deviations with no pass band.
"""
import numpy as np


def columns(problem):
    d = problem.derived()
    c = problem.config
    return dict(sc=d.info_scoring, fr=d.info_freqresp, am=d.info_allmargin, md=d.info_modal,
                ue=float(np.float32(c["unstable_err"])), os=float(np.float32(c["overshoot"])),
                sse=float(np.float32(c["steady_state_error"])), amp=float(np.float32(c["ref_amplitude"])))


def pass_flags(problem, info):
    c = columns(problem)
    sc0, margin = info[:, c["sc"]], info[:, c["sc"] + 2]
    swept = info[:, c["am"] + 3] == 1.0
    os_ok = info[:, c["fr"] + 2] <= c["os"]
    ss_ok = np.abs(info[:, c["fr"] + 3] / c["amp"] - 1.0) <= c["sse"]
    return (sc0 < c["ue"]) & swept & (margin == 0.0) & os_ok & ss_ok


def margin_pass_flags(problem, info):
    """The margin part of the pass rule alone: stable, swept, GM, PM and b at or above their targets
    (the cells the mixed chain's FP32 sweep decides)."""
    c = columns(problem)
    return (info[:, c["sc"]] < c["ue"]) & (info[:, c["am"] + 3] == 1.0) & (info[:, c["sc"] + 2] == 0.0)


def ranking(problem, info):
    """Candidate indices by ascending score, ties by index; rows whose score is not a number last."""
    s = info[:, columns(problem)["sc"]]
    key = np.where(np.isnan(s), np.inf, s)
    order = np.lexsort((np.arange(s.size), np.isnan(s), key))
    return order


def compare(problem, ref, got):
    c = columns(problem)
    am, sc, fr = c["am"], c["sc"], c["fr"]
    both = (ref[:, am + 3] == 1.0) & (got[:, am + 3] == 1.0) & (ref[:, sc + 2] < c["ue"]) & (got[:, sc + 2] < c["ue"])
    both &= (ref[:, sc + 2] < 1e17) & (got[:, sc + 2] < 1e17)
    fitted = (ref[:, sc] < c["ue"]) & (got[:, sc] < c["ue"]) & np.isfinite(ref[:, fr]) & np.isfinite(got[:, fr])

    def dmax(col, rows=both, rel=False):
        a, b = ref[rows, col], got[rows, col]
        sentinel = (a > 1e38) & (b > 1e38)
        d = np.abs(a - b)
        if rel:
            d = d / np.maximum(np.abs(a), 1e-300)
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
    rk_ref, rk_got = ranking(problem, ref), ranking(problem, got)

    def ranked(flags_ref, flags_got):
        """The ranked lists of the flagged candidates: same members in the same order?  With the
        FP64 score gaps of adjacent members whose order differs."""
        a_ref, a_got = rk_ref[flags_ref[rk_ref]], rk_got[flags_got[rk_got]]
        swaps = []
        if np.array_equal(np.sort(a_ref), np.sort(a_got)):
            pos = np.empty(ref.shape[0], dtype=np.int64)
            pos[a_got] = np.arange(a_got.size)
            for a, b in zip(a_ref[:-1], a_ref[1:]):
                if pos[a] > pos[b]:
                    swaps.append(float(abs(s_ref[b] - s_ref[a]) / max(abs(s_ref[a]), 1e-300)))
        return bool(np.array_equal(a_ref, a_got)), swaps

    pass_identical, swaps = ranked(pf_ref, pf_got)
    mp_ref, mp_got = margin_pass_flags(problem, ref), margin_pass_flags(problem, got)
    mp_identical, mp_swaps = ranked(mp_ref, mp_got)
    return dict(
        n=int(ref.shape[0]), n_margin_rows=int(both.sum()),
        max_dgm_db=dmax(am + 0), max_dpm_deg=dmax(am + 1), max_db=dmax(am + 4), max_ddm_ms=dmax(am + 2),
        max_dovershoot=dmax(fr + 2, fitted), max_rel_dwn=dmax(fr + 0, fitted, rel=True),
        max_rel_dyss=dmax(fr + 3, fitted, rel=True),
        passfail_mismatch=int(np.sum(pf_ref != pf_got)), n_pass=int(pf_ref.sum()),
        pass_rank_identical=pass_identical,
        pass_rank_swaps=len(swaps), pass_rank_swap_max_rel_gap=max(swaps) if swaps else 0.0,
        margin_pass_mismatch=int(np.sum(mp_ref != mp_got)), n_margin_pass=int(mp_ref.sum()),
        margin_pass_rank_identical=mp_identical, margin_pass_rank_swaps=len(mp_swaps),
        margin_pass_rank_swap_max_rel_gap=max(mp_swaps) if mp_swaps else 0.0,
        stable_mismatch=int(np.sum(stable_ref != stable_got)),
        behaviour_rel_err_median=float(np.median(rel_beh)) if rel_beh.size else 0.0,
        behaviour_rel_err_p99=float(np.quantile(rel_beh, 0.99)) if rel_beh.size else 0.0,
        score_rank_spearman=spear,
        top1pct_overlap=float(len(set(order_ref) & set(order_got)) / top),
        rank_identical=bool(np.array_equal(np.argsort(s_ref, kind="stable"), np.argsort(s_got, kind="stable"))),
    )
