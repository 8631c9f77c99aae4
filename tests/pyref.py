"""Second, independent restatement of the reference pass loop in plain Python.

Test infrastructure only.  It follows the kernels of
spikesorting_fullpursuit/kernels/binary_pursuit.cl and the host loop of
spikesorting_fullpursuit/parallel/binary_pursuit_parallel.py:532-736 line by
line, with scalar float32 arithmetic emulated exactly (fma32 below rounds the
exact a*b+c once to float32), for tiny inputs only.  Comparing it bit for bit
with the C oracle (oracle/fbp_oracle.c) catches transcription mistakes in
either: they were written separately, in different styles.
"""
from fractions import Fraction

import numpy as np

f32 = np.float32


def fma32(a, b, c):
    """float32(a*b + c) with a single rounding (IEEE fusedMultiplyAdd)."""
    a, b, c = float(a), float(b), float(c)
    p = a * b                      # exact: 24 x 24 bit product fits a double
    s = p + c
    bp = s - c                     # TwoSum: err is the exact rounding error of s
    ap = s - bp
    err = (p - bp) + (c - ap)
    r = f32(s)
    if err != 0.0 and float(r) != s:
        lo = np.nextafter(r, f32(-np.inf))
        hi = np.nextafter(r, f32(np.inf))
        for m in ((float(lo) + float(r)) / 2.0, (float(r) + float(hi)) / 2.0):
            if m == s:   # s sits on a float32 midpoint: the exact value decides
                exact = Fraction(a) * Fraction(b) + Fraction(c)
                below = lo if m < float(r) else r
                above = r if m < float(r) else hi
                return above if exact > Fraction(m) else below
    return r


def likelihood_at(V, Tm, bias, t, i, K, N, C):
    """bp.cl:216-252."""
    if i + N >= K:
        return f32(0.0)
    acc = f32(bias[t])
    for c in range(C):
        for j in range(N):
            acc = fma32(Tm[t, c, j], V[c, i + j], acc)
    return acc


def overlap_score(V, Tm, thr, L, lab, bidx, t, f, s, K, N, C, M):
    """bp.cl:516-632 for one work item."""
    if L[lab, bidx] < thr[lab] or bidx >= K - N:
        return f32(0.0)
    if f + bidx < M or f + bidx >= K - N + M or s + bidx < M or s + bidx >= K - N + M:
        return f32(0.0)
    if abs(f - s) > N:
        return f32(0.0)
    af, as_ = bidx + f - M, bidx + s - M
    if L[lab, af] <= 0 and L[t, as_] <= 0:
        return f32(0.0)
    fixed_align, shift_align = bidx - af, bidx - as_
    acc = f32(0.0)
    for c in range(C):
        ss = f32(0.0)
        for j in range(N):
            fv = Tm[lab, c, fixed_align + j] if 0 <= fixed_align + j < N else f32(0.0)
            sv = Tm[t, c, shift_align + j] if 0 <= shift_align + j < N else f32(0.0)
            summed = f32(fv + sv)
            ss = fma32(summed, summed, ss)
            acc = fma32(summed, V[c, bidx + j], acc)
        if ss > 0:
            acc = f32(float(acc) - 0.5 * float(ss))
    return acc


def pursuit(V, templates, bias, thr, N, M, do_overlap=True, WG=16, max_passes=1000):
    """Returns (onsets, labels, residual, passes)."""
    V = np.array(V, dtype=np.float32)
    C, K = V.shape
    T = templates.shape[0]
    Tm = np.asarray(templates, dtype=np.float32).reshape(T, C, N)
    bias = np.asarray(bias, dtype=np.float32)
    thr = np.asarray(thr, dtype=np.float32)
    W = -(-K // N)
    nfloor = K // N
    S = 2 * M + 1
    items = S * S * T
    nl = (items - 1) // WG + 1
    L = np.zeros((T, K), dtype=np.float32)
    best_L = np.zeros(W, dtype=np.float32)
    best_lab = np.zeros(W, dtype=np.int64)
    best_idx = np.zeros(W, dtype=np.int64)
    ov_idx = np.zeros(W, dtype=np.int64)
    ov_lab = np.zeros(W, dtype=np.int64)
    nxt = np.zeros(W + 2, dtype=np.int64)
    ovf = np.zeros(W + 2, dtype=np.int64)
    windows = list(range(W))
    out_idx, out_lab = [], []
    passes = 0
    while passes < max_passes:
        passes += 1
        for w in windows:                                   # compute_full_likelihood
            for t in range(T):
                for i in range(w * N, min((w + 1) * N, K)):
                    L[t, i] = likelihood_at(V, Tm, bias, t, i, K, N, C)
        for t in range(T):                                  # T launches of the argmax kernel
            for w in windows:
                sw, ew = w * N, (w + 1) * N
                if ew >= K - N:
                    continue
                start = 0 if N > sw else sw - N
                stop = min(ew + N, K)
                b = f32(0.0) if t == 0 else best_L[w]
                lab, idx = best_lab[w], best_idx[w]
                chk = False
                for i in range(start, stop):
                    v = L[t, i]
                    if v > b:
                        b, lab, idx = v, t, i
                    if v > 0 and sw <= i < ew:
                        chk = True
                if chk:
                    nxt[w] = 1
                ovf[w] = 1 if (b > 0 and sw <= idx < ew) else 0
                best_L[w], best_lab[w], best_idx[w] = b, lab, idx
                ov_idx[w], ov_lab[w] = idx, lab
        ov_list = [w for w in range(W) if ovf[w]]
        if ov_list and do_overlap:
            for w in ov_list:                               # overlap_recheck_indices + parse
                lab, bidx = best_lab[w], best_idx[w]
                groups = []
                for g in range(nl):
                    vals = []
                    for lid in range(WG):
                        q = (g * WG + lid) % items
                        vals.append(overlap_score(V, Tm, thr, L, lab, bidx, q % T, (q // T) % S,
                                                  (q // (S * T)) % S, K, N, C, M))
                    ids = list(range(WG))
                    half = WG // 2
                    while half > 0:
                        for lid in range(half):
                            if vals[lid + half] > vals[lid]:
                                vals[lid], ids[lid] = vals[lid + half], ids[lid + half]
                        half //= 2
                    groups.append((vals[0], ids[0] + WG * g))
                bl = best_L[w]
                bg, bsid = bl, 0
                for v, gid in groups:
                    if v > bg:
                        bg, bsid = v, gid
                if bg <= bl or bg <= 0 or bsid % (nl * WG) >= items:
                    ovf[w] = 0
                    continue
                t, f, s = bsid % T, (bsid // T) % S, (bsid // (S * T)) % S
                af, as_ = bidx + f - M, bidx + s - M
                Lm, Lo = L[lab, af], L[t, as_]
                if bg < Lm or bg < Lo or Lm <= 0 or Lo <= 0:
                    ovf[w] = 0
                elif Lm >= Lo and Lm > thr[lab]:
                    best_L[w], ov_lab[w], ov_idx[w] = Lm, lab, af
                elif Lo > Lm and Lo > thr[t]:
                    best_L[w], ov_lab[w], ov_idx[w] = Lo, t, as_
                elif Lm > thr[lab]:
                    best_L[w], ov_lab[w], ov_idx[w] = Lm, lab, af
                else:
                    ov_idx[w], ov_lab[w], ovf[w] = best_idx[w], best_lab[w], 0
            parsed = ovf.copy()                             # check_overlap_reassignments (snapshot)
            for w in ov_list:
                if parsed[w] == 0:
                    continue
                best_idx[w], best_lab[w] = ov_idx[w], ov_lab[w]
                if ov_idx[w] < w * N and w > 1:
                    if parsed[w - 2] == 1:
                        ovf[w] = 0
                        best_idx[w] = 0
                        nxt[w - 1] = nxt[w] = nxt[w + 1] = 2
                    else:
                        nxt[w - 2] = 2
                if ov_idx[w] >= (w + 1) * N and w + 2 < nfloor:
                    nxt[w + 2] = 2
        added, has = [], {}
        for w in windows:                                   # binary_pursuit: adds
            sw, ew = w * N, (w + 1) * N
            ml, lab, idx = (best_L[w], best_lab[w], best_idx[w]) if ew < K - N else (f32(0.0), 0, 0)
            hs = (ml > thr[lab] or ovf[w] == 1) and ((sw <= idx < ew) or ovf[w] == 1)
            has[w] = hs
            if hs:
                added.append((idx, lab))
                nxt[w] = 3
                if w > 0:
                    nxt[w - 1] = 3
                if w + 1 < nfloor:
                    nxt[w + 1] = 3
        snap = nxt.copy()                                   # binary_pursuit: pruning
        for w in windows:
            sw, ew = w * N, (w + 1) * N
            if not ew < K - N or has[w]:
                continue
            own = snap[w]
            if best_idx[w] < sw and own == 1 and w > 0 and snap[w - 1] == 0:
                own = 0
            if best_idx[w] >= ew and own == 1 and w + 1 < nfloor and snap[w + 1] == 0:
                own = 0
            if best_L[w] <= thr[best_lab[w]] and own < 3:
                own = 0
            nxt[w] = own
        if not added:
            break
        ovf[:] = 0
        windows = [w for w in range(W) if nxt[w]]
        nxt[:] = 0
        for idx, lab in added:
            out_idx.append(idx)
            out_lab.append(lab)
        for idx, lab in sorted(added):                      # compute_residual, ordered
            for c in range(C):
                for j in range(N):
                    if idx + j >= K:
                        break
                    V[c, idx + j] = f32(V[c, idx + j] - Tm[lab, c, j])
    return np.array(out_idx, dtype=np.uint32), np.array(out_lab, dtype=np.uint32), V, passes
