"""Hand-built zero-noise scenarios that pin the decision rules of binary
pursuit whose outcome is not fixed by a closed form: the +/-2-window
reassignment of check_overlap_reassignments (binary_pursuit.cl:822-875) and
the flag writes and pruning of the binary_pursuit kernel (binary_pursuit.cl:
918-1063).

Every template is a few delta samples with dyadic amplitudes, so every
likelihood value below is exact in float32 whatever the summation order, and
every expected value is written down from the construction and the rules as
the reference states them, not computed by any restatement of the algorithm.
Notation: N = 16 samples per template, window w covers [16w, 16w + 16), the
argmax of window w scans [16w - 16, 16w + 32), templates first, then ascending
sample, strict > (binary_pursuit.cl:350-441).

Each scenario gives, per pass p = 1, 2, ..., the spikes that pass adds (onset,
label) in ascending window order, and the window list that pass p + 1
evaluates (check_window_on_next_pass != 0 after pass p, binary_pursuit_
parallel.py:683-689).  A test runs the pursuit with max_passes = p and
compares the spike list and the pending window list after each pass.
"""
import numpy as np

N = 16
M = 4


def _delta_templates(C, spec):
    """spec: list of [(channel, sample, amplitude), ...] per template."""
    tm = np.zeros((len(spec), C, N), dtype=np.float32)
    for t, deltas in enumerate(spec):
        for c, j, a in deltas:
            tm[t, c, j] = a
    return tm


def _place(V, tm, onset, label, scale=1.0):
    V[:, onset:onset + N] += np.float32(scale) * tm[label]


def _finish(name, V, tm, thr, passes, lists, truth, do_overlap=True, residual_zero=True):
    T = tm.shape[0]
    ss = np.sum(tm.astype(np.float64).reshape(T, -1) ** 2, axis=1)
    return dict(name=name, voltage=V, templates=tm.reshape(T, -1), bias=np.float32(-0.5 * ss),
                thr=np.float32(thr), N=N, M=M, do_overlap=do_overlap, passes=passes, lists=lists,
                truth=truth, residual_zero=residual_zero)


# Templates shared by the reassignment scenarios (one channel):
#   A (label 0): deltas of 1 at samples 0 and 4, |A|^2 = 2, bias -1
#   B (label 1): delta of 1.25 at sample 0, |B|^2 = 1.5625, bias -0.78125
# so L_A(i) = V[i] + V[i+4] - 1 and L_B(i) = 1.25 V[i] - 0.78125.
_AB = [[(0, 0, 1.0), (0, 4, 1.0)], [(0, 0, 1.25)]]
_AB_THR = [0.5, 0.5]


def give_up_left():
    """Windows 3 and 5 both move their spike left out of the window; window 5
    gives up (binary_pursuit.cl:847-864) and its spike is added a pass later.

    Pattern at o: A at o, B at o + 8 (V = 1 at o, 1 at o+4, 1.25 at o+8).
    L_A(o) = 1, L_A(o+4) = 1.25, L_A(o+8) = 0.25; L_B(o) = L_B(o+4) = 0.46875,
    L_B(o+8) = 0.78125.  The best single match is A at o+4.  Over the
    recheck range [o+4, o+20) the pair (A at o, B at o+8) reproduces the
    voltage exactly: score (1 + 1.5625)/2 = 1.28125 > 1.25, unique maximum.
    Parse: L_main = L_A(o) = 1 >= L_other = L_B(o+8) = 0.78125 and 1 > 0.5,
    so the window's spike becomes A at o (binary_pursuit.cl:752-759).

    Patterns at o = 46 (best A at 50, window 3) and o = 78 (best A at 82,
    window 5).  Pass 1, check_overlap_reassignments: window 3 moves to 46 < 48
    and sets next_check[1] = 2 (overlap_recheck[1] = 0); window 5 moves to
    78 < 80 with overlap_recheck[3] == 1, so it gives up: best index 0, flag
    cleared, next_check[4..6] = 2.  binary_pursuit adds only A at 46 (window 3
    via its overlap flag; window 5 has index 0 outside its window and no
    flag) and sets next_check[2..4] = 3.  No flag is pruned (flags 2 and 3,
    every best above threshold): pass 2 evaluates windows 1..6.

    Pass 2 (B at 54 and pattern 78 left): window 3's scan [32, 80) now sees
    A at 78 (1) above B at 54 (0.78125), so B is blocked; window 5 moves its
    spike to 78 again, overlap_recheck[3] is now 0: next_check[3] = 2 and A at
    78 is added, next_check[4..6] = 3.  Window 3 keeps flag 2 (the rule of
    binary_pursuit.cl:1040 needs flag < 3 and a best below threshold; its best
    A at 78 is 1 > 0.5).  Windows 1, 2 see no positive likelihood (best 0,
    pruned).  Pass 3 evaluates 3..6 and adds B at 54 (window 3) and B at 86
    (window 5); each is alone, the pair recheck cannot beat its single score
    (a pair must exceed it strictly, binary_pursuit.cl:707-711).  Pass 4 adds
    nothing.
    """
    C, K = 1, 16 * 40 + 5
    tm = _delta_templates(C, _AB)
    V = np.zeros((C, K), np.float32)
    truth = []
    for o in (46, 78):
        _place(V, tm, o, 0)
        _place(V, tm, o + 8, 1)
        truth += [(o, 0), (o + 8, 1)]
    passes = [[(46, 0)], [(78, 0)], [(54, 1), (86, 1)], []]
    lists = [[1, 2, 3, 4, 5, 6], [3, 4, 5, 6], [2, 3, 4, 5, 6]]
    return _finish("give_up_left", V, tm, _AB_THR, passes, lists, truth)


def _right_pattern(V, tm, o):
    # B at o and A at o + 4: V = 1.25 at o, 1 at o+4, 1 at o+8.
    _place(V, tm, o, 1)
    _place(V, tm, o + 4, 0)


def move_right():
    """A spike moved right past the end of its window flags window id + 2
    (binary_pursuit.cl:867-871).

    Pattern: B at o, A at o + 4.  L_A(o-4) = 0.25, L_A(o) = 1.25,
    L_A(o+4) = 1; L_B(o) = 0.78125, L_B(o+4) = L_B(o+8) = 0.46875.  The best
    single match is A at o; over [o, o+16) the pair (A at o+4, B at o)
    reproduces the voltage: score (1.5625 + 2)/2 = 1.78125 > 1.25; parse
    keeps the main label at o+4 (L_A(o+4) = 1 >= 0.78125, 1 > 0.5).

    o = 94: window 5 moves its spike to 98 >= 96, so next_check[7] = 2
    (7 < K/N).  Window 7 sees no positive likelihood in its own window (its
    argmax flag is 0), so only that rule puts it on the list; its best is A
    at 98 (1 > 0.5) and the flag 2 survives the pruning.  binary_pursuit adds
    A at 98, next_check[4..6] = 3: pass 2 evaluates 4..7 and adds B at 94;
    pass 3 evaluates 4..6 and adds nothing.
    """
    C, K = 1, 16 * 40 + 5
    tm = _delta_templates(C, _AB)
    V = np.zeros((C, K), np.float32)
    _right_pattern(V, tm, 94)
    passes = [[(98, 0)], [(94, 1)], []]
    lists = [[4, 5, 6, 7], [4, 5, 6]]
    return _finish("move_right", V, tm, _AB_THR, passes, lists, [(94, 1), (98, 0)])


def move_right_last_window():
    """The same move next to the end of the chunk: with K = 122 (K/N = 7 by
    integer division, 8 windows) the spike of window 5 moves to 98, and
    id + 2 = 7 is not < K/N, so no flag is written (binary_pursuit.cl:867);
    window 7 exists, so a bound of ceil(K/N) would list it.  Windows 6 and 7
    end at or after K - N and are never scored.  Pass 1 adds A at 98 and pass
    2 evaluates 4..6 and adds B at 94; pass 3 adds nothing."""
    C, K = 1, 122
    tm = _delta_templates(C, _AB)
    V = np.zeros((C, K), np.float32)
    _right_pattern(V, tm, 94)
    passes = [[(98, 0)], [(94, 1)], []]
    lists = [[4, 5, 6], [4, 5, 6]]
    return _finish("move_right_last_window", V, tm, _AB_THR, passes, lists, [(94, 1), (98, 0)])


def pruning():
    """Flag pruning of binary_pursuit.cl:1009-1045.

    Four one-delta templates on their own channels (no cross terms):
      P (0): 2 on channel 0, own L = 2, threshold 0
      Y (1): 1.5 on channel 1, own L = 1.125, threshold 1.125 (= its L)
      X (2): 1.25 on channel 2, own L = 0.78125, threshold 0.5
      Q (3): 1.75 on channel 3, own L = 1.53125, threshold 0
    Events: Y at 70 (window 4), X at 90 (window 5), 0.75 X at 100 (window
    6, L_X = 0.390625), P at 116 (window 7), Q at 140 (window 8).

    Pass 1: window 4's best is Y at 70, exactly at threshold: not added
    (strict >, binary_pursuit.cl:982) and its flag 1 is pruned (<=,
    :1040-1044); window 5's best is Y at 70 too (pruned the same way).
    Window 7 adds P at 116 (next_check[6..8] = 3).  Window 8's best is P at
    116, so Q at 140 waits.  Pass 2 evaluates 6, 7, 8.

    Pass 2: window 6's scan reads the likelihood of window 5 as it stood in
    pass 1 (not recomputed): its best is X at 90 (0.78125 > 0.5, left of its
    window) and its own flag is 1 (0.75 X at 100); window 5 is not on the
    list, so its flag is 0 and window 6's flag is pruned (:1024-1031).
    Window 8 adds Q at 140 (next_check[7..9] = 3).  Pass 3 evaluates 7, 8, 9
    and adds nothing: Y, X and 0.75 X are never added.
    """
    C, K = 4, 16 * 40 + 5
    tm = _delta_templates(C, [[(0, 0, 2.0)], [(1, 0, 1.5)], [(2, 0, 1.25)], [(3, 0, 1.75)]])
    V = np.zeros((C, K), np.float32)
    _place(V, tm, 70, 1)
    _place(V, tm, 90, 2)
    _place(V, tm, 100, 2, scale=0.75)
    _place(V, tm, 116, 0)
    _place(V, tm, 140, 3)
    passes = [[(116, 0)], [(140, 3)], []]
    lists = [[6, 7, 8], [7, 8, 9]]
    return _finish("pruning", V, tm, [0.0, 1.125, 0.5, 0.0], passes, lists, [(116, 0), (140, 3)],
                   residual_zero=False)


SCENARIOS = [give_up_left, move_right, move_right_last_window, pruning]


def chained_train(seed, n_units=4, C=4, n_spikes=24):
    """A dense train of sparse dyadic templates (three positive deltas each,
    on different channels and samples) in zero noise: spikes come in pairs
    5 samples apart (inside the shift range, so each pair is one overlap
    recheck) and pairs are 22..31 samples apart, so the scans of neighbouring
    windows chain through every pair.  Positive deltas keep every true spike's
    own likelihood at or above half its squared norm, above the 0.25 |t|^2
    threshold.  Returns the recording and the true (onset, label) list."""
    rng = np.random.default_rng(seed)
    spec = []
    for u in range(n_units):
        deltas = []
        for c in rng.choice(C, size=3, replace=False):
            deltas.append((int(c), int(rng.integers(2, N - 2)), float(rng.choice([1.0, 1.25, 1.5, 1.75, 2.0]))))
        spec.append(deltas)
    tm = _delta_templates(C, spec)
    K = 16 * (n_spikes * 2 + 12) + 5
    V = np.zeros((C, K), np.float32)
    truth = []
    o = 3 * N
    for k in range(n_spikes):
        lab = int(rng.integers(0, n_units))
        _place(V, tm, o, lab)
        truth.append((o, lab))
        o += int(rng.integers(5, M + 2)) if k % 2 == 0 else int(rng.integers(N + 6, 2 * N))
    ss = np.sum(tm.astype(np.float64).reshape(n_units, -1) ** 2, axis=1)
    thr = 0.25 * ss
    rec = _finish("chained_%d" % seed, V, tm, thr, None, None, truth)
    return rec
