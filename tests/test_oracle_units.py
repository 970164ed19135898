"""The oracle's step primitives against the reference's own unit literals.

Each test restates one unit test of the reference on the oracle's exported
test hooks (oracle/wfa_oracle.h, ``wfo_t_*``): the same cells in, the same
decisions out. CPU only.
"""
import ctypes

import numpy as np
import pytest

OFFSET_NULL = -(2 ** 31) // 2
DIAG_NULL = 2 ** 31 - 1
I32 = ctypes.c_int32
P32 = ctypes.POINTER(ctypes.c_int32)

WFADAPTIVE, XDROP, ZDROP, BANDED_STATIC, BANDED_ADAPTIVE = 1, 2, 3, 4, 5
C_M, C_I1, C_D1 = 0, 1, 2


@pytest.fixture(scope="module")
def L(oracle, N):
    return oracle.lib(N.Attr)


def arr(values):
    a = np.asarray(values, dtype=np.int32)
    return a, a.ctypes.data_as(P32)


class Row:
    """n cells over the diagonals [base, base + n - 1], all zero (as
    Wavefront::allocate leaves them) unless given."""

    def __init__(self, base, n, fill=0):
        self.base, self.n = base, n
        self.a = np.full(n, fill, dtype=np.int32)

    def __setitem__(self, k, v):
        self.a[k - self.base] = v

    def __getitem__(self, k):
        return int(self.a[k - self.base])

    @property
    def p(self):
        return self.a.ctypes.data_as(P32)


def state(heuristic=0, hp=(0, 0, 0)):
    """HeuristicState::default, or reset for a strategy."""
    st, pst = arr([0, 0, DIAG_NULL, OFFSET_NULL, 0, 0])
    return st, pst


# ---- src/heuristic.rs unit tests -----------------------------------------

def test_wf_distance_end2end(L):
    # src/heuristic.rs:545-568
    assert L.wfo_t_wf_distance(5, 0, 10, 10) == 5
    assert L.wfo_t_wf_distance(8, 2, 10, 10) == 4
    assert L.wfo_t_wf_distance(OFFSET_NULL, 0, 10, 10) == 2 ** 31 - 1
    assert L.wfo_t_wf_distance(10, 0, 10, 10) == 0


def test_wfadaptive_reduce_basic(L):
    # src/heuristic.rs:570-598: wf over [-5, 5] with limits [-3, 3]
    d = Row(-5, 11)
    for k, v in ((-3, 100), (-2, 5), (-1, 2), (0, 2), (1, 2), (2, 5), (3, 100)):
        d[k] = v
    lohi, plohi = arr([-3, 3])
    L.wfo_t_wfadaptive_reduce(plohi, d.p, -5, 2, 10, 0, 0)
    assert list(lohi) == [-2, 2]


def test_wfadaptive_too_short(L):
    # src/heuristic.rs:600-625: length 3 < min_wavefront_length 5
    wf = Row(-5, 11)
    d = Row(-5, 11)
    st, pst = state()
    lohi, plohi = arr([-1, 1])
    L.wfo_t_wfadaptive_cutoff(pst, wf.p, -5, 11, plohi, d.p, 10, 10, 5, 10, 1)
    assert list(lohi) == [-1, 1]
    assert st[0] == 0  # the wait is not reset


def test_banded_static(L):
    # src/heuristic.rs:627-635 and :705-713
    lohi, plohi = arr([-3, 3])
    L.wfo_t_banded_static(plohi, -2, 2)
    assert list(lohi) == [-2, 2]
    lohi, plohi = arr([-1, 1])
    L.wfo_t_banded_static(plohi, -3, 3)
    assert list(lohi) == [-1, 1]


def test_compute_sw_scores_c_formula(L):
    # src/heuristic.rs:637-664
    wf = Row(-5, 11)
    wf[-1], wf[0], wf[1] = 4, 6, OFFSET_NULL
    sw = Row(-5, 11)
    out, pout = arr([0, 0, 0])
    drop, pdrop = arr([-3, 0, 0, 0, 2])
    L.wfo_t_compute_sw_scores(wf.p, -5, 11, -1, 1, sw.p, pdrop, 10, pout)
    assert sw[-1] == (3 * 9 - 10) // 2
    assert tuple(out) == ((3 * 12 - 10) // 2, 0, 6)
    # match == 0: +1 per base, so the furthest cell is the best
    drop, pdrop = arr([0, 0, 0, 0, 2])
    L.wfo_t_compute_sw_scores(wf.p, -5, 11, -1, 1, sw.p, pdrop, 10, pout)
    assert (out[0], out[1]) == ((12 - 10) // 2, 0)
    # ends-free with match == 0: the free leading gap on k = -1 is no progress
    drop, pdrop = arr([0, 1, 3, 3, 2])
    L.wfo_t_compute_sw_scores(wf.p, -5, 11, -1, 1, sw.p, pdrop, 10, pout)
    assert sw[-1] == int((8 - 10) / 2)  # truncation toward zero, as i32 division
    # with match != 0 the begin-free lengths do not apply: same as no ends-free
    drop, pdrop = arr([-3, 1, 3, 3, 2])
    L.wfo_t_compute_sw_scores(wf.p, -5, 11, -1, 1, sw.p, pdrop, 10, pout)
    assert sw[-1] == (3 * 9 - 10) // 2


def test_zdrop_must_persist(L):
    # src/heuristic.rs:666-690: one cell on k = 0, match -1, max_penalty 4
    st, pst = state()
    drop, pdrop = arr([-1, 0, 0, 0, 5])
    sw = Row(-1, 3)

    def step(off, score):
        wf = Row(-1, 3)
        wf[0] = off
        lohi, plohi = arr([0, 0])
        st[0] = 0  # check every call
        return L.wfo_t_zdrop_cutoff(pst, wf.p, -1, 3, plohi, sw.p, pdrop, score, 10, 1)

    assert not step(50, 0)   # best: sw 50
    assert not step(50, 18)  # sw 41: drop 9, no Z-drop
    assert not step(50, 21)  # sw 39: drop 11, for 3 scores
    assert step(50, 22)      # for 4 scores


def test_drop_state_reset_uses_steps(L):
    # src/heuristic.rs:692-703
    st, pst = state()
    hp, php = arr([20, 7, 0])
    L.wfo_t_hstate_reset(pst, XDROP, php)
    assert st[0] == 7 and st[2] == DIAG_NULL
    hp, php = arr([20, 3, 0])
    L.wfo_t_hstate_reset(pst, ZDROP, php)
    assert st[0] == 3


def test_heuristic_equate(L):
    # src/heuristic.rs:715-743
    dst, pdst = arr([-4, 4, 0])
    src, psrc = arr([-2, 2])
    L.wfo_t_heuristic_equate(pdst, psrc)
    assert list(dst) == [-2, 2, 0]
    dst, pdst = arr([-1, 1, 0])
    src, psrc = arr([2, 1])
    L.wfo_t_heuristic_equate(pdst, psrc)
    assert dst[2] == 1


def test_xdrop_first_call_records_best_then_trims(L):
    """xdrop_cutoff (src/heuristic.rs:410-448): the first call only records
    the best; later calls trim cells at least `xdrop` below it from both
    ends (skipping NULL cells) and raise the best. Cells worked by hand."""
    st, pst = state()
    drop, pdrop = arr([-1, 0, 0, 0, 3])
    wf = Row(-3, 7)
    # sw = (v + h - score) / 2 with v = off - k, h = off
    wf[-2], wf[-1], wf[0], wf[1], wf[2] = 2, OFFSET_NULL, 10, 9, 1
    sw = Row(-3, 7)
    lohi, plohi = arr([-2, 2])
    L.wfo_t_xdrop_cutoff(pst, wf.p, -3, 7, plohi, sw.p, pdrop, 0, 5, 1)
    assert list(lohi) == [-2, 2]
    assert (st[1], st[2]) == (10, 0)  # sw(0) = (10 + 10) / 2
    st[0] = 0
    # score 4: sw(-2) = (4 + 2 - 4) / 2 = 1, sw(0) = 8, sw(1) = (8 + 9 - 4) / 2 = 6,
    # sw(2) = (-1 + 1 - 4) / 2 = -2: -2 and 2 drop >= 5 below the best 10
    L.wfo_t_xdrop_cutoff(pst, wf.p, -3, 7, plohi, sw.p, pdrop, 4, 5, 1)
    assert list(lohi) == [0, 1]
    assert (st[1], st[2]) == (10, 0)


def test_banded_adaptive_shifts_window(L):
    """banded_adaptive_cutoff (src/heuristic.rs:355-394): a window of the
    band's width, shifted by the leeway toward the end whose sampled
    diagonals are closer to the end of the alignment. Worked by hand."""
    st, pst = state()
    plen = tlen = 100
    wf = Row(-10, 21)
    lo, hi = -8, 7  # 16 diagonals, band [-2, 2] of 5: leeway 5, quarter 4
    for k in range(lo, hi + 1):
        wf[k] = 60 + k  # h = 60 + k, v = 60: higher k is closer to the end
    lohi, plohi = arr([lo, hi])
    L.wfo_t_banded_adaptive(pst, wf.p, -10, 21, plohi, plen, tlen, -2, 2, 1)
    # d(lo) > d(hi) and d(lo + 4) > d(lo + 8): shifted twice
    assert list(lohi) == [lo + 10, lo + 10 + 4]
    assert st[0] == 1
    # shorter than 4: untouched, and the wait is not reset
    st[0] = 0
    lohi, plohi = arr([0, 2])
    L.wfo_t_banded_adaptive(pst, wf.p, -10, 21, plohi, plen, tlen, -2, 2, 9)
    assert list(lohi) == [0, 2] and st[0] == 0


# ---- kernel cells: src/compute/affine.rs, src/compute/edit.rs -------------

def test_affine_kernel_gap_open(L):
    # src/compute/affine.rs:364-416
    n, base = 11, -5
    m_open = Row(base, n)
    m_open[0] = 0
    for k in (-2, -1, 1, 2):
        m_open[k] = OFFSET_NULL
    m_misms, i1_ext, d1_ext = (Row(base, n, OFFSET_NULL) for _ in range(3))
    m, i1, d1 = Row(base, n), Row(base, n), Row(base, n)
    rows = (ctypes.POINTER(I32) * 12)(m_misms.p, m_open.p, i1_ext.p, d1_ext.p, None, None, None,
                                      m.p, i1.p, d1.p, None, None)
    L.wfo_t_affine_cells(4, 4, rows, base, n, -1, 1, 0)
    assert d1[-1] == 0 and m[-1] == 0
    assert i1[1] == 1 and m[1] == 1


def _edit_prev():
    prev = Row(-3, 7)
    prev[0] = 0
    for k in (-2, -1, 1, 2):
        prev[k] = OFFSET_NULL
    return prev


def test_edit_kernel_basic(L):
    # src/compute/edit.rs:240-278
    prev, cur = _edit_prev(), Row(-3, 7)
    L.wfo_t_edit_cells(4, 4, prev.p, cur.p, -3, 7, -1, 1, 0)
    assert (cur[-1], cur[0], cur[1]) == (0, 1, 1)


def test_indel_kernel_basic(L):
    # src/compute/edit.rs:280-306
    prev, cur = _edit_prev(), Row(-3, 7)
    L.wfo_t_edit_cells(4, 4, prev.p, cur.p, -3, 7, -1, 1, 1)
    assert (cur[-1], cur[0], cur[1]) == (0, OFFSET_NULL, 1)


# ---- BiWFA breakpoints: src/bialign.rs unit tests ---------------------------

def scan(L, w0, lo0, hi0, w1, lo1, hi1, plen, tlen, s0, s1, gap_open=0, comp=C_M, fwd=1):
    bp, pbp = arr([2 ** 31 - 1, 0, 0, 0, 0, 0, 0, C_M])
    L.wfo_t_breakpoint_scan(w0.p, w0.base, w0.n, lo0, hi0, w1.p, w1.base, w1.n, lo1, hi1,
                            plen, tlen, s0, s1, gap_open, comp, fwd, pbp)
    return bp


def test_breakpoint_m2m_found(L):
    # src/bialign.rs:535-559
    f, r = Row(-3, 7), Row(-3, 7)
    f[0], r[0] = 6, 5
    bp = scan(L, f, 0, 0, r, 0, 0, 10, 10, 3, 2)
    assert list(bp) == [5, 3, 2, 0, 0, 6, 5, C_M]


def test_breakpoint_m2m_not_found(L):
    # src/bialign.rs:561-576
    f, r = Row(-3, 7), Row(-3, 7)
    f[0], r[0] = 4, 4
    assert scan(L, f, 0, 0, r, 0, 0, 10, 10, 3, 2)[0] == 2 ** 31 - 1


def test_breakpoint_m2m_different_lengths(L):
    # src/bialign.rs:578-600
    f, r = Row(-5, 11), Row(-5, 11)
    f[2], r[2] = 9, 4
    bp = scan(L, f, 2, 2, r, 2, 2, 8, 12, 5, 3)
    assert bp[0] == 8


def test_breakpoint_m2m_no_k_overlap(L):
    # src/bialign.rs:602-627
    f, r = Row(-5, 11), Row(-5, 11)
    f[4], f[5], r[-5], r[-4] = 1, 1, 1, 1
    assert scan(L, f, 4, 5, r, -5, -4, 10, 10, 3, 2)[0] == 2 ** 31 - 1


def test_breakpoint_indel2indel_found(L):
    # src/bialign.rs:629-659
    f, r = Row(-3, 7), Row(-3, 7)
    f[0], r[0] = 6, 5
    bp = scan(L, f, 0, 0, r, 0, 0, 10, 10, 3, 2, gap_open=6, comp=C_D1)
    assert bp[0] == -1 and bp[7] == C_D1


def test_scan_overlap_first_hit(L):
    """The first hit of scan_overlap is the lowest diagonal of wf_0 whose
    offsets reach the text length (src/bialign.rs:667-740: the reference's
    naive ascending scan, same xorshift generator and offsets)."""
    seed = [0x2545F491]

    def nxt():
        s = seed[0]
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        seed[0] = s
        return s

    plen, tlen, n, c0, c1 = 90, 100, 80, 40, 40
    for trial in range(400):
        def offsets():
            out = []
            for _ in range(n):
                r = nxt()
                out.append(OFFSET_NULL if r % 7 == 0 else 30 + (r % 25) + (trial % 3) * 3)
            return out
        a0, a1 = offsets(), offsets()
        lo = -12 + nxt() % 8
        hi = lo + nxt() % 26
        hits = [k for k in range(lo, hi + 1)
                if a0[c0 + k] + a1[c1 + (tlen - plen) - k] >= tlen]
        w0, w1 = Row(-c0, n), Row(-c1, n)
        w0.a[:] = a0
        w1.a[:] = a1
        # w1's range: every diagonal whose inverse lies in [lo, hi]
        bp = scan(L, w0, lo, hi, w1, (tlen - plen) - hi, (tlen - plen) - lo, plen, tlen, 1, 1)
        if hits:
            assert bp[3] == hits[0], (trial, lo, hi)
            assert (bp[5], bp[6]) == (a0[c0 + hits[0]], a1[c1 + (tlen - plen) - hits[0]])
        else:
            assert bp[0] == 2 ** 31 - 1, (trial, lo, hi)
