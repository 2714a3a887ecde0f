"""float32 transcription of the reference's banded SW row loop and traceback.

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

This restates extension.py:78-302 in numpy float32 with the reference's own
sentinel (-1e30) and its closed-form E (an exclusive max-plus cummax over
``M + d*extend``, extension.py:14-33, 161-167).  Its only job is to pin the
int32 sentinel mapping of ``gmfa_oracle.c`` (ONEG with saturating add): on
the same inputs both must give identical scores, ends, matrices (with -1e30
<-> ONEG) and CIGARs.  Pure numpy per row; keep inputs small.
"""

from __future__ import annotations

import numpy as np

NEG32 = np.float32(-1e30)


def sw_rows_f32(q, t, qlen, tlen, off, hb, ma, mi, go, ge, xdrop=0.0):
    """One call = one batch (B pairs, batch-wide x-drop)."""
    q = np.asarray(q, np.int16)
    t = np.asarray(t, np.int16)
    B, M = q.shape
    N = t.shape[1]
    W = 2 * hb + 1
    d_idx = np.arange(W)
    f32 = np.float32
    scan_add = d_idx.astype(f32) * f32(ge)
    scan_sub = f32(go) + (d_idx.astype(f32) - f32(1.0)) * f32(ge)
    H = np.zeros((B, M + 1, W), f32)
    E = np.full((B, M + 1, W), NEG32, f32)
    F = np.full((B, M + 1, W), NEG32, f32)
    hprev = np.zeros((B, W + 1), f32)
    fprev = np.full((B, W + 1), NEG32, f32)
    best = np.zeros(B, f32)
    bi = np.zeros(B, np.int64)
    bj = np.zeros(B, np.int64)
    seen = np.zeros(B, f32)
    qlen = np.asarray(qlen, np.int64)
    tlen = np.asarray(tlen, np.int64)
    off = np.asarray(off, np.int64)
    rows = int(qlen.max()) if B else 0
    for i in range(1, rows + 1):
        j = i + off[:, None] - hb + d_idx[None, :]
        inb = (j >= 1) & (j <= tlen[:, None]) & (i <= qlen[:, None])
        qb = q[:, i - 1][:, None]
        tb = np.take_along_axis(t, np.clip(j - 1, 0, max(N - 1, 0)), axis=1)
        match = (qb == tb) & (qb >= 0) & (qb != 5)
        sub = np.where(match, f32(ma), f32(-mi)).astype(f32)
        diag = hprev[:, :W] + sub
        vert = np.maximum(hprev[:, 1:] - f32(go), fprev[:, 1:] - f32(ge))
        m = np.maximum(np.maximum(diag, vert), f32(0.0))
        m = np.where(inb, m, NEG32).astype(f32)
        incl = np.maximum.accumulate(m + scan_add, axis=1)
        excl = np.concatenate([np.full((B, 1), NEG32, f32), incl[:, :-1]], axis=1)
        e = (excl - scan_sub).astype(f32)
        e = np.where(inb, e, NEG32).astype(f32)
        h = np.maximum(np.maximum(m, e), f32(0.0))
        h = np.where(inb, h, f32(0.0)).astype(f32)
        fv = np.where(inb, vert, NEG32).astype(f32)
        H[:, i] = h
        E[:, i] = e
        F[:, i] = fv
        rb = h.max(axis=1)
        slot = h.argmax(axis=1)
        imp = rb > best
        best = np.where(imp, rb, best)
        bi = np.where(imp, i, bi)
        bj = np.where(imp, j[np.arange(B), slot], bj)
        seen = np.maximum(seen, rb)
        hprev = np.concatenate([h, np.zeros((B, 1), f32)], axis=1)
        fprev = np.concatenate([fv, np.full((B, 1), NEG32, f32)], axis=1)
        if xdrop > 0 and bool(((seen - rb) > xdrop).all()):
            break
    return dict(score=best, query_end=bi, target_end=bj, H=H, E=E, F=F, half_band=hb,
                band_offset=off)


def traceback_f32(res, index, q, t, ma, mi, go, ge, tol=1e-3):
    H, E, F = res["H"][index], res["E"][index], res["F"][index]
    hb = res["half_band"]
    off = int(res["band_offset"][index])
    rows, W = H.shape

    def cell(X, i, j):
        d = j - i - off + hb
        if i < 0 or i >= rows or d < 0 or d >= W:
            return -1e30
        return float(X[i, d])

    i = int(res["query_end"][index])
    j = int(res["target_end"][index])
    ops = []
    state = "H"
    while i > 0 and j > 0:
        if state == "H":
            h = cell(H, i, j)
            if h <= 0.0:
                break
            mt = q[i - 1] == t[j - 1] and q[i - 1] >= 0 and q[i - 1] != 5
            sub = ma if mt else -mi
            if abs(h - (max(cell(H, i - 1, j - 1), 0.0) + sub)) <= tol:
                ops.append("=" if mt else "X")
                i -= 1
                j -= 1
            elif abs(h - cell(F, i, j)) <= tol:
                state = "F"
            else:
                state = "E"
        elif state == "F":
            cur = cell(F, i, j)
            ops.append("I")
            i -= 1
            if abs(cur - (cell(H, i, j) - go)) <= tol:
                state = "H"
            elif abs(cur - (cell(F, i, j) - ge)) > tol:
                state = "H"
        else:
            cur = cell(E, i, j)
            ops.append("D")
            j -= 1
            if abs(cur - (cell(H, i, j) - go)) <= tol:
                state = "H"
            elif abs(cur - (cell(E, i, j) - ge)) > tol:
                state = "H"
    ops.reverse()
    cig = []
    for op in ops:
        if cig and cig[-1][0] == op:
            cig[-1] = (op, cig[-1][1] + 1)
        else:
            cig.append((op, 1))
    return cig, i, j
