"""Size-independent checks of banded local alignments (scores, ends, starts and
BAM CIGAR runs), run on the device over every alignment of a batch or a served
stream:

* each CIGAR consumes exactly the query span [read_start, query_end) and the
  target span [ref_start, target_end) it reports, and its op counts match;
* every '=' / 'X' column agrees with the actual bases;
* rescoring each CIGAR with the affine scheme (a gap of length L costs
  open + (L-1) * extend, extension.py:154-156) gives exactly the reported
  score, and a positive-score local alignment begins and ends with '=';
  a score of 0 has no CIGAR.

Used by the GPU tests only (needs a device for the vectorised checks)."""

import torch

OP_EQ, OP_X, OP_I, OP_D = 7, 8, 1, 2


def check_alignments(label, *, q_codes, q_base, t_codes, t_base, score, qend, tend, rs, ts,
                     n_cigar, cigar_off, cigar, counts, scoring):
    """Every argument is a tensor on one device.  q_codes / t_codes: uint8
    base codes; q_base / t_base (int64, per alignment): where the alignment's
    query / target starts in those arrays (ends and starts count from there);
    cigar: int32 or int64 runs (len << 4 | op) at cigar_off[k], n_cigar[k] of
    them; counts: (n, 4) of =, X, I, D.  Returns the number of aligned
    columns compared against the bases."""
    dev = score.device
    n = int(score.numel())
    score, qe, te = score.long(), qend.long(), tend.long()
    rs, ts, nc = rs.long(), ts.long(), n_cigar.long()
    assert (score >= 0).all() and (nc >= 0).all(), label
    total = int(nc.sum())
    pair = torch.arange(n, device=dev)
    owner = torch.repeat_interleave(pair, nc)
    first = torch.cumsum(nc, 0) - nc
    within = torch.arange(total, device=dev) - first[owner]
    runs = cigar[cigar_off.long()[owner] + within].long() & 0xFFFFFFFF
    op, ln = runs & 15, runs >> 4
    assert (ln >= 1).all(), label
    assert torch.isin(op, torch.tensor([OP_I, OP_D, OP_EQ, OP_X], device=dev)).all(), label
    is_eq, is_x, is_i, is_d = op == OP_EQ, op == OP_X, op == OP_I, op == OP_D
    zero = torch.zeros(n, dtype=torch.long, device=dev)

    def per_pair(v):
        return zero.clone().index_add_(0, owner, v)

    q_cons = torch.where(is_eq | is_x | is_i, ln, 0)
    t_cons = torch.where(is_eq | is_x | is_d, ln, 0)
    assert torch.equal(per_pair(q_cons), qe - rs), (label, "query span")
    assert torch.equal(per_pair(t_cons), te - ts), (label, "target span")
    cnt = counts.long()
    for k, m in enumerate((is_eq, is_x, is_i, is_d)):
        assert torch.equal(per_pair(torch.where(m, ln, 0)), cnt[:, k]), (label, "counts", k)
    ma, mi, go, ge = (int(x) for x in scoring)
    contrib = torch.where(is_eq, ma * ln, torch.where(is_x, -mi * ln, -(go + (ln - 1) * ge)))
    assert torch.equal(per_pair(contrib), score), (label, "rescoring")
    pos = score > 0
    last = first + nc - 1
    assert (op[first[pos]] == OP_EQ).all() and (op[last[pos]] == OP_EQ).all(), label
    assert (nc[~pos] == 0).all(), label
    # '=' / 'X' columns against the bases
    gq = torch.cumsum(q_cons, 0) - q_cons
    gt = torch.cumsum(t_cons, 0) - t_cons
    q_run = gq - gq[first[owner]]  # query offset of each run inside its alignment
    t_run = gt - gt[first[owner]]
    diag = torch.nonzero(is_eq | is_x).squeeze(1)
    dl = ln[diag]
    col_run = torch.repeat_interleave(diag, dl)
    col_k = torch.arange(int(dl.sum()), device=dev) - \
        (torch.cumsum(dl, 0) - dl).repeat_interleave(dl)
    p = owner[col_run]
    qi = q_base.long()[p] + rs[p] + q_run[col_run] + col_k
    ti = t_base.long()[p] + ts[p] + t_run[col_run] + col_k
    same = q_codes[qi] == t_codes[ti]
    assert torch.equal(same, op[col_run] == OP_EQ), (label, "= / X vs bases")
    return int(dl.sum())
