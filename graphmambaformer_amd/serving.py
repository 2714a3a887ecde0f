"""Continuous serving of banded SW + CIGAR requests on one device.

The reference aligns reads one after another: `pipeline.py:355-363` fans
reads over host threads, each calling `ExtensionEngine.extend_chains`
(`extension.py:511-641`) with its own small batch.  Here a request stream is
served by `gmfa_serve_run`: one persistent forward kernel per band class, a
wavefront per request, the CIGAR walked in the same launch, requests published
by the host as they arrive (continuous batching; no batch boundaries, no launch
per request).  Targets are windows of a reference kept resident on the device
as a 2-bit stream; a request names its window by base offset.

Memory per class: one traceback-plane slab per wavefront slot (sized for the
class's longest query), per-request outputs (score, ends, CIGAR runs) and the
request table.  The library allocates only a small host-coherent control block
for the duration of a run.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib, kernels
from ._lib import ServeClass, ServeReport, SwTile, check, lib, ptr


def serve_tile(half_band: int) -> SwTile:
    """The compiled tile with one pair per wavefront (64 lanes) and the least
    padding for this band (for W = 64*C + 1, C cells per lane plus the scalar
    cell)."""
    t = kernels.resolve_tile(half_band)
    if t.lanes_per_pair == 64:
        return t
    t = kernels.narrow_tile(half_band)
    if t is None:
        raise RuntimeError(f"no 64-lane tile holds a band of width {2 * int(half_band) + 1}")
    return t


@dataclass
class PageGeometry:
    """Paged traceback planes of one class: pages of `page_blocks` 8-row
    plane blocks (a power of two) from a pool of `n_pages`."""

    page_blocks: int
    n_pages: int
    max_pages: int
    page_words: int

    @property
    def pool_bytes(self) -> int:
        return self.n_pages * self.page_words * 4


def page_geometry(tile: SwTile, max_qlen: int, budget_bytes: int,
                  page_blocks: Optional[int] = None) -> PageGeometry:
    """Pages for a plane budget: up to about 1 MB each (the longest query's
    planes in one page where they fit), the longest query within 256 pages."""
    blkw = int(tile.lanes_stored) * int(tile.cells_per_lane)  # words per 8-row block
    max_blocks = max((int(max_qlen) + 7) // 8, 1)
    if page_blocks is None:
        pb = 1
        while pb * 2 * blkw * 4 <= (1 << 20) and pb < max_blocks:
            pb *= 2
        while -(-max_blocks // pb) > 256:
            pb *= 2
    else:
        pb = int(page_blocks)
    page_words = pb * blkw
    max_pages = -(-max_blocks // pb)
    if max_pages > 256:
        raise ValueError(f"the longest query needs {max_pages} pages of {pb} blocks; at most 256")
    n_pages = int(budget_bytes) // (page_words * 4)
    if n_pages < max_pages:
        raise ValueError(f"a plane budget of {budget_bytes} bytes holds {n_pages} pages; the "
                         f"longest query needs {max_pages}")
    return PageGeometry(page_blocks=pb, n_pages=n_pages, max_pages=max_pages,
                        page_words=page_words)


@dataclass
class ResidentReference:
    """A reference held on the device as one 2-bit stream (16 bases per word)."""

    words: torch.Tensor
    n_bases: int

    @classmethod
    def upload(cls, codes: np.ndarray, device) -> "ResidentReference":
        """codes: 0..3 (A, C, G, T)."""
        w, _, _, _ = kernels.pack_host([np.asarray(codes, np.uint8)], alphabet="2bit")
        t = torch.from_numpy(w.view(np.int32)).to(torch.device(device))
        return cls(words=t, n_bases=int(len(codes)))


@dataclass
class RequestPool:
    """Queries of one class packed on the device, with their reference windows."""

    half_band: int
    qseq: torch.Tensor
    q_off: np.ndarray   # base offset of each query in qseq
    qlen: np.ndarray
    t_start: np.ndarray  # base offset of each window in the reference
    t_len: np.ndarray
    band_offset: np.ndarray

    @classmethod
    def upload(cls, queries: Sequence[np.ndarray], t_start, t_len, band_offset, half_band: int,
               device) -> "RequestPool":
        w, _, qo, _ = kernels.pack_host([np.asarray(q, np.uint8) for q in queries], "2bit")
        return cls(half_band=int(half_band),
                   qseq=torch.from_numpy(w.view(np.int32)).to(torch.device(device)),
                   q_off=qo.astype(np.int64),
                   qlen=np.array([len(q) for q in queries], np.int32),
                   t_start=np.asarray(t_start, np.int64), t_len=np.asarray(t_len, np.int32),
                   band_offset=np.asarray(band_offset, np.int32))


@dataclass
class ServeClassRun:
    """One band class of a serving run: its request table (pool entries in
    arrival order), outputs, slabs and arrival times."""

    name: str
    inp: kernels.SwInputs
    scoring: tuple
    half_band: int
    tile: SwTile
    n_slots: int
    slab_words: int
    arrival_s: np.ndarray
    complete_s: np.ndarray
    score: torch.Tensor
    qend: torch.Tensor
    tend: torch.Tensor
    tb: torch.Tensor
    counter: torch.Tensor
    walk: kernels.SwTraceback
    picks: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    pages: Optional[PageGeometry] = None
    page_owner: Optional[torch.Tensor] = None
    page_wait_s: float = 0.0
    prio_rows: int = 0  # requests with at least this many query rows run at raised priority
    split_slices: int = 0  # > 0: every request runs as this many band slices (workgroup slots)

    @property
    def n(self) -> int:
        return self.inp.n_pairs

    def struct(self) -> ServeClass:
        s = ServeClass()
        s.batch = self.inp.batch_struct(None)
        ma, mi, go, ge = self.scoring
        flags = _lib.SW_TRACEBACK | _lib.SW_CIGAR
        s.params = kernels.make_params(ma, mi, go, ge, self.half_band, flags, self.tile,
                                       split_rows=1 if self.split_slices else 0,
                                       split_slices=self.split_slices, prio_rows=self.prio_rows)
        o = s.out
        o.score, o.query_end, o.target_end = ptr(self.score), ptr(self.qend), ptr(self.tend)
        o.tb, o.tb_off = ptr(self.tb), None
        o.rowkey, o.rowkey_off = None, None
        o.work_counter = ptr(self.counter)
        w = self.walk
        o.cigar_off, o.cigar, o.n_cigar = ptr(w.cigar_off), ptr(w.cigar), ptr(w.n_cigar)
        o.read_start, o.ref_start, o.counts = ptr(w.read_start), ptr(w.ref_start), ptr(w.counts)
        s.slab_words = int(self.slab_words)
        s.n_slots = int(self.n_slots)
        if self.pages is not None:
            s.page_blocks, s.n_pages = self.pages.page_blocks, self.pages.n_pages
            s.page_owner = ptr(self.page_owner)
            s.page_wait_s = float(self.page_wait_s)
        s.arrival_s = self.arrival_s.ctypes.data
        s.complete_s = self.complete_s.ctypes.data
        return s


def make_class(name: str, reference: ResidentReference, pool: RequestPool, picks: np.ndarray,
               arrival_s: np.ndarray, scoring: tuple, n_slots: int,
               tile: Optional[SwTile] = None,
               plane_budget_bytes: Optional[int] = None,
               page_blocks: Optional[int] = None,
               page_wait_s: float = 0.0, prio_rows: int = 0,
               split_slices: int = 0) -> ServeClassRun:
    """Request table for pool entries `picks` (in arrival order).  With
    `plane_budget_bytes`, the traceback planes come from a pool of pages of
    that size and requests are admitted by their planes' footprint; without,
    every slot holds a slab for the class's longest query.  Requests with at
    least `prio_rows` query rows (> 0) run at raised wave priority.
    `split_slices` > 0 (a band with a compiled split variant, no N masks):
    every request of the class runs as that many band slices on a workgroup
    of as many wavefronts; `n_slots` then counts workgroups."""
    dev = reference.words.device
    picks = np.asarray(picks, np.int64)
    n = len(picks)
    if len(arrival_s) != n:
        raise ValueError("one arrival time per request")
    tile = tile or serve_tile(pool.half_band)
    if tile.lanes_per_pair != 64:
        raise ValueError("serving needs a 64-lane tile (one pair per wavefront)")
    if n and (pool.t_start[picks].min() < 0 or
              (pool.t_start[picks] + pool.t_len[picks]).max() > reference.n_bases):
        raise ValueError("a request window leaves the reference")

    def up(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    ql, tl = pool.qlen[picks], pool.t_len[picks]
    bo = pool.band_offset[picks]
    inp = kernels.SwInputs(qseq=pool.qseq, tseq=reference.words, qmask=None, tmask=None,
                           q_off=up(pool.q_off[picks]), t_off=up(pool.t_start[picks]),
                           qlen=up(ql), tlen=up(tl), band_offset=up(bo), qlen_host=ql,
                           tlen_host=tl, boff_host=bo, has_n=False)
    max_q = int(ql.max()) if n else 0
    slab = max(int(kernels.tb_bytes(np.array([max_q]), tile)[0]) // 4, 1)
    pages = owner = None
    if plane_budget_bytes is not None:
        pages = page_geometry(tile, max_q, plane_budget_bytes, page_blocks)
        owner = torch.zeros(pages.n_pages, dtype=torch.int32, device=dev)
    cap = ql.astype(np.int64) + tl.astype(np.int64) + 1
    coff = np.zeros(n + 1, np.int64)
    np.cumsum(cap, out=coff[1:])
    walk = kernels.SwTraceback(
        cigar=torch.empty(max(int(coff[-1]), 1), dtype=torch.int32, device=dev),
        cigar_off=up(coff[:-1]), n_cigar=torch.zeros(n, dtype=torch.int32, device=dev),
        read_start=torch.zeros(n, dtype=torch.int32, device=dev),
        ref_start=torch.zeros(n, dtype=torch.int32, device=dev),
        counts=torch.zeros((n, 4), dtype=torch.int32, device=dev))
    return ServeClassRun(
        name=name, inp=inp, scoring=tuple(int(x) for x in scoring), half_band=pool.half_band,
        tile=tile, n_slots=int(n_slots), slab_words=slab,
        arrival_s=np.ascontiguousarray(arrival_s, dtype=np.float64),
        complete_s=np.full(n, -1.0), score=torch.zeros(n, dtype=torch.int32, device=dev),
        qend=torch.zeros(n, dtype=torch.int32, device=dev),
        tend=torch.zeros(n, dtype=torch.int32, device=dev),
        tb=torch.empty(pages.n_pages * pages.page_words if pages else int(n_slots) * slab,
                       dtype=torch.int32, device=dev),
        counter=torch.zeros(16, dtype=torch.int32, device=dev), walk=walk, picks=picks,
        pages=pages, page_owner=owner, page_wait_s=page_wait_s, prio_rows=int(prio_rows),
        split_slices=int(split_slices))


@dataclass
class ServeResult:
    wall_s: float
    completed: int
    timed_out: bool
    waves_gave_up: bool


def run(classes: Sequence[ServeClassRun], timeout_s: float = 60.0) -> ServeResult:
    """Serve every class's requests at their arrival times; blocks until all
    completed (or the timeout).  complete_s of each class is filled in."""
    arr = (ServeClass * len(classes))(*[c.struct() for c in classes])
    rep = ServeReport()
    torch.cuda.synchronize()
    rc = lib().gmfa_serve_run(arr, len(classes), float(timeout_s), ctypes.byref(rep))
    check(rc, "gmfa_serve_run")
    return ServeResult(wall_s=rep.wall_s, completed=int(rep.completed),
                       timed_out=bool(rep.timed_out), waves_gave_up=bool(rep.waves_gave_up))


def latency_summary(arrival_s: np.ndarray, complete_s: np.ndarray, skip: int = 0) -> dict:
    """p50 / p99 / mean latency (ms) of completed requests after the first
    `skip` arrivals."""
    a = np.asarray(arrival_s)[skip:]
    c = np.asarray(complete_s)[skip:]
    ok = c >= 0
    lat = (c[ok] - a[ok]) * 1e3
    if lat.size == 0:
        return {"n": 0}
    return {"n": int(lat.size), "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_ms": round(float(lat.mean()), 3), "max_ms": round(float(lat.max()), 3)}


# ---------------------------------------------------------------------------
# Online sessions: requests submitted one at a time while the kernels run
# ---------------------------------------------------------------------------

@dataclass
class OnlineClass:
    """One band class of an online session.  cigar_cap=None keeps every run
    (a banded local alignment has at most 2*max_qlen + 2*half_band + 2);
    ring=None holds the next power of two >= 2 * n_slots requests in flight."""

    half_band: int
    max_qlen: int
    n_slots: int
    ring: Optional[int] = None
    cigar_cap: Optional[int] = None
    tile: Optional[SwTile] = None
    plane_budget_bytes: Optional[int] = None  # paged planes (admission by footprint)
    page_blocks: Optional[int] = None  # 8-row blocks per page (default: about 1 MB pages)
    page_wait_s: float = 0.0  # bound of one wait for pages (0: the library's default)
    prio_rows: int = 0  # requests with at least this many query rows run at raised priority


@dataclass
class ServedAlignment:
    cls: int
    ticket: int
    score: int
    query_end: int
    target_end: int
    read_start: int
    ref_start: int
    cigar: list
    counts: tuple
    truncated: bool
    no_pages: bool = False  # no traceback pages within the wait bound: no outputs


def _pow2_at_least(n: int) -> int:
    p = 1
    while p < n:
        p *= 2
    return p


class ServeSession:
    """Online serving of banded SW + CIGAR requests (gmfa_serve_open /
    submit / poll / close) over a resident reference.  One host thread per
    session.  The device memory of every class (traceback slabs, query stage
    slots, ticket head) is held here as torch tensors; the library holds the
    request rings in host-coherent memory.

        with ServeSession(ref, [OnlineClass(32, 200, 512)], (2, 4, 4, 2)) as s:
            t = s.submit(0, query_codes, t_start, t_len, band_offset)
            done = s.poll()
    """

    def __init__(self, reference: ResidentReference, classes: Sequence[OnlineClass],
                 scoring: tuple, idle_timeout_s: float = 1.0, poll_results: int = 4096):
        L = lib()
        self.reference = reference
        dev = reference.words.device
        ma, mi, go, ge = (int(x) for x in scoring)
        self.classes = list(classes)
        self._keep = []
        cfgs = (_lib.ServeClassCfg * len(self.classes))()
        max_cap = 1
        for k, oc in enumerate(self.classes):
            tile = oc.tile or serve_tile(oc.half_band)
            if tile.lanes_per_pair != 64:
                raise ValueError("serving needs a 64-lane tile (one pair per wavefront)")
            ring = int(oc.ring) if oc.ring is not None else _pow2_at_least(2 * int(oc.n_slots))
            cap = (int(oc.cigar_cap) if oc.cigar_cap is not None
                   else 2 * int(oc.max_qlen) + 2 * int(oc.half_band) + 2)
            max_cap = max(max_cap, cap)
            slab = max(int(kernels.tb_bytes(np.array([int(oc.max_qlen)]), tile)[0]) // 4, 1)
            sw = int(L.gmfa_serve_stage_words(int(oc.max_qlen)))
            pages = None
            if oc.plane_budget_bytes is not None:
                pages = page_geometry(tile, int(oc.max_qlen), oc.plane_budget_bytes,
                                      oc.page_blocks)
                owner = torch.zeros(pages.n_pages, dtype=torch.int32, device=dev)
                self._keep.append(owner)
                tb = torch.empty(pages.n_pages * pages.page_words, dtype=torch.int32, device=dev)
            else:
                tb = torch.empty(int(oc.n_slots) * slab, dtype=torch.int32, device=dev)
            stage = torch.zeros(int(oc.n_slots) * 2 * sw, dtype=torch.int32, device=dev)
            head = torch.zeros(8, dtype=torch.int64, device=dev)  # 64-byte control block
            self._keep += [tb, stage, head]
            c = cfgs[k]
            c.params = kernels.make_params(ma, mi, go, ge, int(oc.half_band),
                                           _lib.SW_TRACEBACK | _lib.SW_CIGAR, tile,
                                           prio_rows=int(oc.prio_rows))
            c.max_qlen, c.ring, c.n_slots, c.cigar_cap = int(oc.max_qlen), ring, int(oc.n_slots), cap
            c.tb, c.slab_words, c.stage, c.head = ptr(tb), slab, ptr(stage), ptr(head)
            if pages is not None:
                c.page_blocks, c.n_pages = pages.page_blocks, pages.n_pages
                c.page_owner = ptr(owner)
                c.page_wait_s = float(oc.page_wait_s)
        self._h = None
        self._results = (_lib.ServeResultRec * int(poll_results))()
        self._cigar = np.zeros(max(max_cap, 1 << 16), np.uint32)
        handle = ctypes.c_void_p()
        torch.cuda.synchronize(dev)
        check(L.gmfa_serve_open(ptr(reference.words), None, int(reference.n_bases), cfgs,
                                len(self.classes), float(idle_timeout_s), ctypes.byref(handle)),
              "gmfa_serve_open")
        self._h = handle
        self._pending = []

    # -- requests ----------------------------------------------------------
    def submit(self, cls: int, query: np.ndarray, t_start: int, t_len: int, band_offset: int,
               two_bit: bool = True, wait_s: float = 30.0) -> int:
        """Publish one request; returns its ticket (per class).  query: codes
        0..3 (two_bit) or A,C,G,T = 1..4.  When the class's ring is full the
        call polls (keeping the results for the next poll()) until an entry
        frees, for at most wait_s."""
        q = np.ascontiguousarray(query, dtype=np.uint8)
        tk = ctypes.c_int64()
        deadline = time.monotonic() + float(wait_s)
        while True:
            rc = lib().gmfa_serve_submit(self._handle(), int(cls), q.ctypes.data, len(q),
                                         1 if two_bit else 0, int(t_start), int(t_len),
                                         int(band_offset), ctypes.byref(tk))
            if rc != _lib.GMFA_E_AGAIN:
                check(rc, "gmfa_serve_submit")
                return int(tk.value)
            self._pending.extend(self._poll_raw(len(self._results)))
            if time.monotonic() > deadline:
                raise RuntimeError(f"class {cls}: ring still full after {wait_s} s")

    def submit_many(self, cls, codes: np.ndarray, code_off, qlen, t_start, t_len, band_offset,
                    two_bit: bool = True):
        """Publish a burst of requests (gmfa_serve_submit_many); stops early
        at a full ring.  Returns the tickets of the requests published (a
        prefix of the burst)."""
        n = len(qlen)
        arrs = [np.ascontiguousarray(cls, np.int32), np.ascontiguousarray(code_off, np.int64),
                np.ascontiguousarray(qlen, np.int32), np.ascontiguousarray(t_start, np.int64),
                np.ascontiguousarray(t_len, np.int32), np.ascontiguousarray(band_offset, np.int32)]
        codes = np.ascontiguousarray(codes, np.uint8)
        tickets = np.empty(n, np.int64)
        got = ctypes.c_int32()
        c, co, ql, to, tl, bo = (a.ctypes.data for a in arrs)
        rc = lib().gmfa_serve_submit_many(self._handle(), n, c, codes.ctypes.data, co, ql,
                                          1 if two_bit else 0, to, tl, bo, tickets.ctypes.data,
                                          ctypes.byref(got))
        if rc != _lib.GMFA_E_AGAIN:
            check(rc, "gmfa_serve_submit_many")
        return tickets[:got.value]

    RESULT_DTYPE = np.dtype([("ticket", np.int64), ("cls", np.int32), ("status", np.int32),
                             ("score", np.int32), ("query_end", np.int32),
                             ("target_end", np.int32), ("read_start", np.int32),
                             ("ref_start", np.int32), ("n_cigar", np.int32),
                             ("counts", np.int32, (4,)), ("cigar_off", np.int64),
                             ("claim_tick", np.int64), ("done_tick", np.int64)])

    def poll_arrays(self):
        """Completed requests as a structured array (RESULT_DTYPE) plus a copy
        of their CIGAR runs ((length << 4) | op; cigar_off indexes it)."""
        got = ctypes.c_int32()
        check(lib().gmfa_serve_poll(self._handle(), self._results, len(self._results),
                                    self._cigar.ctypes.data, len(self._cigar),
                                    ctypes.byref(got)), "gmfa_serve_poll")
        n = got.value
        rec = np.frombuffer(self._results, dtype=self.RESULT_DTYPE, count=n).copy()
        end = int((rec["cigar_off"] + np.where(rec["status"] == 0, rec["n_cigar"], 0)).max()) if n else 0
        return rec, self._cigar[:end].copy()

    def poll(self, max_results: Optional[int] = None) -> list:
        """Completed requests (never blocks)."""
        out, self._pending = self._pending, []
        n = len(self._results) if max_results is None else int(max_results)
        return out + self._poll_raw(max(0, n - len(out)))

    def wait(self, tickets: int, timeout_s: float = 30.0) -> list:
        """Poll until `tickets` results have come back (or the timeout)."""
        got = []
        end = time.monotonic() + timeout_s
        while len(got) < tickets:
            got += self.poll()
            if time.monotonic() > end:
                raise RuntimeError(f"{len(got)} of {tickets} requests completed in {timeout_s} s")
        return got

    def stats(self) -> dict:
        st = _lib.ServeStats()
        check(lib().gmfa_serve_session_stats(self._handle(), ctypes.byref(st)), "stats")
        return {"submitted": st.submitted, "returned": st.returned,
                "in_flight": st.in_flight, "launches": st.launches,
                "ticks_per_s": st.ticks_per_s}

    def close(self) -> ServeResult:
        if self._h is None:
            return ServeResult(0.0, 0, False, False)
        rep = ServeReport()
        h, self._h = self._h, None
        check(lib().gmfa_serve_close(h, ctypes.byref(rep)), "gmfa_serve_close")
        return ServeResult(wall_s=0.0, completed=int(rep.completed), timed_out=False,
                           waves_gave_up=bool(rep.waves_gave_up))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None:
                self.close()
        except Exception:
            pass

    # -- internals ---------------------------------------------------------
    def _handle(self):
        if self._h is None:
            raise RuntimeError("serving session is closed")
        return self._h

    def _poll_raw(self, n: int) -> list:
        if n <= 0:
            return []
        got = ctypes.c_int32()
        check(lib().gmfa_serve_poll(self._handle(), self._results, min(n, len(self._results)),
                                    self._cigar.ctypes.data, len(self._cigar),
                                    ctypes.byref(got)), "gmfa_serve_poll")
        out = []
        for k in range(got.value):
            r = self._results[k]
            trunc = r.status == _lib.GMFA_SERVE_TRUNCATED
            nopages = r.status == _lib.GMFA_SERVE_NO_PAGES
            runs = self._cigar[r.cigar_off:r.cigar_off + (0 if trunc or nopages else r.n_cigar)]
            cig = [(_lib.OP_CHAR[int(v) & 15], int(v) >> 4) for v in runs]
            out.append(ServedAlignment(cls=int(r.cls), ticket=int(r.ticket), score=int(r.score),
                                       query_end=int(r.query_end), target_end=int(r.target_end),
                                       read_start=int(r.read_start), ref_start=int(r.ref_start),
                                       cigar=cig, counts=tuple(int(x) for x in r.counts),
                                       truncated=trunc, no_pages=nopages))
        return out
