"""CPU checks of the C ABI: the library loads and exports every declared symbol,
host-side entry points work without a GPU."""

import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "gmfa.h")) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gmfa_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    from graphmambaformer_amd import _lib

    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    from graphmambaformer_amd import _lib

    L = _lib.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name
    assert L.gmfa_abi_version() == 2


def test_tile_resolution_and_tb_bytes():
    from graphmambaformer_amd import kernels as K

    t = K.resolve_tile(256)  # W = 513 = 8 x 64 + the scalar cell (plane lane 64)
    assert (t.cells_per_lane, t.lanes_per_pair, t.lanes_stored) == (8, 64, 65)
    t = K.resolve_tile(256, 9, 64)
    assert (t.cells_per_lane, t.lanes_per_pair, t.lanes_stored) == (9, 64, 57)
    t = K.resolve_tile(255, 8, 64)  # W = 511: the general 8x64 tile
    assert (t.cells_per_lane, t.lanes_per_pair, t.lanes_stored) == (8, 64, 64)
    t = K.resolve_tile(32)
    assert t.cells_per_lane * t.lanes_per_pair >= 65
    for hb in (0, 1, 5, 64, 200, 511, 751, 1023):
        t = K.resolve_tile(hb)
        assert t.cells_per_lane * t.lanes_per_pair >= 2 * hb + 1
        assert t.lanes_stored == -(-(2 * hb + 1) // t.cells_per_lane)
    from graphmambaformer_amd._lib import lib

    assert lib().gmfa_sw_tb_bytes(13, t) == 2 * t.lanes_stored * t.cells_per_lane * 4
    assert K.tb_bytes(np.array([13]), t)[0] == 2 * t.lanes_stored * t.cells_per_lane * 4


def test_invalid_arguments_raise_value_error():
    import pytest

    from graphmambaformer_amd import kernels as K

    with pytest.raises(ValueError):
        K.resolve_tile(5000)
    with pytest.raises(RuntimeError):
        K.resolve_tile(300, 5, 16)  # tile too small for the band


def test_host_backtrack_matches_oracle():
    from graphmambaformer_amd import kernels as K
    from oracle import oracle as O

    rng = np.random.default_rng(4)
    for _ in range(30):
        n = int(rng.integers(0, 80))
        f = rng.integers(0, 60, size=n).astype(np.float64)
        par = np.array([int(rng.integers(-1, i)) if i else -1 for i in range(n)], np.int64)
        got = K.chain_backtrack(f, par, 1, 10.0)
        exp = O.chain_backtrack(f, par, 1, 10.0)
        assert [(m.tolist(), s) for m, s in got] == [(m.tolist(), s) for m, s in exp]


def test_pack_host_layout():
    from graphmambaformer_amd import kernels as K

    seqs = [np.array([1, 2, 3, 4, 5, -1, 1], np.int16), np.array([4] * 17, np.int16)]
    w, m, off, has_n = K.pack_host(seqs)
    assert has_n and off.tolist() == [0, 16]
    assert (w[0] & 0xFF) == 0b11100100  # A C G T -> 0 1 2 3
    assert (m[0] >> 8) & 0xF == 0b0101  # N and padding masked
    assert w[1] == 0xFFFFFFFF and (w[2] & 3) == 3


def test_serve_run_validates_before_touching_the_device():
    import ctypes

    import pytest

    from graphmambaformer_amd import _lib

    L = _lib.lib()
    rep = _lib.ServeReport()
    assert L.gmfa_serve_run(None, 1, 1.0, ctypes.byref(rep)) == _lib.GMFA_E_INVALID
    arr = (_lib.ServeClass * 1)()
    assert L.gmfa_serve_run(arr, 0, 1.0, ctypes.byref(rep)) == _lib.GMFA_E_INVALID
    assert L.gmfa_serve_run(arr, 1, -1.0, ctypes.byref(rep)) == _lib.GMFA_E_INVALID
    # a class with requests but no arrival times
    arr[0].batch.n_pairs = 3
    assert L.gmfa_serve_run(arr, 1, 1.0, ctypes.byref(rep)) == _lib.GMFA_E_INVALID
    t = (ctypes.c_double * 3)(0.0, 2.0, 1.0)  # not nondecreasing
    c = (ctypes.c_double * 3)()
    arr[0].arrival_s = ctypes.cast(t, ctypes.c_void_p)
    arr[0].complete_s = ctypes.cast(c, ctypes.c_void_p)
    assert L.gmfa_serve_run(arr, 1, 1.0, ctypes.byref(rep)) == _lib.GMFA_E_INVALID
    assert b"nondecreasing" in L.gmfa_last_error()
    with pytest.raises(ValueError):
        _lib.check(L.gmfa_serve_run(arr, 1, 1.0, ctypes.byref(rep)))


def test_entry_points_reject_bad_arguments_without_a_device():
    """Every entry point validates its arguments before any HIP call, so bad
    input is a ValueError (GMFA_E_INVALID) even on a host without a GPU."""
    import ctypes

    from graphmambaformer_amd import _lib

    L = _lib.lib()
    E = _lib.GMFA_E_INVALID
    ix = _lib.KmerIndex()
    assert L.gmfa_kmer_lookup(None, None, 1, None, None, None) == E
    ix.k, ix.dir_bits = 0, 0
    assert L.gmfa_kmer_lookup(ctypes.byref(ix), None, 1, None, None, None) == E
    ix.k, ix.dir_bits = 15, 40  # dir_bits above 2k
    assert L.gmfa_kmer_lookup(ctypes.byref(ix), None, 1, None, None, None) == E
    ix.dir_bits = 8
    rd = _lib.KmerReads()
    rd.window = 0
    assert L.gmfa_kmer_seed(ctypes.byref(ix), ctypes.byref(rd), None, None, None, None, None,
                            None) == E
    assert L.gmfa_wfa_distance(None, 10, 10, None, None, None) == E
    b = _lib.SwBatch()
    b.n_pairs = 1  # pointers missing
    assert L.gmfa_wfa_distance(ctypes.byref(b), 10, 10, None, None, None) == E
    assert L.gmfa_wfa_align(ctypes.byref(b), None, None, None, None, None, None, None, None,
                            None) == E
    assert L.gmfa_chain_sort(None, -1, 0, *([None] * 13), None) == E
    assert L.gmfa_chain_backtrack_device(None, 1, -5, *([None] * 7), 1, 0.0, *([None] * 8),
                                         None) == E
    assert L.gmfa_chain_select(1, *([None] * 5), 0.6, 0.5, 0, *([None] * 4), None) == E
    assert L.gmfa_cigar_bam(-1, *([None] * 13), None) == E
    assert L.gmfa_pack_flat(None, -1, 0, None, None, None, None) == E
    assert L.gmfa_sw_xdrop(None, None, None, None, None, 1, 0, 0, 10, None, None, None, None,
                           None) == E
    assert L.gmfa_sw_xdrop(None, None, None, None, None, 1, 4, 0, 5000, None, None, None, None,
                           None) == E
    assert b"half_band" in L.gmfa_last_error()


def test_serve_session_validates_before_touching_the_device():
    """gmfa_serve_open / submit / poll / close reject bad arguments (no
    session, bad class configuration) before any HIP call."""
    import ctypes

    from graphmambaformer_amd import _lib

    L = _lib.lib()
    E = _lib.GMFA_E_INVALID
    h = ctypes.c_void_p()
    cfg = (_lib.ServeClassCfg * 1)()
    assert L.gmfa_serve_open(None, None, 100, cfg, 1, 1.0, ctypes.byref(h)) == E
    ref = ctypes.c_void_p(0x1000)  # never dereferenced: validation fails first
    assert L.gmfa_serve_open(ref, None, 0, cfg, 1, 1.0, ctypes.byref(h)) == E
    assert L.gmfa_serve_open(ref, None, 100, cfg, 0, 1.0, ctypes.byref(h)) == E
    assert L.gmfa_serve_open(ref, None, 100, cfg, 1, 0.0, ctypes.byref(h)) == E
    c = cfg[0]
    c.max_qlen, c.ring, c.n_slots, c.cigar_cap = 100, 6, 4, 100  # ring not a power of two
    assert L.gmfa_serve_open(ref, None, 100, cfg, 1, 1.0, ctypes.byref(h)) == E
    assert b"power of two" in L.gmfa_last_error()
    c.ring = 8  # tb / stage / head missing
    assert L.gmfa_serve_open(ref, None, 100, cfg, 1, 1.0, ctypes.byref(h)) == E
    c.tb = c.stage = c.head = 0x1000
    c.params.flags = _lib.SW_NMASK  # N masks without a reference mask
    assert L.gmfa_serve_open(ref, None, 100, cfg, 1, 1.0, ctypes.byref(h)) == E
    assert h.value is None
    tk = ctypes.c_int64()
    assert L.gmfa_serve_submit(None, 0, None, 0, 1, 0, 0, 0, ctypes.byref(tk)) == E
    n = ctypes.c_int32()
    assert L.gmfa_serve_poll(None, None, 0, None, 0, ctypes.byref(n)) == E
    assert L.gmfa_serve_close(None, None) == E
    assert L.gmfa_serve_stage_words(100) == (100 + 15) // 16 + 2


def test_noalloc_validates_before_touching_the_device():
    """gmfa_sw_forward_noalloc rejects a batch whose planes exceed its static
    scratch, and caller-supplied plane / counter buffers, before any HIP call."""
    import ctypes

    from graphmambaformer_amd import _lib

    L = _lib.lib()
    E = _lib.GMFA_E_INVALID
    b = _lib.SwBatch()
    p = _lib.SwParams()
    o = _lib.SwOut()
    b.n_pairs, b.max_qlen = 1, 1_000_000
    p.match, p.mismatch, p.gap_open, p.gap_extend = 2, 4, 4, 2
    p.half_band = 1023
    p.flags = _lib.SW_TRACEBACK  # no fused walk
    assert L.gmfa_sw_forward_noalloc(ctypes.byref(b), ctypes.byref(p), ctypes.byref(o), None) == E
    p.flags = _lib.SW_TRACEBACK | _lib.SW_CIGAR
    o.tb = 0x1000  # planes are the library's
    assert L.gmfa_sw_forward_noalloc(ctypes.byref(b), ctypes.byref(p), ctypes.byref(o), None) == E
    assert b"must be NULL" in L.gmfa_last_error()
    o.tb = None
    # 125,000 blocks x 2,048 plane dwords: about 1 GB for one wavefront
    assert L.gmfa_sw_forward_noalloc(ctypes.byref(b), ctypes.byref(p), ctypes.byref(o), None) == E
    assert b"static scratch" in L.gmfa_last_error()
    assert _lib.GMFA_NOALLOC_SCRATCH_BYTES == 256 << 20


def test_drop_ins_reject_bands_and_lookbacks_beyond_the_kernels():
    """The reference takes any half_band (extension.py:105) and max_lookback
    (chaining.py:154); the kernels here hold half_band <= 1023 and
    max_lookback <= 256 (the reference defaults are 512 and 64, config.py:681,
    760).  Beyond the caps the drop-ins raise ValueError before any device
    work; at the caps they accept."""
    import torch

    from graphmambaformer_amd import kernels as K
    from graphmambaformer_amd.alignment import chaining as C
    from graphmambaformer_amd.alignment import extension as E
    from graphmambaformer_amd.config import ChainingConfig, ExtensionConfig

    with pytest.raises(ValueError, match="1023"):
        K.resolve_tile(1024)
    t = K.resolve_tile(E.MAX_HALF_BAND)
    assert t.cells_per_lane * t.lanes_per_pair >= 2 * E.MAX_HALF_BAND + 1
    q = torch.ones((1, 10), dtype=torch.int16)
    n = torch.tensor([10])
    with pytest.raises(ValueError, match="half_band 1024"):
        E.banded_affine_sw_batch(q, q, n, n, ExtensionConfig(), half_band=1024)
    with pytest.raises(ValueError, match="half_band 1100"):
        E.banded_affine_sw_batch(q, q, n, n, ExtensionConfig(half_band=1100))
    with pytest.raises(ValueError, match="max_lookback 257"):
        C.AffineChainer(ChainingConfig(max_lookback=257))
    x = np.arange(5, dtype=np.int64)
    with pytest.raises(ValueError, match="max_lookback 257"):
        C.chain_dp_numpy(x, x, np.ones(5), ChainingConfig(max_lookback=257))
    C.AffineChainer(ChainingConfig(max_lookback=C.MAX_LOOKBACK))


def test_narrow_and_batch_tiles_on_the_host():
    """The 64-lane tile a small batch or a serving class gets for each band
    (host-side tile tables only, no device work): the scalar-cell tiles for
    W = 64*C + 1, else the smallest general 64-lane tile that holds the band;
    large batches keep the library's default."""
    from graphmambaformer_amd import kernels as K

    def shape(t):
        return (t.cells_per_lane, t.lanes_per_pair, t.lanes_stored)

    assert shape(K.narrow_tile(32)) == (1, 64, 65)     # W = 65: 1 x 64 + the scalar cell
    assert shape(K.narrow_tile(64)) == (4, 64, 33)     # W = 129: 2 x 64 holds 128 only
    assert shape(K.narrow_tile(256)) == (8, 64, 65)    # W = 513: 8 x 64 + the scalar cell
    assert shape(K.narrow_tile(751)) == (24, 64, 63)
    assert shape(K.narrow_tile(2)) == (2, 64, 3)
    # the default for W = 65 packs four pairs to a wavefront; a small batch
    # (fewer than 4 wavefronts per SIMD of a 256-CU device) takes the latency
    # tile: the skewed 2 x 63 tile up to 126 diagonals, else the narrow tile
    assert shape(K.resolve_tile(32))[:2] == (5, 16)
    assert shape(K.batch_tile(32, 1024)) == (2, 63, 33)
    assert shape(K.batch_tile(32, 64 * 4 * 4 * 256))[:2] == (5, 16)
    assert shape(K.latency_tile(62)) == (2, 63, 63)    # W = 125
    assert shape(K.latency_tile(63)) == (2, 64, 64)    # W = 127: past the skewed tile
    assert shape(K.latency_tile(256)) == (8, 64, 65)
    for hb in (2, 7, 31, 32, 33, 64, 100, 255, 256, 300, 751, 1023):
        t = K.narrow_tile(hb)
        assert t is not None and t.lanes_per_pair == 64, hb
        assert t.lanes_stored * t.cells_per_lane >= 2 * hb + 1, hb
    # the skewed tile is never a default, holds up to 126 diagonals and has no
    # form with per-row keys
    assert (2, 63) in K.list_tiles()
    for hb in range(0, 1024, 7):
        assert K.resolve_tile(hb).lanes_per_pair != 63, hb
    with pytest.raises((RuntimeError, ValueError)):
        K.resolve_tile(63, 2, 63)



def test_online_session_rejects_a_split_class():
    """Online sessions take no split classes: gmfa_serve_open refuses one with
    GMFA_E_INVALID while validating its arguments, before any device call."""
    import ctypes

    from graphmambaformer_amd import _lib, kernels as K
    from graphmambaformer_amd._lib import lib

    cfg = (_lib.ServeClassCfg * 1)()
    c = cfg[0]
    c.params = K.make_params(2, 4, 4, 2, 751, _lib.SW_TRACEBACK | _lib.SW_CIGAR,
                             K.resolve_tile(751), split_rows=1, split_slices=3)
    c.max_qlen, c.ring, c.n_slots, c.cigar_cap = 1000, 8, 2, 100
    dummy = ctypes.c_void_p(0x1000)  # never dereferenced: validation fails first
    c.tb, c.stage, c.head = dummy, dummy, dummy
    sess = ctypes.c_void_p()
    rc = lib().gmfa_serve_open(dummy, None, 1000, cfg, 1, 1.0, ctypes.byref(sess))
    assert rc == _lib.GMFA_E_INVALID
    assert b"split" in lib().gmfa_last_error()
    assert not sess.value
