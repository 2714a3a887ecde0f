"""GPU parity of the k-mer seeding kernels (gmfa_kmer_lookup / seed / emit)
against the oracle and the reference's golden vectors, exact."""

import numpy as np
import pytest
import torch

from oracle import seeding as S
from tests.golden.load import load

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _say(*a):
    print(*a, flush=True)


def test_kmer_query_golden():
    from graphmambaformer_amd.alignment import GPUKmerIndex, encode_bases

    for n, case in enumerate(load("seeding.json")["queries"]):
        _say("case: seeding golden", n, case["k"], case["window"], case["max_occ"])
        idx = GPUKmerIndex(encode_bases(case["ref"]), k=case["k"], window=case["window"],
                           max_occ=case["max_occ"], device=DEV)
        assert int(idx.keys.numel()) == case["n_keys"]
        got = idx.query_batch([encode_bases(r["read"]) for r in case["reads"]])
        for r, (rp, fp, ln) in zip(case["reads"], got):
            assert rp.tolist() == r["read_pos"], n
            assert fp.tolist() == r["ref_pos"], n
            assert ln.tolist() == r["length"], n
        one = idx.query(encode_bases(case["reads"][0]["read"]))
        assert one[0].tolist() == case["reads"][0]["read_pos"]


def _random_ref(rng, n):
    ref = rng.integers(1, 5, size=n).astype(np.int8)
    unit = rng.integers(1, 5, size=60).astype(np.int8)
    for s in rng.integers(0, n - 3000, size=20):
        ref[s:s + 1800] = np.tile(unit, 30)
    ref[rng.integers(0, n, size=200)] = 5
    return ref


@pytest.mark.parametrize("k,w,occ,dir_bits", [(15, 10, 0, 22), (21, 5, 0, 8), (11, 16, 5, 0),
                                              (13, 1, 0, 26)])
def test_kmer_batch_vs_oracle(k, w, occ, dir_bits):
    from graphmambaformer_amd.alignment import GPUKmerIndex

    _say("case: seeding batch", k, w, occ, dir_bits)
    rng = np.random.default_rng(k * 100 + w)
    ref = _random_ref(rng, 1_000_000)
    idx = GPUKmerIndex(ref, k=k, window=w, max_occ=occ, device=DEV, dir_bits=dir_bits)
    tab = S.build_index(ref, k, w, occ)
    reads = []
    for r in range(160):
        L = int(rng.integers(1, 3000)) if r % 10 else int(rng.integers(0, k + 2))
        s0 = int(rng.integers(0, len(ref) - L - 1))
        rd = ref[s0:s0 + L].copy()
        rd[rng.integers(0, max(L, 1), size=min(L, 8))] = rng.integers(1, 6, size=min(L, 8))
        reads.append(rd)
    got = idx.query_batch(reads)
    for r, (rp, fp, ln) in enumerate(got):
        erp, efp, eln = S.kmer_query(tab, reads[r], w)
        assert np.array_equal(rp, erp), r
        assert np.array_equal(fp, efp), r
        assert np.array_equal(ln, eln), r
    # the lookup entry point: present keys, absent keys, out-of-range values
    q = np.concatenate([tab.keys[::5], rng.integers(-3, 1 << (2 * k), size=3000),
                        [(1 << (2 * k)), (1 << 62)]]).astype(np.int64)
    st, ct = idx.lookup(torch.from_numpy(q))
    es, ec = tab.lookup(q)
    assert np.array_equal(st.cpu().numpy(), es.astype(np.int32))
    assert np.array_equal(ct.cpu().numpy(), ec)


def test_kmer_rejects_bad_inputs():
    from graphmambaformer_amd.alignment import GPUKmerIndex

    _say("case: seeding invalid inputs")
    ref = np.random.default_rng(0).integers(1, 5, size=5000).astype(np.int8)
    with pytest.raises(ValueError):
        GPUKmerIndex(ref, k=15, window=200, device=DEV)
    idx = GPUKmerIndex(ref, k=15, window=10, device=DEV)
    with pytest.raises(ValueError):
        idx.query(np.array([1, 2, 0, 3] * 10, np.int8))
    torch.cuda.synchronize()
