"""Drop-in stage modules (seeding lookup, chaining, extension) of the alignment path."""

from .chaining import AffineChainer, chain_dp_batched, chain_dp_numpy  # noqa: F401
from .extension import (  # noqa: F401
    BandedDPResult,
    ExtensionEngine,
    WavefrontAligner,
    banded_affine_sw_batch,
    traceback_banded,
)
from .seeding import (  # noqa: F401
    N_CODE,
    GPUKmerIndex,
    KmerTable,
    decode_bases,
    encode_bases,
    hash64,
    minimizer_mask,
    pack_kmers,
    reverse_complement_codes,
)
from .types import AnchorSet, Chain, ExtensionResult, merge_cigar, run_length_encode  # noqa: F401
