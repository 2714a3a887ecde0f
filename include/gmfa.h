/*
 * gmfa.h -- C ABI of libgmfa.so, the MI355X (gfx950) banded affine-gap
 * Smith-Waterman + traceback and anchor-chaining library.
 *
 * Every entry point is extern "C", takes plain pointers and sizes (device
 * pointers unless stated), returns 0 on success or a negative GMFA_E_* code,
 * and leaves a message for gmfa_last_error() (thread-local) on failure.
 * Streams are hipStream_t passed as void*; NULL means the null stream.
 * The library keeps no mutable global state besides the thread-local error
 * string and per-device read-only properties, so calls are reentrant.
 *
 * The reference path these entry points replace (files of
 * sarakh1999/GraphMambaFormer):
 *   gmfa_sw_forward    graphmambaformer/alignment/extension.py:78-204
 *                      banded_affine_sw_batch (and the NVIDIA tier
 *                      accel/cuda_kernels.py:504-566 banded_sw)
 *   gmfa_sw_xdrop      graphmambaformer/alignment/extension.py:186-193
 *                      (batch-wide x-drop, one group per reference batch)
 *   gmfa_sw_traceback  graphmambaformer/alignment/extension.py:207-302
 *                      traceback_banded
 *   gmfa_pack_codes    graphmambaformer/alignment/seeding.py:46-80 code
 *                      alphabet (A,C,G,T=1..4, N=5, padding<0) -> 2-bit + N mask
 *   gmfa_chain_dp_f64  graphmambaformer/alignment/chaining.py:129-180
 *                      chain_dp_numpy (and chaining.py:183-258
 *                      chain_dp_batched, cuda_kernels.py:440-501 chain_dp)
 *   gmfa_chain_dp_i32  the same DP on the integer profile
 *   gmfa_chain_backtrack graphmambaformer/alignment/chaining.py:530-576
 *                      AffineChainer._backtrack (host code)
 *   gmfa_wfa_distance / gmfa_wfa_align  graphmambaformer/alignment/extension.py:308-463
 *                      WavefrontAligner.align for a batch of pairs
 *   gmfa_cigar_bam     graphmambaformer/data/formats.py:52, 343-382 CIGAR of a
 *                      BAM record (soft clips, strand, SAM length rule)
 *   gmfa_chain_sort    graphmambaformer/alignment/types.py:149-152 (DP order)
 *   gmfa_chain_backtrack_device / gmfa_chain_select  chaining.py:530-617
 *                      (_backtrack and select on the device)
 *   gmfa_kmer_lookup   graphmambaformer/alignment/seeding.py:751-775
 *                      GPUKmerIndex.lookup (accel/cuda_kernels.py:44-74)
 *   gmfa_kmer_seed / gmfa_kmer_emit  seeding.py:777-798 GPUKmerIndex.query
 *                      for a batch of reads
 *   gmfa_serve_run     the per-read extension calls of
 *                      graphmambaformer/alignment/pipeline.py:355-363
 *                      (ExtensionEngine.extend_chains, extension.py:511-641,
 *                      one read after another) as a continuously batched
 *                      server: requests are aligned as they arrive
 *   gmfa_serve_open / gmfa_serve_submit / gmfa_serve_poll / gmfa_serve_close
 *                      the same per-read calls as an online session fed by a
 *                      front end one request at a time
 */
#ifndef GMFA_H
#define GMFA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GMFA_ABI_VERSION 2

#define GMFA_OK 0
#define GMFA_E_INVALID -1   /* bad argument: maps to ValueError */
#define GMFA_E_RUNTIME -2   /* HIP runtime / launch failure: maps to RuntimeError */
#define GMFA_E_UNSUPPORTED -3 /* shape outside the compiled variants: RuntimeError */
#define GMFA_E_AGAIN -4       /* serving ring full: poll, then retry */

/* flags for gmfa_sw_params.flags */
#define GMFA_SW_TRACEBACK 1u  /* write the 4-bit traceback planes */
#define GMFA_SW_NMASK 2u      /* sequences carry N masks (qmask/tmask valid) */
#define GMFA_SW_ROWKEYS 4u    /* write per-row (max H, first column) keys */
#define GMFA_SW_CIGAR 8u      /* walk each pair right after its forward pass and
                                 write its CIGAR (needs GMFA_SW_TRACEBACK, no ROWKEYS) */

/* BAM CIGAR op codes used in traceback output (len << 4 | op) */
#define GMFA_OP_I 1
#define GMFA_OP_D 2
#define GMFA_OP_S 4
#define GMFA_OP_EQ 7
#define GMFA_OP_X 8

typedef struct gmfa_sw_params {
    int32_t match;       /* cfg.match_score (integer) */
    int32_t mismatch;    /* cfg.mismatch_penalty (integer, positive) */
    int32_t gap_open;    /* cfg.gap_open; a gap of length L costs open+(L-1)*extend */
    int32_t gap_extend;  /* cfg.gap_extend, <= gap_open */
    int32_t half_band;   /* band is 2*half_band+1 wide, <= 1023 */
    uint32_t flags;      /* GMFA_SW_* */
    int32_t cells_per_lane; /* 0 = library default for the band; else a compiled tile */
    int32_t lanes_per_pair; /* 0 = library default; else 16, 32 or 64 */
    /* Split launch (gmfa_sw_forward with exactly GMFA_SW_TRACEBACK |
     * GMFA_SW_CIGAR, a 64-lane tile and batch->order set): split_rows > 0 runs
     * the longest-first prefix of batch->order whose queries have at least
     * split_rows rows (at most 512 pairs) each on a workgroup of several
     * wavefronts, one band slice per wavefront (at most 3/4 of the resident
     * wavefronts in such workgroups; further split pairs queue behind the
     * first ones), in a kernel on the library's
     * own stream beside the plain kernel for the rest (forked from and joined
     * back into the caller's stream with events, so a graph capture of the
     * caller's stream holds both; the plain kernel waits, at most 5 ms, until
     * the split workgroups are resident); out->work_counter must then point
     * to 4 uint32.  Outputs and planes are those of the plain launch.
     * split_slices: 0 = the first compiled slice count for the band, else
     * that count.  Bands without a compiled split variant run the plain
     * launch.  0 / 0: the plain launch. */
    int32_t split_rows;
    int32_t split_slices;
    /* prio_rows > 0 (gmfa_sw_forward, and a serving class's params): pairs
     * (requests) with at least prio_rows query rows run at raised wave
     * priority, so the longest pairs of an uneven batch, or a class's longest
     * requests, keep their SIMD's issue slots and the other waves fill their
     * stalls. */
    int32_t prio_rows;
    /* max_waves > 0 (gmfa_sw_forward): the persistent launch holds at most this
     * many wavefronts (default: as many as fit on the device at once). */
    int32_t max_waves;
} gmfa_sw_params;

/* A batch of pairs whose sequences live in 2-bit packed streams
 * (16 bases per uint32, base k of a word at bits 2k..2k+1, A=0 C=1 G=2 T=3).
 * N masks use the same layout with value 1 marking a base that never matches. */
typedef struct gmfa_sw_batch {
    int32_t n_pairs;
    int32_t max_qlen;      /* host-side bound: every qlen <= max_qlen */
    const uint32_t *qseq;  /* query stream */
    const uint32_t *tseq;  /* target stream */
    const uint32_t *qmask; /* N masks or NULL (requires GMFA_SW_NMASK) */
    const uint32_t *tmask;
    int64_t q_words;       /* words in qseq / qmask */
    int64_t t_words;       /* words in tseq / tmask */
    const int64_t *q_off;  /* [n_pairs] first base of each query in qseq */
    const int64_t *t_off;  /* [n_pairs] first base of each target in tseq */
    const int32_t *qlen;   /* [n_pairs] */
    const int32_t *tlen;   /* [n_pairs] */
    const int32_t *band_offset; /* [n_pairs] diagonal of the band centre */
    const int32_t *order;  /* [n_pairs] processing order (longest first: a
                            * permutation of 0..n_pairs-1) or NULL; a launch
                            * that gives every pair a wavefront of its own
                            * does not read it */
} gmfa_sw_batch;

typedef struct gmfa_sw_out {
    int32_t *score;        /* [n_pairs] best local score */
    int32_t *query_end;    /* [n_pairs] 1-based row of the best cell (0 if none) */
    int32_t *target_end;   /* [n_pairs] 1-based column of the best cell */
    uint32_t *tb;          /* traceback planes (GMFA_SW_TRACEBACK) */
    const int64_t *tb_off; /* [n_pairs] dword offset of each pair's planes */
    int32_t *rowkey;       /* per-row keys (GMFA_SW_ROWKEYS): qlen int32 per pair,
                              key of row i = H << 11 | (2047 - d) for the row's
                              maximum H at its first band diagonal d */
    const int64_t *rowkey_off; /* [n_pairs] element offset of each pair's keys */
    uint32_t *work_counter;    /* one device uint32 of scratch for the work queue (4 for
                                  split launches): zero before the first launch that uses
                                  it; every launch leaves word 0 at zero (its last
                                  wavefront resets it), so it needs no clearing between
                                  launches */
    /* GMFA_SW_CIGAR outputs, same layout as gmfa_sw_traceback's */
    const int64_t *cigar_off;
    uint32_t *cigar;
    int32_t *n_cigar;
    int32_t *read_start;
    int32_t *ref_start;
    int32_t *counts;
} gmfa_sw_out;

typedef struct gmfa_sw_tile {
    int32_t cells_per_lane; /* band diagonals per lane */
    int32_t lanes_per_pair; /* lanes of a wavefront that share one pair */
    int32_t lanes_stored;   /* ceil(band / cells_per_lane): lanes with planes */
    int32_t half_band;
} gmfa_sw_tile;

/* library / device */
const char *gmfa_last_error(void);
int gmfa_abi_version(void);
int gmfa_device_info(int32_t *cu_count, int32_t *warp_size, char *arch_name, int32_t arch_len);

/* Tile (cells per lane, lanes per pair) the forward kernel uses for a band of
 * half-width half_band given a requested tile (0 = default); fails with
 * GMFA_E_UNSUPPORTED if no compiled variant covers it. */
int gmfa_sw_resolve_tile(int32_t half_band, int32_t cells_per_lane, int32_t lanes_per_pair,
                         gmfa_sw_tile *tile);

/* Compiled (cells_per_lane, lanes_per_pair) tile shapes, each once: the
 * general tiles (any band up to cells_per_lane x lanes_per_pair), then the
 * shapes compiled only for one band (1 x 64 plus a scalar cell: W = 65);
 * returns how many exist and fills up to cap entries.  gmfa_sw_resolve_tile
 * tells whether a shape holds a given band. */
int gmfa_sw_list_tiles(int32_t *cells_per_lane, int32_t *lanes_per_pair, int32_t cap);

/* Slice counts of the compiled split variants for a band (half_band, with the
 * regular tile of cells_per_lane cells per lane, 0 = default): returns how
 * many exist and fills up to cap entries (cells per lane of a slice, slices
 * per pair). */
int gmfa_sw_split_variants(int32_t half_band, int32_t cells_per_lane, int32_t *slice_cells,
                           int32_t *slices, int32_t cap);

/* Wavefronts of the forward kernel that params selects (band, tile, flags)
 * the current device holds at once: the persistent launch's grid.  A batch's
 * mean cells per resident wavefront is the share a pair must exceed to set
 * the batch's end (the pairs worth a split launch).  No counterpart in the
 * reference (its batches are one pair per workgroup). */
int gmfa_sw_resident_waves(const gmfa_sw_params *params, int32_t *waves);

/* Bytes of traceback planes one pair of query length qlen needs with tile. */
int64_t gmfa_sw_tb_bytes(int32_t qlen, const gmfa_sw_tile *tile);

/* Pack int16 base codes (rows of a (B, L) row-major matrix, row b of length
 * lens[b]) into the 2-bit stream at word offset word_off[b] (a multiple of
 * 16 bases).  Codes 1..4 -> A..T; code 5 and negative codes -> N mask.
 * *bad_codes (device int32, zeroed by the caller) receives nonzero if a code
 * outside {<0, 1..5} occurs within a row's length. */
int gmfa_pack_codes(const int16_t *codes, int64_t B, int64_t L, const int32_t *lens,
                    const int64_t *word_off, int64_t total_words, uint32_t *seq,
                    uint32_t *mask, int32_t *bad_codes, void *stream);

/* Pack a flat stream of n byte codes (sequences back to back) into the 2-bit
 * stream + N mask (words of 16 bases; padding past n is masked).  two_bit: codes
 * 0..3; else reference codes (1..4, 5 = N).  *flags (zeroed by the caller): bit
 * 0 = a code outside the alphabet, bit 1 = an N. */
int gmfa_pack_flat(const uint8_t *codes, int64_t n, int32_t two_bit, uint32_t *seq,
                   uint32_t *mask, int32_t *flags, void *stream);

/* Forward pass: scores, ends and (optionally) traceback planes / row keys. */
int gmfa_sw_forward(const gmfa_sw_batch *batch, const gmfa_sw_params *params,
                    const gmfa_sw_out *out, void *stream);

/* Size of the static device scratch of gmfa_sw_forward_noalloc (per GPU). */
#define GMFA_NOALLOC_SCRATCH_BYTES (256ll << 20)

/* gmfa_sw_forward with GMFA_SW_TRACEBACK | GMFA_SW_CIGAR, allocating nothing.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).  The
 * work counter and the traceback planes live in a static device array of
 * GMFA_NOALLOC_SCRATCH_BYTES per GPU, the planes as one slab per wavefront
 * sized for batch->max_qlen and reused pair after pair (the CIGAR walk runs
 * in the same launch), so the launch takes as many wavefronts as the slabs
 * fit: short queries run at full occupancy, long ones on fewer wavefronts.
 * out: score, query_end, target_end and the CIGAR outputs as for
 * gmfa_sw_forward; out->tb, tb_off, work_counter and rowkey must be NULL.
 * GMFA_E_INVALID if one wavefront's slabs exceed the scratch.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (other streams
 * or threads) race. */
int gmfa_sw_forward_noalloc(const gmfa_sw_batch *batch, const gmfa_sw_params *params,
                            const gmfa_sw_out *out, void *stream);

/* Batch-wide x-drop (graphmambaformer/alignment/extension.py:186-193) per
 * group of pairs, from the per-row keys of a GMFA_SW_ROWKEYS forward pass.
 * Group g = pairs group_off[g] .. group_off[g+1]-1 (device int32 array, at
 * most max_group pairs per group), one group = one reference batch:
 * after the first row where every pair of the group has best-so-far - row
 * best > x_drop, the group stops; each pair's result is then its first strict
 * maximum over the rows up to that one.  x_drop = floor of the reference's
 * x_drop, or -1 when the reference's x_drop <= 0 (no stop).  Groups above 64
 * pairs need scratch (4 int32 per pair; NULL is fine otherwise).  Writes
 * score, 1-based query_end / target_end (0 when the score is 0). */
int gmfa_sw_xdrop(const int32_t *rowkey, const int64_t *rowkey_off, const int32_t *qlen,
                  const int32_t *band_offset, const int32_t *group_off, int32_t n_groups,
                  int32_t max_group, int32_t x_drop, int32_t half_band, int32_t *scratch,
                  int32_t *score, int32_t *query_end, int32_t *target_end, void *stream);

/* Walk the traceback planes from (query_end, target_end) of every pair and
 * write BAM-style CIGAR runs (len << 4 | op, forward order) at
 * cigar[cigar_off[p]], their count n_cigar[p], the start (0-based
 * read_start / ref_start) and op tallies counts[4p..4p+3] = (=, X, I, D).
 * score/query_end/target_end are the (possibly x-drop adjusted) best cells. */
int gmfa_sw_traceback(const gmfa_sw_batch *batch, const gmfa_sw_params *params,
                      const int32_t *score, const int32_t *query_end,
                      const int32_t *target_end, const uint32_t *tb, const int64_t *tb_off,
                      const int64_t *cigar_off, uint32_t *cigar, int32_t *n_cigar,
                      int32_t *read_start, int32_t *ref_start, int32_t *counts,
                      void *stream);

/* Chaining DP over reads stored back to back (segment s = anchors
 * seg_off[s] .. seg_off[s+1]-1, already in DP order).  Float64 profile:
 * pen_lut[g] = penalty of gap g for g < max_gap (device array); bonus is an
 * optional (n_anchors, max_lookback) row-major graph bonus.
 * Requires 1 <= max_lookback <= 256 (chaining.py:154 has no cap; one
 * wavefront holds 64 predecessors per window, up to 4 windows). */
int gmfa_chain_dp_f64(const int64_t *read_end, const int64_t *ref_end, const double *weight,
                      const int64_t *seg_off, int32_t n_segments, int32_t max_lookback,
                      int64_t max_gap, const double *pen_lut, const double *bonus,
                      double *f, int64_t *parent, void *stream);

/* Integer profile: penalty(g) = g > 0 ? gap_open + gap_extend*g : 0. */
int gmfa_chain_dp_i32(const int64_t *read_end, const int64_t *ref_end, const int32_t *weight,
                      const int64_t *seg_off, int32_t n_segments, int32_t max_lookback,
                      int64_t max_gap, int32_t gap_open, int32_t gap_extend,
                      int32_t *f, int64_t *parent, void *stream);

/* Host: AffineChainer._backtrack over one segment (host pointers).
 * members[] receives DP-local indices chain by chain in read order,
 * chain_start[0..n] the boundaries, chain_score[c] the scores.
 * Returns the number of chains (>= 0) or a negative error. */
int64_t gmfa_chain_backtrack(int64_t n, const double *f, const int64_t *parent,
                             int64_t min_anchors, double min_score, int64_t *members,
                             int64_t *chain_start, double *chain_score);

/* BAM CIGAR records (graphmambaformer/data/formats.py:52, 343-382, with the
 * soft clips of extension.py:628-640): per record, S(read_start) + the runs +
 * S(read_len - read_end) merged, in the BAM binary encoding (len << 4 | op);
 * a mapped record (runs, ref_end > ref_start, read_len > 0) on strand -1 is
 * reversed to the forward strand with flag 0x10; if its query-consuming
 * length differs from read_len it becomes unmapped: flag |= 0x4, no CIGAR.
 * out[out_off[r]] needs room for n_cigar[r] + 2 runs. */
int gmfa_cigar_bam(int64_t n, const uint32_t *cigar, const int64_t *cigar_off,
                   const int32_t *n_cigar, const int32_t *read_start, const int32_t *read_end,
                   const int32_t *read_len, const int64_t *ref_start, const int64_t *ref_end,
                   const int32_t *strand, uint32_t *out, const int64_t *out_off, int32_t *out_n,
                   int32_t *flag, void *stream);

/* ---------------------------------------------------------------------------
 * Unit-cost wavefront alignment (graphmambaformer/alignment/extension.py:308-463
 * WavefrontAligner; accel/cuda_kernels.py:290-357 wfa_distance) for a batch
 * of pairs in the gmfa_sw_batch layout (band fields unused).
 * ------------------------------------------------------------------------- */

/* Edit distance of every pair, -1 when above max_distance (the reference's
 * RuntimeError).  cap >= min(max_distance, max over pairs of max(qlen, tlen));
 * scratch: 2 * (2 cap + 3) int32 per pair. */
int gmfa_wfa_distance(const gmfa_sw_batch *batch, int32_t max_distance, int32_t cap,
                      int32_t *scratch, int32_t *distance, void *stream);

/* CIGAR of every pair with distance >= 0: BAM runs (len << 4 | op) at
 * cigar[cigar_off[p]] (capacity qlen + tlen + 1), n_cigar[p] runs, counts
 * (=, X, I, D).  hist: (distance + 1)^2 int32 per pair with both sides
 * non-empty, at hist_off[p].  *status (zeroed by the caller) is set if a
 * traceback leaves the first wavefront (the reference would fail there). */
int gmfa_wfa_align(const gmfa_sw_batch *batch, const int32_t *distance, const int64_t *hist_off,
                   int32_t *hist, const int64_t *cigar_off, uint32_t *cigar, int32_t *n_cigar,
                   int32_t *counts, int32_t *status, void *stream);

/* ---------------------------------------------------------------------------
 * Chaining around the DP, on the device, for a batch of segments (a segment =
 * one (read, strand) anchor set; segment s = elements seg_off[s] ..
 * seg_off[s+1]-1, all arrays device memory).
 * ------------------------------------------------------------------------- */

/* AnchorSet.sorted_by_ref (graphmambaformer/alignment/types.py:149-152):
 * stable sort of each segment by (ref_end, read_end); the six arrays are
 * gathered into DP order.  scratch: 2n int32. */
int gmfa_chain_sort(const int64_t *seg_off, int32_t n_segments, int64_t n,
                    const int64_t *read_end, const int64_t *ref_end, const double *weight,
                    const int64_t *read_pos, const int64_t *ref_pos, const int64_t *anchor_id,
                    int32_t *scratch, int64_t *o_read_end, int64_t *o_ref_end, double *o_weight,
                    int64_t *o_read_pos, int64_t *o_ref_pos, int64_t *o_anchor_id, void *stream);

/* AffineChainer._backtrack (graphmambaformer/alignment/chaining.py:530-576)
 * per segment on the device (f / parent of gmfa_chain_dp_f64, DP order,
 * segment-local parents).  Chain records of segment s are stored from index
 * seg_off[s] on: n_chains[s] of them, chain_score, chain_span (read_start,
 * read_end, ref_start, ref_end), chain_members (absolute offset into members,
 * count); members hold anchor_id values in read order.  scratch: 2n int32,
 * claimed: n bytes; *bad_parent (zeroed by the caller) is set if a parent
 * leaves its segment. */
int gmfa_chain_backtrack_device(const int64_t *seg_off, int32_t n_segments, int64_t n,
                                const double *f, const int64_t *parent, const int64_t *read_pos,
                                const int64_t *read_end, const int64_t *ref_pos,
                                const int64_t *ref_end, const int64_t *anchor_id,
                                int64_t min_anchors, double min_score, int32_t *scratch,
                                uint8_t *claimed, int32_t *n_chains, double *chain_score,
                                int64_t *chain_span, int64_t *chain_members, int64_t *members,
                                int32_t *bad_parent, void *stream);

/* AffineChainer.select (chaining.py:579-617) per read: candidates are the
 * chains of segment read_seg[2r] (strand +1) then read_seg[2r+1] (strand -1;
 * -1 = no segment); kept[r * max_chains + x] = chain record indices, the
 * primary first, n_kept[r] of them.  scratch: at read_base[r] room for the
 * read's candidates (its anchor count suffices). */
int gmfa_chain_select(int32_t n_reads, const int32_t *read_seg, const int64_t *seg_off,
                      const int32_t *n_chains, const double *chain_score,
                      const int64_t *chain_span, double score_ratio, double max_overlap,
                      int32_t max_chains, int64_t *scratch, const int64_t *read_base,
                      int32_t *n_kept, int64_t *kept, void *stream);

/* ---------------------------------------------------------------------------
 * Stage-1 k-mer seeding (graphmambaformer/alignment/seeding.py:720-798
 * GPUKmerIndex; accel/cuda_kernels.py:44-74 kmer_lookup).  The table is the
 * reference's KmerTable (sorted unique packed k-mers, CSR positions) plus a
 * directory over the top dir_bits of the 2k-bit key: dir[b] = first key
 * index with key >> (2k - dir_bits) >= b, for b in [0, 2^dir_bits].
 * ------------------------------------------------------------------------- */
typedef struct gmfa_kmer_index {
    const int64_t *keys;      /* [n_keys] sorted unique packed k-mers (device) */
    const int64_t *offsets;   /* [n_keys + 1] CSR offsets into positions */
    const int64_t *positions; /* reference offsets */
    const int32_t *dir;       /* [2^dir_bits + 1] bucket starts */
    int64_t n_keys;
    int32_t k;                /* 1..31 */
    int32_t dir_bits;         /* 0..min(2k, 28) */
} gmfa_kmer_index;

/* Reads of one seeding batch: a 2-bit stream + optional N mask (the layout of
 * gmfa_sw_batch), per-read base offsets and lengths, k-mer position offsets
 * pos_off[r] = sum over r' < r of max(0, len - k + 1) (n_reads + 1 entries)
 * and 256-position tiles (tile t covers positions tile_start[t] .. +255 of
 * read tile_read[t]; every position of every read in exactly one tile). */
typedef struct gmfa_kmer_reads {
    const uint32_t *seq;
    const uint32_t *mask;     /* NULL: no N */
    int64_t n_words;
    const int64_t *read_off;
    const int32_t *read_len;
    const int64_t *pos_off;
    const int32_t *tile_read;
    const int32_t *tile_start;
    int32_t n_reads;
    int32_t n_tiles;
    int32_t window;           /* minimizer window, 1..128 */
    int32_t reserved;
} gmfa_kmer_reads;

/* (start, count) of each query k-mer in the table, (0, 0) if absent. */
int gmfa_kmer_lookup(const gmfa_kmer_index *index, const int64_t *queries, int64_t n,
                     int64_t *out_start, int32_t *out_count, void *stream);

/* Minimizer selection + lookup of every read k-mer: per position (start,
 * count) (count 0 unless selected, valid and present), pos_aoff = exclusive
 * scan of count within its read, read_total[r], and read_anchor_off[0..n_reads]
 * = exclusive scan of read_total (read_anchor_off[n_reads] = all anchors). */
int gmfa_kmer_seed(const gmfa_kmer_index *index, const gmfa_kmer_reads *reads,
                   int64_t *pos_start, int32_t *pos_count, int64_t *pos_aoff,
                   int64_t *read_total, int64_t *read_anchor_off, void *stream);

/* Anchors of every read at read_anchor_off[r]: read_pos (k-mer position in
 * the read) and ref_pos (reference offset), in position order and table order
 * within a position, as GPUKmerIndex.query returns them. */
int gmfa_kmer_emit(const gmfa_kmer_index *index, const gmfa_kmer_reads *reads,
                   const int64_t *pos_start, const int32_t *pos_count, const int64_t *pos_aoff,
                   const int64_t *read_anchor_off, int64_t *read_pos, int64_t *ref_pos,
                   void *stream);

/* ---------------------------------------------------------------------------
 * Continuous serving.  One class = requests sharing one band width and
 * scoring.  Each class runs one persistent forward kernel of n_slots
 * wavefronts on a stream of its own; a wavefront takes the next published
 * request, computes it, walks its CIGAR, and marks it complete.  The host
 * thread publishes request k of a class at arrival_s[k] (seconds after the
 * start of the run) and records complete_s[k] when it sees the completion.
 *
 * batch:  every request of the class, in arrival order (device memory, all
 *         written before the call); batch.order must be NULL.
 * params: flags GMFA_SW_TRACEBACK | GMFA_SW_CIGAR (+ GMFA_SW_NMASK), a tile
 *         with 64 lanes per pair (one pair per wavefront).
 * out:    per-request outputs as for gmfa_sw_forward with GMFA_SW_CIGAR;
 *         out.tb holds n_slots plane slabs of slab_words uint32 each
 *         (slab_words >= gmfa_sw_tb_bytes(batch.max_qlen) / 4); out.tb_off
 *         is unused; out.work_counter points to 64 bytes of 64-byte aligned
 *         device memory (the class's queue head and device-side mirrors of
 *         the control words, which idle wavefronts poll instead of host
 *         memory).
 * A run ends when every request completed or after timeout_s; a wavefront
 * never waits longer than timeout_s + 5 s for a request.  The control words
 * (queue tails and completion flags) are allocated in host-coherent memory
 * for the duration of the call.  One run at a time per device.
 * ------------------------------------------------------------------------- */
typedef struct gmfa_serve_class {
    gmfa_sw_batch batch;
    gmfa_sw_params params;
    gmfa_sw_out out;
    int64_t slab_words;
    int32_t n_slots;
    int32_t reserved;
    const double *arrival_s; /* host [batch.n_pairs], nondecreasing */
    double *complete_s;      /* host [batch.n_pairs], out (-1: not completed) */
    /* Paged traceback planes (page_blocks > 0; else per-slot slabs as above):
     * out.tb is a pool of n_pages pages of page_blocks (a power of two) 8-row
     * plane blocks, page_blocks * gmfa_sw_tile.lanes_stored *
     * gmfa_sw_tile.cells_per_lane words each.  A wavefront takes the pages of
     * its request before the forward pass and frees them after the walk:
     * requests are admitted by the planes' HBM footprint (waiting while the
     * pool is short).  page_owner: device int32 [n_pages].  The longest query
     * may need at most 256 pages.  page_wait_s bounds one request's wait for
     * pages (0: timeout_s + 5 s); past it the request completes with no
     * outputs (out.n_cigar = -1) and the report's waves_gave_up is set. */
    int32_t page_blocks;
    int32_t n_pages;
    int32_t *page_owner;
    double page_wait_s;
} gmfa_serve_class;

typedef struct gmfa_serve_report {
    double wall_s;          /* start of the run .. last completion seen */
    int64_t completed;      /* requests completed over all classes */
    int32_t timed_out;      /* 1: timeout_s ended the run before all completed */
    int32_t waves_gave_up;  /* 1: a wavefront's wait reached its deadline */
} gmfa_serve_report;

int gmfa_serve_run(const gmfa_serve_class *classes, int32_t n_classes, double timeout_s,
                   gmfa_serve_report *report);

/* ---------------------------------------------------------------------------
 * Online serving session: a front end submits requests one at a time while
 * the kernels run, and polls for completed ones (the gmfa_serve_submit /
 * gmfa_serve_poll pair).  The same per-band-class persistent kernels as
 * gmfa_serve_run, fed from a ring per class in host-coherent memory:
 * gmfa_serve_submit packs the query into the next free ring entry, writes the
 * request (reference window, band offset) and publishes it; a wavefront claims
 * it, copies the query into its device stage slot, aligns it, walks its CIGAR
 * and writes the outputs straight into the entry; gmfa_serve_poll returns the
 * completed entries and frees them.  A wavefront claims only published
 * requests, so waves that idle for idle_timeout_s exit (freeing the CUs) and
 * the class is relaunched by the next submit or poll that finds requests
 * pending and no wave left.  One host thread per session; sessions on one
 * device share its CUs.
 *
 * Device memory per class is the caller's: n_slots traceback slabs (tb,
 * slab_words >= gmfa_sw_tb_bytes(max_qlen) / 4 each), n_slots query stage
 * slots (stage, 2 * gmfa_serve_stage_words(max_qlen) words each) and the
 * 64-byte control block (head, zeroed by gmfa_serve_open).  The library allocates
 * the rings and control words in host-coherent memory for the session's life.
 * ------------------------------------------------------------------------- */
typedef struct gmfa_serve_session gmfa_serve_session;

typedef struct gmfa_serve_class_cfg {
    gmfa_sw_params params;  /* GMFA_SW_TRACEBACK | GMFA_SW_CIGAR (+ GMFA_SW_NMASK),
                               a tile with 64 lanes per pair */
    int32_t max_qlen;       /* longest query the class accepts, <= 2^20 */
    int32_t ring;           /* requests in flight (submitted, not yet polled): a power of two <= 2^20 */
    int32_t n_slots;        /* persistent wavefronts */
    int32_t cigar_cap;      /* CIGAR runs kept per request; a longer CIGAR is
                               reported with status GMFA_SERVE_TRUNCATED */
    uint32_t *tb;           /* device: n_slots slabs of slab_words */
    int64_t slab_words;
    uint32_t *stage;        /* device: n_slots * 2 * gmfa_serve_stage_words(max_qlen) */
    uint64_t *head;         /* device: 64 bytes, 64-byte aligned (ticket head and
                               device-side mirrors of the control words) */
    /* paged traceback planes, as in gmfa_serve_class (page_blocks 0: tb holds
     * n_slots slabs of slab_words) */
    int32_t page_blocks;
    int32_t n_pages;
    int32_t *page_owner;
    double page_wait_s;  /* bound of one wait for pages (0: the longer of the
                            idle bound and 60 s); past it: GMFA_SERVE_NO_PAGES */
} gmfa_serve_class_cfg;

#define GMFA_SERVE_TRUNCATED 1  /* n_cigar > cigar_cap: no runs returned */
#define GMFA_SERVE_NO_PAGES 2   /* no traceback pages within the wait bound: no outputs */

typedef struct gmfa_serve_result {
    int64_t ticket;         /* as returned by gmfa_serve_submit (per class, from 0) */
    int32_t cls;
    int32_t status;         /* 0 or GMFA_SERVE_TRUNCATED */
    int32_t score;
    int32_t query_end;      /* 1-based end row / column of the best cell */
    int32_t target_end;
    int32_t read_start;     /* 0-based start of the local alignment */
    int32_t ref_start;      /* within the request's reference window */
    int32_t n_cigar;        /* runs ((length << 4) | op, GMFA_OP_*) */
    int32_t counts[4];      /* bases under =, X, I, D */
    int64_t cigar_off;      /* first run in the caller's cigar buffer */
    int64_t claim_tick;     /* device wall clock when a wavefront took the request */
    int64_t done_tick;      /* ... and when its outputs were complete */
} gmfa_serve_result;

typedef struct gmfa_serve_stats {
    int64_t submitted;      /* requests published over all classes */
    int64_t returned;       /* requests returned by gmfa_serve_poll */
    int64_t in_flight;      /* submitted - returned */
    int64_t launches;       /* class kernel launches (1 per class + relaunches) */
    double ticks_per_s;     /* rate of the device wall clock of claim_tick / done_tick */
} gmfa_serve_stats;

/* Query words of one stage slot for queries up to max_qlen bases. */
int32_t gmfa_serve_stage_words(int32_t max_qlen);

/* ref_seq / ref_mask: the resident reference, 2-bit stream of ref_bases bases
 * (device; ref_mask needed by classes with GMFA_SW_NMASK).  idle_timeout_s in
 * (0, 3600]: how long a wavefront waits for a request before it exits. */
int gmfa_serve_open(const uint32_t *ref_seq, const uint32_t *ref_mask, int64_t ref_bases,
                    const gmfa_serve_class_cfg *classes, int32_t n_classes,
                    double idle_timeout_s, gmfa_serve_session **session);

/* Submit one request to class cls: the query as qlen host codes (two_bit: 0..3;
 * else A,C,G,T = 1..4 and N = 5, N only for GMFA_SW_NMASK classes), aligned
 * against reference bases [t_off, t_off + tlen) with the band centred on
 * diagonal band_offset.  *ticket receives the request's ticket.  Returns
 * GMFA_E_AGAIN when the class's ring is full (poll, then submit again). */
int gmfa_serve_submit(gmfa_serve_session *session, int32_t cls, const uint8_t *codes,
                      int32_t qlen, int32_t two_bit, int64_t t_off, int32_t tlen,
                      int32_t band_offset, int64_t *ticket);

/* Submit n requests in order (request k: class cls[k], query codes[code_off[k]
 * .. code_off[k] + qlen[k]), window t_off[k] / tlen[k], band_offset[k]), with
 * one publish per class for the whole burst.  Stops at the first request
 * that is invalid (GMFA_E_INVALID) or meets a full ring (GMFA_E_AGAIN); the
 * requests before it are published.  *n_submitted receives their count and
 * tickets[0 .. *n_submitted) their tickets. */
int gmfa_serve_submit_many(gmfa_serve_session *session, int32_t n, const int32_t *cls,
                           const uint8_t *codes, const int64_t *code_off, const int32_t *qlen,
                           int32_t two_bit, const int64_t *t_off, const int32_t *tlen,
                           const int32_t *band_offset, int64_t *tickets, int32_t *n_submitted);

/* Completed requests, at most max_results, their CIGAR runs packed into the
 * host buffer cigar (cigar_cap runs); each returned request frees its ring
 * entry.  Never blocks.  *n_results receives the count. */
int gmfa_serve_poll(gmfa_serve_session *session, gmfa_serve_result *results,
                    int32_t max_results, uint32_t *cigar, int64_t cigar_cap,
                    int32_t *n_results);

int gmfa_serve_session_stats(const gmfa_serve_session *session, gmfa_serve_stats *stats);

/* Close the queues, wait for every submitted request to finish (results not
 * yet polled are dropped) and free the session.  report->completed counts the
 * requests completed during the session. */
int gmfa_serve_close(gmfa_serve_session *session, gmfa_serve_report *report);

#ifdef __cplusplus
}
#endif
#endif /* GMFA_H */
