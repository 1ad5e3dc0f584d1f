/*
 * amethyst.h -- C ABI of the MI355X (gfx950) packed-column operator engine.
 *
 * Every entry point takes plain device pointers, sizes and a HIP stream
 * passed as `void*` (a hipStream_t; NULL = the default stream). No operator
 * allocates device memory; every operator is stream-ordered and asynchronous
 * and reentrant. The only global state is the fixed per-GPU scratch array of
 * am_compressstore_noalloc. Errors are returned as AM_* status codes; nothing
 * aborts the process.
 *
 * The reference (yogi-tud/Amethyst) is a CUDA benchmark whose operators are
 * templated kernels `kernel_<op>_<format><block_size, bits, [outbits], op>`
 * launched by per-bench host functions; its intended host wrappers are in
 * src/lib.cuh:4-46 (never compiled there). Each declaration below names the
 * reference kernel/wrapper it replaces.
 *
 * Column formats (reference layouts, src/common_code.cuh):
 *   AM_FMT_BASELINE  one uint64 per element, value = word & mask(bits)
 *                    (get_element_baseline, src/common_code.cuh:51-55)
 *   AM_FMT_SLAB      "element stuffing": floor(64/bits) elements per uint64,
 *                    element i in word i/eps at bit (i%eps)*bits, LSB-first,
 *                    top 64-eps*bits bits of each word dead (zero when we
 *                    write them) (get_element, src/common_code.cuh:43-50)
 *   AM_FMT_BITSTREAM "nogaps": dense LSB-first bitstream, element i starts at
 *                    bit i*bits (get_elem_nogaps, src/common_code.cuh:100-116)
 *   AM_FMT_TIGHT     smallest of u8/u16/u32/u64 holding `bits`, masked
 *                    (tight_type / sentinel, src/common_code.cuh:11-19,
 *                    src/slab_iterator.cuh:83-91)
 *   AM_FMT_U32       one uint32 per element (bits <= 32): the "uncompressed
 *                    uint32" column of the benchmark configurations
 * Every buffer is a whole number of uint64 words (am_packed_words); padding
 * and dead bits written by this library are zero.
 *
 * Bitmasks are LSB-first: element i is bit (i%64) of uint64 word i/64, which
 * on little-endian is bit (i%8) of byte i/8 -- the bit_iterator view
 * (src/bit_iterator.cuh:18-21) and the u32 mask words the reference filters
 * write (src/kernels/filter_slabs.cuh:55). Mask buffers hold ceil(n/64)
 * words (src/main.cu:234); bits past n are written as zero.
 */
#ifndef AMETHYST_H
#define AMETHYST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AM_ABI_VERSION 2

typedef enum {
    AM_OK = 0,
    AM_E_INVALID = 1,            /* bad argument (null pointer, width, format) */
    AM_E_UNSUPPORTED_CONFIG = 2, /* replaces the reference's -1 "skip" return */
    AM_E_HIP = 3,                /* a HIP runtime call failed */
    AM_E_WORKSPACE = 4,          /* workspace missing or too small */
    AM_E_CAPACITY = 5,           /* output buffer too small for the request */
    AM_E_ABORTED = 6,            /* a look-back wait exceeded its time limit; the
                                    workspace is poisoned until re-initialised */
    AM_E_SCRATCH_LIMIT = 7       /* input larger than am_compressstore_noalloc's
                                    fixed device scratch (AM_NOALLOC_MAX_ELEMENTS) */
} am_status;

typedef enum {
    AM_FMT_BASELINE = 0,
    AM_FMT_SLAB = 1,
    AM_FMT_BITSTREAM = 2,
    AM_FMT_TIGHT = 3,
    AM_FMT_U32 = 4
} am_format;

/* A column: `n` elements of `bits` (1..64) bits in `format` at `data`. */
typedef struct {
    void* data;
    uint64_t n;
    uint32_t bits;
    uint32_t format;
} am_column;

/* Filter predicates: element v is selected iff `v OP c`.
 * The reference predicate FILTER_PREDICATE_OP<bits> (src/config.cuh:107-113)
 * is {AM_PRED_GE, 1 << (bits-1)}. */
typedef enum {
    AM_PRED_LT = 0,
    AM_PRED_LE = 1,
    AM_PRED_GT = 2,
    AM_PRED_GE = 3,
    AM_PRED_EQ = 4,
    AM_PRED_NE = 5
} am_pred_op;

typedef struct {
    uint32_t op;
    uint32_t reserved;
    uint64_t c;
} am_predicate;

/* Binary operators, r = OP(l, r) & mask(out.bits) computed in uint64
 * (src/config.cuh:5-42). AND/OR/XOR/SUB are bitwise/wrapping extensions.
 * Division and modulo by zero follow the RISC-V convention: x/0 = all ones,
 * x%0 = x (the reference leaves them undefined). */
typedef enum {
    AM_OP_ADD = 0,          /* op_add          src/config.cuh:5-9   */
    AM_OP_MUL = 1,          /* op_mul          src/config.cuh:27-31 */
    AM_OP_MOD = 2,          /* op_mod          src/config.cuh:17-21 */
    AM_OP_DIV = 3,          /* op_div          src/config.cuh:22-26 */
    AM_OP_MATH = 4,         /* op_math         src/config.cuh:11-15 */
    AM_OP_BOOL_SIMPLE = 5,  /* op_bool_simple  src/config.cuh:32-36 (l || r) */
    AM_OP_BOOL_COMPLEX = 6, /* op_bool_complex src/config.cuh:38-42 (l xor r, logical) */
    AM_OP_AND = 7,
    AM_OP_OR = 8,
    AM_OP_XOR = 9,
    AM_OP_SUB = 10
} am_binop;

typedef enum {
    AM_GB_EXACT = 0, /* u64 SUM + u64 COUNT per key */
    AM_GB_REF = 1    /* SUM mod 2^value_bits, the reference's semantics
                        (src/kernels/groupby_sm_array.cuh:77) */
} am_groupby_mode;

/* ---------------------------------------------------------------- helpers */

int am_abi_version(void);
const char* am_status_string(int status);

/* uint64 words a column of n elements occupies in `format` at `bits`
 * (0 for an invalid combination). */
uint64_t am_packed_words(uint32_t format, uint32_t bits, uint64_t n);

/* Output-width rules op_add_outbits/op_mul_outbits/op_mod_outbits
 * (src/config.cuh:44-55): add -> bits+1, mul -> 2*bits, both capped at 64;
 * every other op keeps `bits`. */
uint32_t am_outbits(uint32_t op, uint32_t bits);

/* Smallest width that holds every value the predicate can select from a
 * column of `bits` bits (e.g. v < 1024 -> 10). */
uint32_t am_pred_width(am_predicate pred, uint32_t bits);

/* gen_bitmask (src/common_code.cuh:38-41), defined for bits in 1..64. */
uint64_t am_gen_bitmask(uint32_t bits);

/* Workspace used by the single-pass operators (compressstore, sum,
 * group-by). It must be zero-filled once before first use (am_workspace_init
 * or hipMemset); the kernels leave it ready for the next call. Calls that
 * share one workspace must be ordered on one stream. */
uint64_t am_workspace_bytes(uint64_t max_elements);
int am_workspace_init(void* ws, uint64_t ws_bytes, void* stream);

/* Copy `bytes` from device memory to host memory after the work queued on
 * `stream` before it, and wait for the copy (e.g. an operator's statistics).
 * Used by the C++ facade's count-returning wrappers. */
int am_read_device(void* host_dst, const void* dev_src, uint64_t bytes, void* stream);

/* Synchronous health check of a workspace: AM_OK, or AM_E_ABORTED when a
 * single-pass launch gave up waiting on a predecessor tile (every wait is
 * bounded at 2 s; such a launch writes statistics[0] = UINT64_MAX and later
 * launches on the poisoned workspace do nothing until am_workspace_init).
 * No reference counterpart: the reference's multi-launch operators do not
 * wait across blocks. */
int am_workspace_status(const void* ws);

/* ---------------------------------------------------------------- data generation */

/* Fill out[0..n) with the fast_prng stream (PCG32 XSH-RR seeded through
 * splitmix64, inc = 1; src/fast_prng.cuh:11-30) starting at draw `offset`,
 * computed in parallel with PCG jump-ahead: out[i] = draw (offset + i). */
int am_prng_fill_u32(uint64_t seed, uint64_t offset, uint32_t* out, uint64_t n, void* stream);

/* Uniform w-bit values v_i = draw(i) & mask(bits) packed into `out`
 * (bits > 32: v_i = (draw(2i) << 32 | draw(2i+1)) & mask). The reference
 * fills its inputs the same way on the host (src/main.cu:239-251). */
int am_gen_uniform(uint64_t seed, am_column* out, void* stream);

/* Inverse-CDF draws: v_i = first b with draw(i) < thresholds_dev[b]
 * (device array of nbins ascending uint32 thresholds, last = 2^32-1);
 * used for Zipf-distributed group-by keys. No reference counterpart. */
int am_gen_from_cdf(uint64_t seed, const uint32_t* thresholds_dev, uint32_t nbins, am_column* out, void* stream);

/* Ascending values with noise: v_i = min(mask, floor((i/n) * mask) +
 * (draw(i) & (2^noise_bits - 1))), i/n and the product each rounded in IEEE
 * double -- the "sorted with noise" columns of the uneven batch workload.
 * No reference counterpart. */
int am_gen_sorted_noise(uint64_t seed, uint32_t noise_bits, am_column* out, void* stream);

/* Bernoulli(selectivity) mask: bit i = draw(i) < round(s * 2^32). LSB-first
 * counterpart of generate_mask_uniform (src/main.cu:33-46), which writes
 * MSB-first bytes from a serial stream. */
int am_gen_mask_bernoulli(uint64_t seed, double selectivity, uint64_t* mask, uint64_t n, void* stream);

/* Clustered mask: alternating runs, 1-runs Geometric(mean mean_run),
 * 0-runs Geometric(mean mean_run*(1-s)/s), from a two-state Markov chain
 * restarted from its stationary distribution every 4096 elements.
 * Counterpart of create_bitmask's clustered masks (src/main.cu:59-108). */
int am_gen_mask_runs(uint64_t seed, double selectivity, double mean_run, uint64_t* mask, uint64_t n, void* stream);

/* The reference driver's own mask generators, byte for byte, on the GPU.
 * generate_mask_uniform (src/main.cu:33-46): bytes buf[offset ..
 * offset+length) from a generator seeded `seed` (42 in the reference), bit j
 * of each byte (7 down to 0: MSB first) set iff rand() <
 * (uint64)(selectivity * UINT32_MAX). create_bitmask (src/main.cu:59-108):
 * total/8 bytes with `cluster_count` runs of
 * (size_t)(selectivity*total)/cluster_count ones (product formed in float),
 * run i starting at bit i*(total/cluster_count), MSB first within a byte.
 * The reference's kernels read these bytes LSB-first, as the operators here
 * do, so passing the buffer as a mask reproduces its selections. */
int am_gen_mask_uniform_ref(uint64_t seed, double selectivity, uint8_t* buf, uint64_t offset, uint64_t length,
                            void* stream);
int am_gen_mask_clusters_ref(float selectivity, uint64_t cluster_count, uint64_t total, uint8_t* out, void* stream);

/* Copy nwords 64-bit words (16-byte aligned buffers) with 16-byte
 * non-temporal loads and stores: the device's sustained read+write rate,
 * against which the operators' HBM fractions are read. No reference
 * counterpart: measurement harness, beside the event timing of
 * CUDA_QUICKTIME (src/cuda_time.cuh:43-58). */
int am_stream_copy(const uint64_t* src, uint64_t* dst, uint64_t nwords, void* stream);

/* ---------------------------------------------------------------- launch-shape table */

/* Launch shapes the autotuner chooses (amethyst_amd/autotune.py sweeps them
 * on the GPU and writes amethyst_amd/tune_table.txt). An entry overrides the
 * built-in shape for one (knob, bits, in_format, out_format); bits is the
 * input width (the key width for group-by) and out_format is 0 where the
 * operator has no output column (filter, sum). Values outside a knob's
 * range are clamped to it. The table is process-wide host memory, read at
 * every launch; set it before launching work that should use it. Counterpart
 * of the reference's per-width grid x block sweep, which picks the best
 * launch shape per width at benchmark time (src/common_code.cuh:179-225,
 * src/utils.cuh:209-257). */
typedef enum {
    AM_KNOB_FILTER_WORDS = 0,    /* filter: 64-word chunks per wave per tile, 1..8 */
    AM_KNOB_COMPRESS_CHUNKS = 1, /* compaction (slab/container input): 64-word chunks per wave per tile, 1..8 */
    AM_KNOB_BINARY_WORDS = 2,    /* binary ops on slab/container words: output words per thread, 2, 4 or 8 */
    AM_KNOB_SUM_BLOCKS = 3,      /* sum: blocks of the partial-sum launch, 64..8192 */
    AM_KNOB_GROUPBY_BLOCKS = 4,  /* group-by over <= 12-bit keys: row blocks, 64..2048 */
    AM_KNOB_PIPELINE_TILE = 5,   /* batched chain: elements per lane per tile (8, 16 or 32) for variant 0 */
    AM_KNOB_COUNT = 6
} am_knob;

/* value 0 removes the entry. AM_E_INVALID for an unknown knob, width or format. */
int am_tune_set(uint32_t knob, uint32_t bits, uint32_t in_format, uint32_t out_format, uint32_t value);
/* The entry, or 0 when there is none (the built-in shape is used). */
uint32_t am_tune_get(uint32_t knob, uint32_t bits, uint32_t in_format, uint32_t out_format);
/* Remove every entry. */
void am_tune_clear(void);
/* Read entries from a text file: one "knob bits in_format out_format value"
 * per line, '#' starts a comment. Returns the number of entries read, or -1
 * if the file cannot be opened or a line is malformed (entries before it
 * stay set). */
int am_tune_load(const char* path);

/* ---------------------------------------------------------------- operators */

/* Convert a column between formats/widths: out[i] = in[i] & mask(out.bits).
 * Replaces the reference's host-side packing and its accessors
 * get_element / get_elem_nogaps / get_element_baseline
 * (src/common_code.cuh:43-55,100-116). out->n must equal in->n. */
int am_convert(const am_column* in, am_column* out, void* stream);

/* Predicate filter -> LSB-first bitmask (ceil(n/64) words).
 * Replaces kernel_filter_baseline / kernel_filter_nogaps / kernel_filter_slabs
 * (src/kernels/filter_baseline.cuh:9, filter_nogaps.cuh:4, filter_slabs.cuh:4)
 * and the wrapper filter_baseline (src/lib.cuh:43-46). */
int am_filter(const am_column* in, am_predicate pred, uint64_t* mask_out, void* stream);

/* Order-preserving compaction of `in` into `out` (any format, out->bits may
 * differ from in->bits). Selection comes from `mask` (LSB-first, ceil(n/64)
 * words) or, when mask == NULL, from `pred` evaluated on the fly (fused
 * filter + compressstore). If mask_out != NULL the selection bitmask is
 * written there too. out->n is the capacity in elements; writes beyond it
 * are dropped. stats_dev (device, 2 x uint64) receives
 * {selected count, number of selected values that did not fit out->bits};
 * values are stored masked to out->bits. The second count saturates at
 * 2^29 - 1 (a planner only needs to know that values did not fit).
 * One launch, decoupled look-back: replaces kernel_compresstore_slabs_chunk_popcount
 * + cub::DeviceScan::ExclusiveSum + kernel_compresstore_slabs_write
 * (src/kernels/compressstore_slabs.cuh:4-206,
 * src/benchs/bench_compressstore_slabs.cu:38-56) and the wrapper
 * compressstore_baseline_cub (src/lib.cuh:24-39). */
int am_compressstore(const am_column* in, const uint64_t* mask, const am_predicate* pred,
                     am_column* out, uint64_t* mask_out, uint64_t* stats_dev,
                     void* ws, uint64_t ws_bytes, void* stream);

/* am_compressstore without a workspace and without any allocation, for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc. No counterpart in the
 * reference. Same arguments, semantics and results as am_compressstore; its
 * look-back descriptors live in a fixed device array of
 * AM_NOALLOC_SCRATCH_BYTES bytes that the library holds on each GPU it runs
 * on until the process ends (looked up and zeroed on the first call on a
 * device, which must be the caller's current device). Inputs longer than
 * AM_NOALLOC_MAX_ELEMENTS return AM_E_SCRATCH_LIMIT. Every call on one GPU
 * shares that array: calls that overlap in time on one GPU (two streams, two
 * host threads) race and give wrong results -- order them on one stream.
 * Calls no hipMalloc/hipFree. */
#define AM_NOALLOC_MAX_ELEMENTS (1ull << 28)
#define AM_NOALLOC_SCRATCH_BYTES ((10ull << 20) + (256ull << 10))
int am_compressstore_noalloc(const am_column* in, const uint64_t* mask, const am_predicate* pred,
                             am_column* out, uint64_t* mask_out, uint64_t* stats_dev, void* stream);

/* Element-wise out[i] = op(a[i], b[i]) & mask(out.bits); a and b share n and
 * bits; any input format x any output format.
 * Replaces kernel_binary_op_baseline / _nogaps / _slabs
 * (src/kernels/binary_op_baseline.cuh:4, binary_op_nogaps.cuh:5,
 * binary_op_slabs.cuh:7) and the wrappers in src/lib.cuh:4-20. */
int am_binary(uint32_t op, const am_column* a, const am_column* b, am_column* out, void* stream);

/* Wrapping uint64 sum of all elements into *sum_dev (device pointer).
 * Replaces kernel_hop_slabs / kernel_hop_nogaps (src/kernels/hop.cuh:7,61)
 * with HOP_INIT = 0, HOP_OP = + (src/config.cuh:63-71). */
int am_sum(const am_column* in, uint64_t* sum_dev, void* ws, uint64_t ws_bytes, void* stream);

/* Workspace bytes am_groupby needs for n rows at these widths (0 for an
 * unsupported width): the fixed workspace for keys of <= 12 bits; keys of
 * 13..24 bits also need room for the partitioned rows (4 or 8 bytes each)
 * and a dense table of 2^key_bits sums and counts. */
uint64_t am_groupby_workspace_bytes(uint64_t n, uint32_t key_bits, uint32_t value_bits);

/* Group-by over keys (bits <= 24) with values (bits <= 32). Rows whose value
 * fails `value_pred` are skipped when value_pred != NULL (the chained
 * filter -> group-by). Output is in ascending key order:
 *   out_keys   : present keys, slab-packed at keys->bits
 *   out_sums   : per-key sums, slab-packed at result[1] bits
 *   out_counts : per-key counts, slab-packed at result[2] bits (may be NULL)
 * Each output buffer must hold am_packed_words(SLAB, bits, 2^keys->bits)
 * words for the widest possible bits (64). result_dev (device, 4 x uint64)
 * receives {group count, sum bits, count bits, rows aggregated}. In
 * AM_GB_EXACT the sums are exact uint64 and the widths are minimal
 * (>= 1); in AM_GB_REF sums are mod 2^values->bits and stored at that width.
 * Replaces kernel_groupby_sm_array_gather / _writeout
 * (src/kernels/groupby_sm_array.cuh:4-81), whose output order is
 * nondeterministic; ours is key-sorted. Keys of 13..24 bits (n < 2^32) take
 * a partitioned path: rows are split by the key's high bits and each
 * partition is aggregated in an LDS table over its low bits -- the
 * counterpart of the reference's sort-based group-by (cub::DeviceRadixSort::
 * SortPairs + cub::DeviceReduce::ReduceByKey, src/benchs/
 * bench_groupby_sort_cub.cu:50-63), with the same outputs as the small-key
 * path. */
int am_groupby(const am_column* keys, const am_column* values, const am_predicate* value_pred,
               uint32_t mode, uint64_t* out_keys, uint64_t* out_sums, uint64_t* out_counts,
               uint64_t* result_dev, void* ws, uint64_t ws_bytes, void* stream);

/* ---------------------------------------------------------------- batched operator chain */

/* One column of a batch: mask = filter(a, pred); s = compressstore(a, mask)
 * + compressstore(b, mask) (add, width bits+1); total = the exact sum of the
 * selected a+b (whatever s.bits is). `s` gives the output format (slab or
 * bitstream), width and capacity (s.n); values wider than s.bits are stored
 * masked and counted. stats (device, 4 x uint64) receives {selected count,
 * values that did not fit s.bits, total, 0}. a and b: same n, bits (<= 31)
 * and format (any format). variant: elements per lane per tile, 8, 16 or
 * 32 (2048-, 4096- or 8192-element tiles), or 0 for the launch-shape
 * table's choice (AM_KNOB_PIPELINE_TILE), else 32. */
typedef struct {
    am_column a;
    am_column b;
    am_column s;
    am_predicate pred;
    uint64_t* stats;
    uint32_t variant;
    uint32_t reserved;
} am_pipeline_col;

/* Workspace for a batch whose columns have lengths n[0..count). */
uint64_t am_pipeline_workspace_bytes(const uint64_t* n, uint32_t count);

/* Validate the batch and upload its job table into the workspace; waits for
 * `stream` first (a previous run on it may still read the table), then
 * copies synchronously. *total_tiles receives the grid size to pass to run.
 * The reference chains operators only by format, one launch each
 * (src/op_chaining_demo.cu:137-152); here the chain is fused. */
int am_pipeline_prepare(const am_pipeline_col* cols, uint32_t count, void* ws, uint64_t ws_bytes,
                        uint32_t* total_tiles, void* stream);

/* Run a prepared batch: ONE launch, one block per tile of every column
 * (tiles numbered column by column), no host work, graph-capturable. The
 * kernel checks the workspace's recorded batch against ws_bytes and
 * total_tiles; on a mismatch it only writes stats[0] = UINT64_MAX for every
 * column. A workspace that was zero-filled again must be prepared again. */
int am_pipeline_run(void* ws, uint64_t ws_bytes, uint32_t total_tiles, void* stream);

/* ---------------------------------------------------------------- learned planner */

/* 32 fp32 features per column of a batch (slab or bitstream inputs; width, length, sampled
 * selectivity, 16-bin histogram, sampled max of selected a+b, sortedness,
 * selection run length, byte estimates) into feat_dev[count][32]. scratch_dev
 * holds am_plan_scratch_bytes(count) bytes. No reference counterpart. */
uint64_t am_plan_scratch_bytes(uint32_t count);
int am_plan_features(const am_pipeline_col* cols, uint32_t count, float* feat_dev, void* scratch_dev, void* stream);

/* MLP inference (32 -> 64 -> 64 -> nplans, ReLU, fp32) over normalised
 * features; plan_dev[c] = argmin of the predicted relative log-times, or
 * fallback_dev[c] when the best beats the runner-up by less than min_margin
 * (fallback_dev may be NULL). score_dev (may be NULL) receives the
 * predictions. Weights: W1[64][32] b1[64] W2[64][64] b2[64] W3[P][64] b3[P]. */
int am_plan_infer(const float* feat_dev, uint32_t count, const float* weights_dev, uint32_t nplans,
                  const float* fmean_dev, const float* fstd_dev, float min_margin, const int32_t* fallback_dev,
                  int32_t* plan_dev, float* score_dev, void* stream);

/* ---------------------------------------------------------------- packed matmul */

/* One product C = A * B of unsigned `bits`-bit matrices (bits <= 7), A M x K
 * and B K x N, both row-major with each row padded to whole slabs -- the
 * layout of get_matrix_element (src/common_code.cuh:124-129). C_packed
 * (may be NULL) receives C as a row-padded slab matrix at out_bits;
 * C_i32 (may be NULL) receives it as row-major int32. The reference has
 * only the layout helpers (src/common_code.cuh:65-86,118-134); its matrix
 * kernels were never released (README.md:20). */
typedef struct {
    const uint64_t* A;
    const uint64_t* B;
    uint64_t* C_packed;
    int32_t* C_i32;
    uint32_t M, N, K;
    uint32_t out_bits; /* 0: am_matmul_out_bits(bits, K) */
} am_matmul_desc;

/* Width of the largest possible entry, K * (2^bits - 1)^2. */
uint32_t am_matmul_out_bits(uint32_t bits, uint32_t K);

/* Workspace for a batch of `count` products with the given shapes. */
uint64_t am_matmul_workspace_bytes(const uint32_t* M, const uint32_t* N, const uint32_t* K, uint32_t count);

/* Exact int32 products on the int8 matrix cores (K * (2^bits-1)^2 < 2^31). */
int am_matmul_packed(uint32_t bits, const uint64_t* A, const uint64_t* B, uint32_t M, uint32_t N, uint32_t K,
                     uint32_t out_bits, uint64_t* C_packed, int32_t* C_i32, void* ws, uint64_t ws_bytes, void* stream);

/* A batch of uneven products (descs in host memory) with one launch per
 * phase for the whole batch. The per-call header (job list, unpack and pack
 * tables) is built on the host and copied into the workspace with a
 * stream-ordered copy from pageable memory, so this call cannot be captured
 * in a graph: use a plan for that. */
int am_matmul_packed_batched(uint32_t bits, const am_matmul_desc* descs, uint32_t count, void* ws, uint64_t ws_bytes,
                             void* stream);

/* A prepared batch: create validates the products, lays out `ws` and copies
 * the header into it (stream-ordered on `stream`); run launches the kernels
 * only (no host work, graph-capturable) and may be called any number of
 * times while ws and the matrices stay in place; destroy frees the host
 * record (no device memory is owned by a plan). */
typedef struct am_matmul_plan am_matmul_plan;
int am_matmul_plan_create(uint32_t bits, const am_matmul_desc* descs, uint32_t count, void* ws, uint64_t ws_bytes,
                          am_matmul_plan** plan, void* stream);
int am_matmul_plan_run(const am_matmul_plan* plan, void* stream);
void am_matmul_plan_destroy(am_matmul_plan* plan);

#ifdef __cplusplus
}
#endif

#endif /* AMETHYST_H */
