/*
 * gsw_hip.h -- C ABI of libgsw_hip.so, the MI355X (gfx950) affine-gap (Gotoh)
 * alignment-scoring library.
 *
 * Plain C: no torch or HIP types cross this boundary.  Device buffers are opaque
 * `void*` obtained from gsw_malloc; host buffers are plain pointers.  Every function
 * returns GSW_OK (0) or an error code; gsw_last_error() gives a message for the
 * calling thread.  Every scoring call is synchronous: when it returns, its outputs
 * are complete.
 *
 * Reference interfaces replaced (freiburgermsu/gpuSW, gpusw 0.2.0):
 *   - Backend contract      src/gpusw/backends/__init__.py:31-87
 *   - CudaBackend           src/gpusw/backends/cuda.py:29-106
 *   - module/SUB cache      src/gpusw/_compile.py:39-93
 *   - kernels               src/gpusw/kernel.py:56-190 (sw_score, sw_cross, sw_pairs)
 */
#ifndef GSW_HIP_H
#define GSW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define GSW_OK            0
#define GSW_ERR_NO_DEVICE 1  /* no AMD GPU / HIP runtime: NoGpuError            */
#define GSW_ERR_INVALID   2  /* bad argument: ValueError / GpuSWError           */
#define GSW_ERR_HIP       3  /* HIP runtime failure: GpuSWError                 */
#define GSW_ERR_LIMIT     4  /* size limit exceeded: GpuSWError                 */
#define GSW_ERR_OVERFLOW  5  /* score range exceeds int32: OverflowRiskError    */
#define GSW_ERR_BUSY      6  /* server admission refused (memory budget)        */
#define GSW_ERR_TIMEOUT   7  /* server request not finished within the timeout  */

/* ---- scheme ------------------------------------------------------------- */
#define GSW_MODE_LOCAL      0
#define GSW_MODE_GLOBAL     1
#define GSW_MODE_SEMIGLOBAL 2

#define GSW_DTYPE_AUTO  0   /* int16 fast path, int32 where the range needs it   */
#define GSW_DTYPE_INT16 1   /* caller checked the envelope (Scheme.resolve_dtype) */
#define GSW_DTYPE_INT32 2   /* 32-bit DP everywhere                               */

/* Scoring scheme passed by value.  Replaces the compile-time #defines of
 * src/gpusw/kernel.py:44-54 (MODE, GOPEN, GEXT, FREE_Q, FREE_R, NS) and the SUB
 * upload of src/gpusw/_compile.py:83-93.  sub[a*ns+b] scores reference code a
 * against query code b (src/gpusw/scheme.py:149-162). */
typedef struct gsw_scheme {
    int32_t mode;        /* GSW_MODE_*                                  */
    int32_t gap_open;    /* score of a gap's first residue, <= 0        */
    int32_t gap_extend;  /* score of each further gap residue, <= 0     */
    int32_t free_q;      /* semiglobal: query end gaps free             */
    int32_t free_r;      /* semiglobal: reference end gaps free         */
    int32_t ns;          /* side of the sentinel-padded table (<= 64)   */
    const int32_t *sub;  /* host pointer, ns*ns int32                   */
} gsw_scheme;

/* Per-call measurements (kernel time from HIP events on the library stream). */
typedef struct gsw_stats {
    double  kernel_ms;    /* summed device time of the scoring kernels          */
    double  wall_ms;      /* host wall time of the whole call                    */
    int64_t cells;        /* sum over scored pairs of len(query) * len(ref)      */
    int64_t pairs;        /* pairs scored                                        */
    int64_t rescored;     /* int16 pairs that saturated and were rescored in int32 */
    int64_t plain_pairs;  /* pairs routed to the one-thread-per-pair kernel      */
    int32_t launches;     /* kernel launches issued                              */
    int32_t variant;      /* dominant batched variant id (-1: none)              */
    double  top_ms;       /* cross: device time of the batched launch with the   */
                          /*   most cells (events before and after it on its    */
                          /*   stream; includes time shared with launches of    */
                          /*   the other stream);                               */
                          /* pairs: device time of the dominant tile's full     */
                          /*   batches (first start to last end) per batch      */
    int64_t top_cells;    /*   cells of that launch (pairs: mean per batch)      */
    int32_t top_variant;  /*   its variant id (-1: none)                         */
    int32_t top_striped;  /*   kernel: 0 cross, 1 stripe, 2 pairs, 3 pairs table, 4 byte-select table */
} gsw_stats;

/* ---- runtime (src/gpusw/_compile.py:30-46; backends/cuda.py:29-41) ----- */
const char *gsw_last_error(void);
const char *gsw_version(void);
int gsw_device_count(int32_t *count);
int gsw_set_device(int32_t device);            /* Backend.device_scope            */
int gsw_get_device(int32_t *device);
int gsw_device_name(int32_t device, char *buf, int32_t buflen);
int gsw_malloc(void **dptr, size_t bytes);     /* Backend.upload (allocation)     */
int gsw_free(void *dptr);
int gsw_memcpy_h2d(void *dst, const void *src, size_t bytes);   /* Backend.upload   */
int gsw_memcpy_d2h(void *dst, const void *src, size_t bytes);   /* Backend.as_host  */
int gsw_memset(void *dptr, int32_t value, size_t bytes);
int gsw_memcpy_d2d(void *dst, const void *src, size_t bytes);  /* device copy (profiling calibration) */
int gsw_synchronize(void);

/* ---- plain comparator: one thread per pair ------------------------------
 * Replaces CudaBackend.run_cross / run_pairs (src/gpusw/backends/cuda.py:47-85)
 * launching sw_cross / sw_pairs (src/gpusw/kernel.py:160-190).  dtype is the
 * resolved store type (GSW_DTYPE_INT16 or GSW_DTYPE_INT32); threads is the block
 * size.  Scores are written as int32:
 *   cross: d_out[a*nr + b] = score(query h_qlist[a], reference b), a < nq, b < nr
 *   pairs: d_out[k]        = score(query h_qi[k], reference h_rj[k])            */
int gsw_score_cross_plain(const gsw_scheme *s, int32_t dtype, int32_t threads,
                          const uint8_t *d_qcodes, const int32_t *d_qoff,
                          const uint8_t *d_rcodes, const int32_t *d_roff,
                          const int32_t *h_qlist, int32_t nq, int32_t nr, int32_t maxq,
                          int32_t *d_out, gsw_stats *stats);
int gsw_score_pairs_plain(const gsw_scheme *s, int32_t dtype, int32_t threads,
                          const uint8_t *d_qcodes, const int32_t *d_qoff,
                          const uint8_t *d_rcodes, const int32_t *d_roff,
                          const int32_t *h_qi, const int32_t *h_rj, int64_t n, int32_t maxq,
                          int32_t *d_out, gsw_stats *stats);

/* ---- indexed sequence sets ------------------------------------------------
 * A set of n sequences already resident on the device (codes + n+1 offsets, as
 * produced by the reference's encode.funnel, src/gpusw/encode.py:164-220).  The
 * set borrows the device buffers (caller keeps them alive, codes unchanged: the set
 * records their largest value) and copies the host offsets to build its length index.  Replaces the Aligner's per-set device state
 * (src/gpusw/aligner.py:69-87). */
typedef struct gsw_seqset_s *gsw_seqset;
int gsw_seqset_create(const uint8_t *d_codes, const int32_t *d_off, const int32_t *h_off,
                      int32_t n, gsw_seqset *out);
int gsw_seqset_destroy(gsw_seqset set);

/* ---- batched wavefront path (backend="hip") ------------------------------
 * Same contract and outputs as the plain calls above (int32 scores, row-major
 * (nq, nr) for cross), computed by the wavefront kernels: two alignments per lane on
 * half2 tiles (binary16, exact for the integers in [-2048, 2048]; local results of 2048
 * or more rescored in int32) or packed int16 tiles, int32 tiles where the score envelope
 * needs them.  Queries wider than every tile run as stripes; explicit pairs longer than
 * every tile on both sides run on the int32 stripe kernels.  dtype_policy is
 * GSW_DTYPE_AUTO/INT16/INT32 (src/gpusw/scheme.py:116-141 decides it on the host; AUTO
 * lets the library pick per pair). */
int gsw_score_cross(const gsw_scheme *s, int32_t dtype_policy,
                    gsw_seqset queries, const int32_t *h_qlist, int32_t nq,
                    gsw_seqset refs, int32_t *d_out, gsw_stats *stats);
int gsw_score_pairs(const gsw_scheme *s, int32_t dtype_policy,
                    gsw_seqset queries, gsw_seqset refs,
                    const int32_t *h_qi, const int32_t *h_rj, int64_t n,
                    int32_t *d_out, gsw_stats *stats);

/* ---- explicit pairs without device allocation -------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * scores as gsw_score_pairs, computed by the same table kernels, with every buffer beyond
 * the inputs and outputs in a file-scope device array of about 18 MiB per GPU (held until
 * the process ends).  The call allocates nothing: no hipMalloc/hipFree, no streams, events or
 * scheme cache entries; it runs on the null stream and returns when its kernels are done.
 * Limits: at most 262,144 pairs per call (GSW_ERR_LIMIT beyond, checked before any device
 * work) and queries of at most 2,048 columns (GSW_ERR_LIMIT otherwise).  Every call on a GPU
 * shares that GPU's array, which keeps the last call's contents; the library serialises calls
 * on one GPU with a lock, so calls that overlap on one GPU wait for each other instead of
 * racing.  stats->kernel_ms is the call's wall time (no events are created). */
int gsw_score_pairs_noalloc(const gsw_scheme *s, int32_t dtype_policy,
                            gsw_seqset queries, gsw_seqset refs,
                            const int32_t *h_qi, const int32_t *h_rj, int64_t n,
                            int32_t *d_out, gsw_stats *stats);

/* ---- device top-k ---------------------------------------------------------
 * Per row of a device int32 (nrows, ncols) score matrix, the k best columns,
 * best first, ties broken toward the lower column index.  Replaces
 * CudaBackend.topk (src/gpusw/backends/cuda.py:93-106) whose tie order is
 * unspecified; the order here is that of the dense result's stable sort
 * (src/gpusw/result.py:37).  Host outputs h_idx/h_sc are (nrows, k). */
int gsw_topk_rows(const int32_t *d_scores, int32_t nrows, int32_t ncols, int32_t k,
                  int32_t *h_idx, int32_t *h_sc, gsw_stats *stats);

/* ---- tile selection ---------------------------------------------------------
 * No counterpart in the reference (its kernel has no tile parameters).  The batched
 * kernels are compiled for a set of (G lanes per reference, R query columns per lane)
 * tiles; a cost model picks one per query length.  The autotuner measures the tiles
 * and installs rules that take precedence: n rules of 6 int32
 * (lq_lo, lq_hi, packed_int16, local_mode, G, R); n = 0 clears them.  Forcing a tile
 * (G, R) applies it to every query it covers ((0, 0) restores automatic choice). */
int gsw_set_tile_table(const int32_t *rules, int32_t n);
int gsw_force_variant(int32_t G, int32_t R);
int gsw_list_variants(int32_t *gr, int32_t cap, int32_t *n);   /* (G, R) pairs */
/* The tile of a variant id as reported in gsw_stats (variant, top_variant): info[0..3] =
 * G, R, element kind (0 int32, 1 int16x2, 2 half2: binary16 pairs, exact for integers in
 * [-2048, 2048]), local mode (1) or global/semiglobal (0). */
int gsw_variant_info(int32_t id, int32_t *info);
/* Tiles that also run queries wider than G*R as stripes (forcing one of them stripes
 * every query it does not cover). */
int gsw_list_stripe_variants(int32_t *gr, int32_t cap, int32_t *n);

/* ---- traceback ------------------------------------------------------------
 * No counterpart in the reference, which scores only and re-aligns its top hits on the
 * CPU (docs/DESIGN.md:94-96, examples/legacy_16s_gpu_align.py:280-292).  For each pair
 * (h_qi[k], h_rj[k]) the alignment behind its score, in int32 (16 or 32 lanes fill a
 * pair's matrix for queries up to 2,048 columns, one thread beyond; one thread walks it):
 * coordinates are 0-based half-open; free end gaps (semiglobal) are left out of the CIGAR
 * and show in the coordinates; h_ops receives the CIGAR runs of pair k at
 * h_ops[out[k].ops_offset ..+ n_ops), each (length << 2) | op with op 0 '=' (identical
 * residues), 1 'X', 2 'I' (query residue against a gap), 3 'D' (reference residue against
 * a gap).  ops_cap must be >= sum over pairs of (lq + lr + 1).  Ties: H prefers zero
 * (local start), then the diagonal, then E, then F; E and F prefer opening; the end cell
 * is the first maximum in row-major order (local) or the corner, then the last column,
 * then the last row (semiglobal). */
typedef struct gsw_aln {
    int32_t score;
    int32_t q_begin, q_end, r_begin, r_end;
    int32_t matches, mismatches, ins, del, gap_opens;
    int32_t n_ops;
    int64_t ops_offset;
} gsw_aln;
int gsw_traceback_pairs(const gsw_scheme *s, gsw_seqset queries, gsw_seqset refs, const int32_t *h_qi,
                        const int32_t *h_rj, int64_t n, gsw_aln *h_out, uint32_t *h_ops, int64_t ops_cap,
                        gsw_stats *stats);

/* ---- host FASTA funnel ------------------------------------------------------
 * Replaces the funnel's FASTA reader (src/gpusw/encode.py:91-121, 164-220) for ASCII
 * files: same records (lines split at the ASCII breaks of str.splitlines, stripped of
 * str.strip whitespace, blank lines skipped, text before the first '>' dropped, id = the
 * first word of the header).  gsw_fasta_scan sizes the output (GSW_ERR_INVALID on a byte
 * >= 0x80); gsw_fasta_encode maps residues through lut[256] into codes[nres], writes
 * offsets[nrec+1] and id_span[2*nrec] = (start, length) of each id in buf (length 0: the
 * header has no word). */
int gsw_fasta_scan(const uint8_t *buf, int64_t n, int64_t *nrec, int64_t *nres);
int gsw_fasta_encode(const uint8_t *buf, int64_t n, const uint8_t *lut, uint8_t *codes,
                     int64_t *offsets, int64_t *id_span);

/* ---- serving -------------------------------------------------------------
 * No counterpart in the reference (its Aligner.top_k, src/gpusw/aligner.py:169-195,
 * is a synchronous batch loop).  A server keeps reference databases resident, admits
 * top-k requests by a device-memory budget (FIFO; a request larger than the whole
 * budget is rejected with GSW_ERR_BUSY), batches the work of all admitted requests of
 * a database into shared persistent-kernel launches, and returns per request the same
 * top-k as Aligner.top_k (score descending, ties to the lower reference index). */
typedef struct gsw_server_s *gsw_server;

typedef struct gsw_request_info {
    double  submit_ms, admit_ms, done_ms;   /* since server creation                 */
    double  latency_ms;                     /* done - submit                         */
    int64_t cells;                          /* sum(query lengths) * sum(ref lengths) */
    int32_t nq, k, launches, status;        /* k actually returned (min(k, n_refs))  */
} gsw_request_info;

typedef struct gsw_server_status {
    int64_t submitted, completed, rejected, launches, cells, rescored;
    int32_t queued, active;
    size_t  memory_budget, memory_used;
    double  uptime_ms;
} gsw_server_status;

/* launch_cells: target cell updates per batched launch (<= 0: default 2e11). */
int gsw_server_create(size_t mem_budget_bytes, double launch_cells, gsw_server *out);
/* The database's scheme is fixed at registration; refs must outlive the server. */
int gsw_server_add_db(gsw_server srv, const gsw_scheme *s, gsw_seqset refs, int32_t *db_id);
/* Copies the request's queries (codes indexed by h_off[0..nq]); returns immediately. */
int gsw_server_submit(gsw_server srv, int32_t db_id, const uint8_t *h_codes, const int32_t *h_off,
                      int32_t nq, int32_t k, int64_t *req_id);
/* Waits (timeout_ms < 0: forever) and copies (nq, info.k) results; consumes the id. */
int gsw_server_wait(gsw_server srv, int64_t req_id, int32_t timeout_ms, int32_t *h_idx,
                    int32_t *h_sc, gsw_request_info *info);
int gsw_server_stats(gsw_server srv, gsw_server_status *st);
/* Finishes all submitted requests, then frees the server. */
int gsw_server_destroy(gsw_server srv);

#ifdef __cplusplus
}
#endif
#endif /* GSW_HIP_H */
