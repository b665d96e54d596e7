/*
 * poyg.h -- C ABI of the MI355X (gfx950) direct-optimization engine.
 *
 * The engine replaces the C stubs behind POY's OCaml `external` declarations
 * for the alignment hot path.  Each per-call entry point below names the
 * reference stub it replaces (paths relative to the lealanko/poyd tree); the
 * OCaml-side stub that forwards a `value` to these functions is shown in
 * INTEGRATION.md.  Sequences are arrays of 5-bit masks (A=1, C=2, G=4, T=8,
 * gap=16) that include POY's leading gap at index 0.  Output sequences are
 * written the way POY's seq_prepend fills a caller-allocated sequence: the
 * produced symbols occupy out[cap-len .. cap-1] of a buffer of capacity cap.
 *
 * All functions return POYG_OK (0) or a negative POYG_E* code; the message of
 * the last error on the calling thread is available from poyg_last_error().
 * The library never longjmps; a stub raises OCaml's Failure on a non-zero
 * code.  A context is used by one host thread at a time (POY enters C from a
 * single thread, src/poyd/PoydThread.ml:40-77).
 *
 * Batched entry points take device pointers and a hipStream_t (passed as
 * void*) so that inputs stay resident in HBM; per-call entry points take host
 * pointers and synchronise.
 */
#ifndef POYG_H
#define POYG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define POYG_OK 0
#define POYG_EINVAL (-1)     /* bad argument (length, alphabet, NULL) */
#define POYG_ENOMEM (-2)     /* host or device allocation failed */
#define POYG_EDEVICE (-3)    /* HIP runtime error */
#define POYG_EUNSUPPORTED (-4)
#define POYG_EBAND (-5)      /* a backtrace left the filled band */

typedef struct poyg_ctx poyg_ctx;
typedef struct poyg_cm poyg_cm;
typedef struct poyg_cm3 poyg_cm3;
typedef struct poyg_plan poyg_plan;

const char *poyg_last_error(void);
const char *poyg_version(void);

/* ---- context -------------------------------------------------------- */
int poyg_ctx_create(int device, poyg_ctx **out);
void poyg_ctx_destroy(poyg_ctx *ctx);
/* Waits for all work queued on the context's device (every stream).  Device
 * blocks released since the previous synchronization (poyg_dev_free, plan
 * destruction) become reusable by later allocations of this library. */
int poyg_ctx_sync(poyg_ctx *ctx);

/* Launch-shape knobs of the plans made on this context afterwards (no
 * reference counterpart; results do not depend on them): "aff_quad_max_lanes"
 * (41..48, default 48: affine bands of at most this many diagonal pairs go
 * to the four-pairs-per-wave fill, wider ones to the one-pair fill),
 * "aff_side_div" (2..64, default 4: the side streams get at most
 * CUs / value compute units), "lin_sys_blocks_per_cu" and
 * "lin_trace_blocks_per_cu" (1..occupancy: resident workgroups per CU of the
 * linear fill and backtrace), "cube_waves_per_cu" (1..64, default 8),
 * "lin_cost_kernel" (1: the 32-bit systolic linear fill for every pair it
 * takes; 2, the default: the packed 16-bit fill where a pair's costs allow
 * it, one wave per pair, or four waves per pair when the batch has at most
 * two such pairs per CU; 3 and 4: the packed fill with one, four waves per
 * pair whatever the batch size), "aff_narrow_kernel" (1: the 32-bit four-pair affine fill for the
 * narrow bands; 2, the default: the packed eight-pair fill).  A negative
 * value restores the default.  poyd_amd/autotune.py picks them by timed
 * trials. */
int poyg_ctx_set_tuning(poyg_ctx *ctx, const char *knob, int value);
int poyg_ctx_get_tuning(poyg_ctx *ctx, const char *knob);
/* "<device name>:<gfx arch>:<CUs>cu", the key of tuned settings. */
int poyg_ctx_device_name(poyg_ctx *ctx, char *buf, int cap);

/* ---- device buffers (plumbing for callers without their own allocator;
 * no reference counterpart) ------------------------------------------- */
int poyg_dev_alloc(poyg_ctx *ctx, int64_t bytes, void **out);
/* Returns a block from poyg_dev_alloc.  Blocks go to a per-device cache
 * (up to 64 MiB in power-of-two classes, at most 1 GiB of them; larger ones
 * in classes a quarter power of two apart, up to 60% of the device's
 * memory) and are handed out again only after the next device-wide
 * synchronization (poyg_ctx_sync). */
int poyg_dev_free(poyg_ctx *ctx, void *p);
int poyg_dev_upload(poyg_ctx *ctx, void *dst, const void *src, int64_t bytes);
int poyg_dev_download(poyg_ctx *ctx, void *dst, const void *src, int64_t bytes);
/* Zeroes `bytes` (a multiple of 4) at the 4-byte aligned device address p
 * with a kernel on `stream` (NULL: the context's stream); usable inside a
 * graph capture, where a memset node is not. */
int poyg_dev_zero(poyg_ctx *ctx, void *p, int64_t bytes, void *stream);
/* The context's stream as a hipStream_t, and its device. */
void *poyg_ctx_stream(poyg_ctx *ctx);
int poyg_ctx_device(poyg_ctx *ctx);

/* ---- recorded passes (no reference counterpart) ----------------------
 * A pass of fixed shape over fixed device buffers -- a tree's Fitch down
 * and up passes, or the run of one plan -- recorded once as a HIP graph and
 * replayed.  poyg_graph_begin starts a capture on a stream of the context
 * and returns it; every call made with that stream until poyg_graph_end
 * (poyg_fitch_down / _up, poyg_plan_run, poyg_dev_zero) is recorded, not
 * run.  poyg_graph_launch replays the whole pass on `stream` (NULL: the
 * context's stream); replays of one graph on one stream run in order, and
 * a caller that reads a plan's status after a replay synchronizes that
 * stream first.  The buffers the recorded calls named must stay allocated
 * while the graph is used.  One capture at a time per context. */
typedef struct poyg_graph poyg_graph;
int poyg_graph_begin(poyg_ctx *ctx, void **capture_stream);
int poyg_graph_end(poyg_ctx *ctx, poyg_graph **out);
int poyg_graph_launch(poyg_graph *g, void *stream);
void poyg_graph_destroy(poyg_graph *g);

/* ---- two-dimensional cost matrix (src/cm.c, src/cost_matrix.ml) ------- */
/* cm_CAML_create (src/cm.c:1389-1398); only 5-bit combination alphabets. */
int poyg_cm_create(int a_sz, int combinations, int cost_model, int gap_open, int all_elements,
                   poyg_cm **out);
void poyg_cm_destroy(poyg_cm *cm);
/* cm_CAML_set_cost / set_worst / set_median / set_prepend / set_tail
 * (src/cm.c:1295-1345). */
int poyg_cm_set_cost(poyg_cm *cm, int a, int b, int v);
int poyg_cm_set_worst(poyg_cm *cm, int a, int b, int v);
int poyg_cm_set_median(poyg_cm *cm, int a, int b, int v);
int poyg_cm_set_prepend(poyg_cm *cm, int a, int v);
int poyg_cm_set_tail(poyg_cm *cm, int a, int v);
/* cm_CAML_set_affine (src/cm.c:123-128 via Cost_matrix.Two_D.c_set_aff). */
int poyg_cm_set_affine_flag(poyg_cm *cm, int cost_model, int gap_open);
/* cm_CAML_get_cost / get_median / get_worst / get_prepend / get_tail /
 * get_gap / get_lcm / get_affine / get_gap_opening. */
int poyg_cm_get_cost(const poyg_cm *cm, int a, int b);
int poyg_cm_get_median(const poyg_cm *cm, int a, int b);
int poyg_cm_get_worst(const poyg_cm *cm, int a, int b);
int poyg_cm_get_prepend(const poyg_cm *cm, int a);
int poyg_cm_get_tail(const poyg_cm *cm, int a);
int poyg_cm_get_gap(const poyg_cm *cm);
int poyg_cm_get_affine(const poyg_cm *cm);
int poyg_cm_get_gap_opening(const poyg_cm *cm);
/* Cost_matrix.Two_D.of_list ~use_comb:true (src/cost_matrix.ml:520-538,
 * 596-602): fills every ambiguity combination from a w x w base matrix. */
int poyg_cm_of_list(const int *matrix, int w, poyg_cm **out);
/* Cost_matrix.Two_D.set_affine (src/cost_matrix.ml:405-416): model 0 =
 * Linnear, 1 = Affine gap_open, 2 = No_Alignment; refills the table. */
int poyg_cm_set_affine(poyg_cm *cm, int model, int gap_open);
/* Scripting CM.of_sub_indel / of_sub_indel_affine
 * (src/scripting.ml:4365-4371); gap_open < 0 keeps the linear model. */
int poyg_cm_of_sub_indel(int sub, int indel, int gap_open, poyg_cm **out);
/* Cost_matrix.Three_D.of_two_dim (src/cost_matrix.ml:853-861). */
int poyg_cm3_of_two_dim(const poyg_cm *cm, poyg_cm3 **out);
void poyg_cm3_destroy(poyg_cm3 *cm3);
int poyg_cm3_get_cost(const poyg_cm3 *cm3, int a, int b, int c);
int poyg_cm3_get_median(const poyg_cm3 *cm3, int a, int b, int c);

/* Custom-block wire format, so that a stub whose block holds a poyg_cm keeps
 * poyd's Marshal bytes and checkpoints: the exact byte stream of
 * cm_CAML_serialize / cm_CAML_deserialize (src/cm.c:860-932; identifier
 * "http://www.amnh.org/poy/cost_matrix/two_dimensional.0.1") and of the
 * three-dimensional pair (src/cm.c:890-952; ".../three_dimensional.0.1"):
 * 4-byte big-endian header fields and tables, one byte per median.  The stub
 * moves the buffer with caml_serialize_block_1 / caml_deserialize_block_1. */
int64_t poyg_cm_serialized_size(const poyg_cm *cm);
int poyg_cm_serialize(const poyg_cm *cm, uint8_t *buf, int64_t cap);
int poyg_cm_deserialize(const uint8_t *buf, int64_t size, poyg_cm **out);
int64_t poyg_cm3_serialized_size(const poyg_cm3 *cm3);
int poyg_cm3_serialize(const poyg_cm3 *cm3, uint8_t *buf, int64_t cap);
int poyg_cm3_deserialize(const uint8_t *buf, int64_t size, poyg_cm3 **out);
/* cm_compare (src/cm.c:980-1010): cm_CAML_compare's order; 0 = same content. */
int poyg_cm_compare(const poyg_cm *a, const poyg_cm *b);

/* ---- per-call alignment entry points (host pointers) ------------------ */
/* algn_CAML_simple_2 (src/algn.c:3408-3427): s1 must be at least as long
 * as s2; deltawh is the band parameter Sequence.Align.cost_2 computes. */
int poyg_algn_simple_2(poyg_ctx *ctx, const poyg_cm *cm, const uint8_t *s1, int l1,
                       const uint8_t *s2, int l2, int deltawh, int *cost);
/* algn_CAML_align_2d (src/algn.c:3986-3997): cost + backtrack_2d. */
int poyg_algn_align_2d(poyg_ctx *ctx, const poyg_cm *cm, const uint8_t *s1, int l1,
                       const uint8_t *s2, int l2, int deltawh, int swaped,
                       uint8_t *r1, uint8_t *r2, int cap, int *rlen, int *cost);
/* algn_CAML_align_affine_3 (src/algn.c:2550-2614): lens = median, medianwg,
 * resi, resj lengths; resi follows si whichever sequence is shorter. */
int poyg_algn_align_affine_3(poyg_ctx *ctx, const poyg_cm *cm, const uint8_t *si, int li,
                             const uint8_t *sj, int lj, uint8_t *median, uint8_t *medianwg,
                             uint8_t *resi, uint8_t *resj, int cap, int *lens, int *cost);
/* algn_CAML_cost_affine_3 (src/algn.c:2627-2680). */
int poyg_algn_cost_affine_3(poyg_ctx *ctx, const poyg_cm *cm, const uint8_t *si, int li,
                            const uint8_t *sj, int lj, int *cost);
/* algn_CAML_align_3d (src/algn.c:4004-4012) with the cost of
 * algn_CAML_simple_3 (:3457-3471); r1..r3 receive the aligned triple. */
int poyg_algn_align_3d(poyg_ctx *ctx, const poyg_cm3 *cm3, const uint8_t *s1, int l1,
                       const uint8_t *s2, int l2, const uint8_t *s3, int l3,
                       uint8_t *r1, uint8_t *r2, uint8_t *r3, int cap, int *rlen, int *cost);
/* powell_3D_align (src/ukkCommon.c:110-151; the Iterative ThreeD
 * readjustment, src/sequence.ml:1078-1139): the reference's check-point
 * Ukkonen search (src/ukk.checkp.c:113-781) over Powell's three-way affine
 * model (mismatch mm > 0, gap opening go >= 0, extension ge > 0): its cost
 * (the model's least cost; 1 for three equal strings, as the reference's
 * loop starts at cost 1) and the alignment it picks, right-aligned in
 * ra/rb/rc (cap >= la+lb+lc) with the leading gap.  POYG_EINVAL, as the
 * reference's failure, when a character past the leading gap has no base
 * bit; POYG_EUNSUPPORTED past 65,534 characters a sequence; POYG_ENOMEM when
 * the search's cost levels would exceed 24 GiB. */
int poyg_powell_3d(poyg_ctx *ctx, const uint8_t *a, int la, const uint8_t *b, int lb, const uint8_t *c, int lc,
                   int mm, int go, int ge, uint8_t *ra, uint8_t *rb, uint8_t *rc, int cap, int *rlen, int *cost);
/* The direction matrix that the readers of Matrix.default see after
 * cost_2 / create_edited_2 on (s1, s2) (mat_CAML_get_value,
 * src/matrices.c:236-249; algn_CAML_create_backtrack, src/algn.c:4301-4328;
 * their callers src/charTransform.ml:446-447, src/splitting.ml:225): dm
 * (host, cap >= l1 * l2 entries) receives row i, column j at i * l2 + j --
 * ALIGN 1 | INSERT 2 | DELETE 4 inside the band algn_fill_plane_2 fills, 0
 * outside it (where the reference's matrix holds whatever the previous fill
 * left).  s1 at least as long as s2, linear matrices; *cost as simple_2. */
int poyg_algn_direction_matrix(poyg_ctx *ctx, const poyg_cm *cm, const uint8_t *s1, int l1, const uint8_t *s2,
                               int l2, int deltawh, uint16_t *dm, int64_t cap, int *cost);
/* algn_CAML_median_2_no_gaps / _with_gaps / ancestor_2 / union / median_3
 * (src/algn.c:4186-4235, 4287-4299) on already aligned sequences. */
int poyg_algn_median_2_no_gaps(const poyg_cm *cm, const uint8_t *a, const uint8_t *b, int len,
                               uint8_t *out, int cap, int *olen);
int poyg_algn_median_2_with_gaps(const poyg_cm *cm, const uint8_t *a, const uint8_t *b, int len,
                                 uint8_t *out, int cap, int *olen);
int poyg_algn_ancestor_2(const poyg_cm *cm, const uint8_t *a, const uint8_t *b, int len,
                         uint8_t *out, int cap, int *olen);
int poyg_algn_union(const uint8_t *a, const uint8_t *b, int len, uint8_t *out, int cap, int *olen);
int poyg_algn_median_3(const poyg_cm3 *cm3, const uint8_t *a, const uint8_t *b, const uint8_t *c,
                       int len, uint8_t *out, int cap, int *olen);
/* algn_CAML_worst_2 / algn_CAML_verify_2 (src/algn.c:3372-3406). */
int poyg_algn_worst_2(const poyg_cm *cm, const uint8_t *a, const uint8_t *b, int len, int *cost);
int poyg_algn_verify_2(const poyg_cm *cm, const uint8_t *a, const uint8_t *b, int len, int *cost);

/* ---- batched pairwise alignment (device pointers) --------------------- */
/* Modes of a pair plan. */
#define POYG_MODE_LINEAR_COST 0   /* Sequence.Align.cost_2, linear matrices */
#define POYG_MODE_LINEAR_ALIGN 1  /* align_2 + median_2 (Scripting.DNA.Align.algn) */
#define POYG_MODE_AFFINE_COST 2   /* cost_2 on affine matrices (cost_affine_3) */
#define POYG_MODE_AFFINE_ALIGN 3  /* align_affine_3 */
#define POYG_MODE_TRIPLE 4        /* align_3 + median_3 (poyg_plan_triples) */

/* Plans a batch of n pairs (a_k, b_k).  off/len arrays are HOST arrays that
 * describe where each sequence lives inside the device buffer passed to
 * poyg_plan_run; lengths include the leading gap.  deltaw (host, may be
 * NULL) gives the ?deltaw argument of Sequence.Align.cost_2 per pair for the
 * linear modes (value < 0 = absent); it is ignored for affine modes. */
int poyg_plan_pairs(poyg_ctx *ctx, const poyg_cm *cm, int mode, int32_t n,
                    const int64_t *off_a, const int32_t *len_a, const int64_t *off_b,
                    const int32_t *len_b, const int32_t *deltaw, poyg_plan **out);
void poyg_plan_destroy(poyg_plan *plan);
/* Output layout: pair k owns out[out_off[k] .. out_off[k] + slots*cap_k),
 * cap_k = round16(len_a[k] + len_b[k] + 2), in `slots` consecutive sequences each
 * right-aligned in its cap_k bytes; out_len[slots*k + s] is the length of
 * sequence s.  LINEAR_ALIGN slots: aligned a, aligned b, median_2 (3);
 * AFFINE_ALIGN: median, medianwg, resa, resb (4); COST modes: none. */
int poyg_plan_slots(const poyg_plan *plan);
int64_t poyg_plan_out_bytes(const poyg_plan *plan);
int poyg_plan_out_offsets(const poyg_plan *plan, int64_t *out_off);
/* DP cells the plan evaluates (band cells of the reference algorithm). */
int64_t poyg_plan_cells(const poyg_plan *plan);
/* Runs the plan on `stream` (hipStream_t, or NULL for the context stream).
 * d_seq: device sequences; d_cost: n int32; d_out / d_out_len as above
 * (ignored by cost modes; d_out 16-byte aligned, as hipMalloc returns it).
 * Bytes of a slot left of its right-aligned sequence are unspecified. */
int poyg_plan_run(poyg_plan *plan, const uint8_t *d_seq, int32_t *d_cost, uint8_t *d_out,
                  int32_t *d_out_len, void *stream);
/* Kernel timings of the last run (ms, measured with HIP events on the run's
 * stream): fill kernel and backtrace kernel. */
int poyg_plan_last_times(const poyg_plan *plan, float *fill_ms, float *trace_ms);
int poyg_plan_set_timing(poyg_plan *plan, int enable);

/* Number of pairs of the last run whose results are not to be trusted: a
 * backtrace that left the filled band, or a linear fill whose wait for the
 * previous strip of its pair gave up (a bounded poll; neither happens in a
 * correct run).  0 when every pair completed normally. */
int poyg_plan_status(poyg_plan *plan, int32_t *bad);

/* A batch of pairs whose sequences are in HOST memory, planned, run on the
 * context's stream and read back in one call (no reference counterpart: the
 * batch route of a host-side tree driver, with one upload of descriptors and
 * sequences and one synchronization).  seq / seq_bytes: the sequences; the
 * off / len arrays index into seq (host arrays, as poyg_plan_pairs); deltaw
 * as poyg_plan_pairs, or NULL in the linear modes for the default band of
 * Sequence.Align.cost_2 (?deltaw absent) of every pair.
 * cost[n] receives the costs; in the alignment modes out (out_cap bytes, at
 * least poyg_plan_out_bytes of the same plan: slots of cap_k bytes per pair
 * in pair order, laid out as above) and out_len[n * slots] the outputs.
 * *cells: the band cells evaluated; *failed: as poyg_plan_status (the other
 * pairs' results are valid).  Not for concurrent use of one context. */
int poyg_run_pairs_host(poyg_ctx *ctx, const poyg_cm *cm, int mode, int32_t n, const uint8_t *seq,
                        int64_t seq_bytes, const int64_t *off_a, const int32_t *len_a, const int64_t *off_b,
                        const int32_t *len_b, const int32_t *deltaw, int32_t *cost, uint8_t *out, int64_t out_cap,
                        int32_t *out_len, int64_t *cells, int32_t *failed);

/* ---- batched three-sequence alignment --------------------------------- */
/* Plans n triples for algn_CAML_align_3d + algn_CAML_median_3
 * (src/algn.c:3457-3478, 4004-4017, 4223-4235).  Runs through
 * poyg_plan_run; slots: aligned s1, s2, s3, median_3, each of capacity
 * round16(l1+l2+l3+1).  Direction cubes are processed in batches whose size
 * is bounded by POYG_CUBE_DM_BUDGET_MB (default 16384). */
int poyg_plan_triples(poyg_ctx *ctx, const poyg_cm3 *cm3, int32_t n, const int64_t *off_1,
                      const int32_t *len_1, const int64_t *off_2, const int32_t *len_2,
                      const int64_t *off_3, const int32_t *len_3, poyg_plan **out);

/* ---- Fitch over packed nonadditive characters ------------------------- */
/* nonadd_median (src/nonaddCSc.c:422-462) for n node pairs over `len` u8
 * characters: res[k] = fitch(a[ia[k]], b[ib[k]]), added[k] = changes.
 * Rows of `chars` are len bytes long; device pointers. */
int poyg_fitch_down(poyg_ctx *ctx, uint8_t *d_chars, int32_t len, int32_t n,
                    const int32_t *d_ia, const int32_t *d_ib, const int32_t *d_ires,
                    int64_t *d_added, void *stream);
/* The down pass on host rows (NonaddCS8.median, src/nonaddCS.ml:119-122,
 * through char_nonadd_CAML_basic_median / nonadd_median,
 * src/nonaddCSc.c:422-462, for a batch of row pairs): res
 * (n x len) and added (n) receive the medians and change counts; one upload,
 * one download, one synchronization. */
int poyg_fitch_pairs_host(poyg_ctx *ctx, const uint8_t *a, const uint8_t *b, int32_t n, int32_t len, uint8_t *res,
                          int64_t *added);
/* nonadd_median_3 (src/nonaddCSc.c:472-544): F = uppass(A, N, D1, D2). */
int poyg_fitch_up(poyg_ctx *ctx, uint8_t *d_chars, int32_t len, int32_t n, const int32_t *d_iA,
                  const int32_t *d_iN, const int32_t *d_iD1, const int32_t *d_iD2,
                  const int32_t *d_iF, void *stream);
/* char_nonadd_CAML_median / median_3 / distance on host arrays. */
int poyg_nonadd_median(poyg_ctx *ctx, const uint8_t *a, const uint8_t *b, uint8_t *res, int len,
                       int64_t *added);
int poyg_nonadd_median_3(poyg_ctx *ctx, const uint8_t *A, const uint8_t *N, const uint8_t *D1,
                         const uint8_t *D2, uint8_t *F, int len);
/* The same passes for characters of `width` bytes: 1, 2 or 4, i.e.
 * char_nonadd8/16/32_CAML_basic_median and _median_3 (nonaddCSc.c compiled
 * with CHARSIZE 8, 16, 32: src/myocamlbuild.ml.in:255-269; Node picks the
 * width from the number of states, src/node.ml:1179-1234).  `len` counts
 * characters; rows of `chars` are len * width bytes. */
int poyg_fitch_down_w(poyg_ctx *ctx, int width, uint8_t *d_chars, int32_t len, int32_t n, const int32_t *d_ia,
                      const int32_t *d_ib, const int32_t *d_ires, int64_t *d_added, void *stream);
int poyg_fitch_up_w(poyg_ctx *ctx, int width, uint8_t *d_chars, int32_t len, int32_t n, const int32_t *d_iA,
                    const int32_t *d_iN, const int32_t *d_iD1, const int32_t *d_iD2, const int32_t *d_iF,
                    void *stream);
int poyg_nonadd_median_w(poyg_ctx *ctx, int width, const void *a, const void *b, void *res, int len,
                         int64_t *added);
int poyg_nonadd_median_3_w(poyg_ctx *ctx, int width, const void *A, const void *N, const void *D1, const void *D2,
                           void *F, int len);

/* ---- batched linear costs without device allocation -------------------
 * No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the costs of poyg_plan_pairs +
 * poyg_plan_run in POYG_MODE_LINEAR_COST (Sequence.Align.cost_2 ->
 * algn_CAML_simple_2, src/algn.c:3408-3427), with every buffer the plan
 * would allocate taken from one static device array of
 * poyg_noalloc_scratch_bytes() (32 MiB) per device instead.  Host arrays
 * describe the pairs as for poyg_plan_pairs (deltaw required); d_seq and
 * d_cost are the caller's device buffers.  Pairs whose shorter sequence
 * has more than 2,048 positions, or batches whose descriptors and prefix
 * sums (2 ints per position of each pair's shorter sequence) do not fit the
 * array, are refused with POYG_EUNSUPPORTED.  Every call on a device shares
 * that device's array: calls that overlap on one device race.  The call
 * returns after its work is done. */
int64_t poyg_noalloc_scratch_bytes(void);
int poyg_pairs_cost_linear_noalloc(poyg_ctx *ctx, const poyg_cm *cm, int32_t n, const int64_t *off_a,
                                   const int32_t *len_a, const int64_t *off_b, const int32_t *len_b,
                                   const int32_t *deltaw, const uint8_t *d_seq, int32_t *d_cost, void *stream);

/* ---- tree-pass support (device-resident sequences) ------------------- */
/* Per sequence: stats[2k] = positions carrying the gap bit (seq_CAML_count,
 * src/seq.c:571-581, the count in Sequence.Align.cost_2's band parameter,
 * src/sequence.ml:691-713) and stats[2k + 1] = 1 when every position is
 * the gap (Sequence.is_empty, src/sequence.ml:442-450, which DOS.median and
 * DOS.distance test, src/seqCS.ml:747-776, 818-864).  Device pointers. */
int poyg_seq_stats(poyg_ctx *ctx, const uint8_t *d_seq, const int64_t *d_off, const int32_t *d_len,
                   int32_t n, int32_t *d_stats, void *stream);

/* ---- sequence unions (Sequence.Unions, src/sequence.ml:1810-1895) ------
 * A union of n items is four arrays: seq (the items), end (the reference's
 * `offset` array: 0 on a homologous item, else the item's 1-based place in
 * its run of non-homologous items), ca / cb (`union_c1` / `union_c2`: the
 * child-union positions a homologous item came from).  The reference keeps
 * them as int16 (int32 with USE_LONG_SEQUENCES) bigarrays, right-aligned
 * in their capacity; here they are int32.  A leaf's union is its sequence
 * with end = 0 and ca = cb = 0 .. n-1 (Unions.leaf).
 *
 * union_CAML_make (src/union.c:209-226, union_merge :102-164): the union of
 * two children's unions over their aligned pair (a, b: len columns) and its
 * median with gaps (median == NULL: the matrix's median of each column, as
 * median_2_with_gaps gives for linear matrices).  The children are read
 * left-aligned (na / nb items; their ca / cb are not read); the result is
 * written right-aligned into capacity-cap arrays (cap >= na + nb + 2, the
 * capacity Unions.union allocates) and *out_n is its length.  POYG_EINVAL
 * when the children do not match the alignment.  Host pointers. */
int poyg_union_make(const poyg_cm *cm, const uint8_t *a, const uint8_t *b, const uint8_t *median, int len,
                    const uint8_t *ua_seq, const int32_t *ua_end, int na, const uint8_t *ub_seq,
                    const int32_t *ub_end, int nb, uint8_t *out_seq, int32_t *out_end, int32_t *out_ca,
                    int32_t *out_cb, int cap, int *out_n);
/* One merge of a batch; every pointer is a device pointer. */
typedef struct poyg_union_job {
  const uint8_t *a, *b, *median;      /* aligned pair and median with gaps (NULL: column medians) */
  const uint8_t *ua_seq, *ub_seq;     /* children's unions */
  const int32_t *ua_end, *ub_end;
  uint8_t *out_seq;                   /* result, right-aligned in cap */
  int32_t *out_end, *out_ca, *out_cb;
  int32_t len, na, nb, cap;
} poyg_union_job;
/* The merges of d_jobs[0..n) (a device array) on the device, one thread
 * each; d_out_n[k] = length of union k, or < 0 when its children do not
 * match its alignment or cap is too small.  No counterpart batch call in
 * the reference, which makes one union per node (Node.Union.union,
 * src/node.ml:2355-2408). */
int poyg_union_merge(poyg_ctx *ctx, const poyg_cm *cm, int32_t n, const poyg_union_job *d_jobs,
                     int32_t *d_out_n, void *stream);

/* ---- learned join scorer ----------------------------------------------
 * No counterpart in the reference: POY scores every candidate join exactly
 * (AllDirChar.cost_fn, src/allDirChar.ml:1279-1332).  These entry points
 * let a TBR round order candidates by a predicted join cost and score only
 * the best of them exactly; every reported cost still comes from the exact
 * DP.  The network is 256 -> 256 -> 256 -> 1 with ReLU, bf16 weights and
 * activations, f32 accumulation (MFMA products), trained by mean squared
 * error with Adam.  Parameter vector (f32, poyg_mlp_param_count() entries):
 * W1[256 in][256 out], b1[256], W2[256][256], b2[256], w3[256], b3[1]. */
typedef struct poyg_mlp poyg_mlp;
int64_t poyg_mlp_param_count(void);
/* He-uniform weights from `seed`, zero biases; batches of up to batch_max
 * rows. */
int poyg_mlp_create(poyg_ctx *ctx, int batch_max, uint64_t seed, poyg_mlp **out);
void poyg_mlp_destroy(poyg_mlp *m);
int poyg_mlp_set_params(poyg_mlp *m, const float *host, int64_t n);
int poyg_mlp_get_params(poyg_mlp *m, float *host, int64_t n);
/* Predictions y[r] of n rows of bf16 features x[r][256] (device). */
int poyg_mlp_forward(poyg_mlp *m, const uint16_t *d_x, int64_t n, float *d_y, void *stream);
/* Gradient of the mean squared error of one batch (n <= batch_max rows,
 * device features and f32 targets), kept for poyg_mlp_adam and readable
 * with poyg_mlp_get_grads; *loss (if not NULL) receives the batch loss. */
int poyg_mlp_grad(poyg_mlp *m, const uint16_t *d_x, const float *d_t, int n, float *loss);
int poyg_mlp_get_grads(poyg_mlp *m, float *host, int64_t n);
/* One Adam step on the gradient of the last poyg_mlp_grad. */
int poyg_mlp_adam(poyg_mlp *m, float lr, float beta1, float beta2, float eps);
/* 4-mer profiles: counts[k][b0*64 + b1*16 + b2*4 + b3] over positions 1..
 * of sequence k, a position's base being the lowest base bit of its mask;
 * windows holding a position without a base bit are not counted. */
int poyg_kmer_counts(poyg_ctx *ctx, const uint8_t *d_seq, const int64_t *d_off, const int32_t *d_len, int32_t n,
                     float *d_counts, void *stream);
/* Pair features x[p][k] = bf16(|counts[ia[p]][k] - counts[ib[p]][k]| * scale). */
int poyg_pair_features(poyg_ctx *ctx, const float *d_counts, const int32_t *d_ia, const int32_t *d_ib, int64_t n,
                       float scale, uint16_t *d_x, void *stream);
/* dst row r = src row idx[r], rows of row_bytes (a multiple of 4) bytes:
 * the training loop's batch shuffle. */
int poyg_gather_rows(poyg_ctx *ctx, const void *d_src, int64_t row_bytes, const int32_t *d_idx, int64_t n,
                     void *d_dst, void *stream);
/* C (f32, row stride ldc) = A B over bf16 operands, A(i, k) = a[i*a_rs +
 * k*a_cs], B(k, j) = b[k*b_rs + j*b_cs]: the MFMA tiles the network runs
 * on, exposed for tests. */
int poyg_gemm_bf16_f32(poyg_ctx *ctx, int M, int N, int K, const uint16_t *d_a, int64_t a_rs, int64_t a_cs,
                       const uint16_t *d_b, int64_t b_rs, int64_t b_cs, float *d_c, int64_t ldc, void *stream);

#ifdef __cplusplus
}
#endif
#endif
