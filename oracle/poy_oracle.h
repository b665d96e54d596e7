/*
 * poy_oracle.h -- CPU restatement of POY's direct-optimization hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP engine in poyd_amd/ is compared with
 * this library: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only to check or to time the reference
 * algorithm on the CPU.  The product path never calls it.
 *
 * Each function cites the reference file:line range it restates (paths are
 * relative to the lealanko/poyd source tree).  Sequences are arrays of 5-bit
 * IUPAC masks (A=1, C=2, G=4, T=8, gap=16) that carry POY's leading gap at
 * index 0.  Output sequences are produced the way POY's seq_prepend builds
 * them: written from the end of a caller buffer of capacity `cap`; the
 * functions return the produced length and the result occupies
 * out[cap-len .. cap-1].
 */
#ifndef POY_ORACLE_H
#define POY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PO_HIGH_NUM 1000000           /* src/algn.c:38 */
#define PO_OCAML_MAX_INT 1073741823   /* src/cost_matrix.ml:41 (max_int lsr 1) */
#define PO_CM2_SLOTS 2048             /* 2 * 2^lcm * 2^lcm for lcm = 5, src/cm.c:340 */
#define PO_CM3_SLOTS (64 * 64 * 64)   /* (2^(lcm+1))^3 for lcm = 5, src/cm.c:397 */

/* Two-dimensional transformation cost matrix, src/cm.h:32-46. */
typedef struct po_cm2 {
    int a_sz, lcm, gap, cost_model, combinations, gap_open, is_metric, all_elements;
    int cost[PO_CM2_SLOTS];
    int worst[PO_CM2_SLOTS];
    int prepend[PO_CM2_SLOTS];
    int tail[PO_CM2_SLOTS];
    uint8_t median[PO_CM2_SLOTS];
} po_cm2;

/* Three-dimensional transformation cost matrix, src/cm.h:222-245. */
typedef struct po_cm3 {
    int a_sz, lcm, gap, cost_model, combinations, gap_open, all_elements;
    int cost[PO_CM3_SLOTS];
    uint8_t median[PO_CM3_SLOTS];
} po_cm3;

/* ---- R1: cost matrices (src/cost_matrix.ml, src/cm.c) ---------------- */
po_cm2 *po_cm2_create(int a_sz, int combinations, int cost_model, int gap_open, int all_elements);
void po_cm2_free(po_cm2 *c);
int po_cm2_cost(const po_cm2 *c, int a, int b);
int po_cm2_worst(const po_cm2 *c, int a, int b);
int po_cm2_median(const po_cm2 *c, int a, int b);
/* Cost_matrix.Two_D.fill_cost_matrix (src/cost_matrix.ml:520-538) on a w*w
 * row-major input matrix. */
po_cm2 *po_cm2_of_list(const int *mat, int w, int use_comb, int all_elements);
/* Cost_matrix.Two_D.of_transformations_and_gaps true 5 sub indel 31
 * (src/cost_matrix.ml:645-656), i.e. Scripting CM.of_sub_indel. */
po_cm2 *po_cm2_of_sub_indel(int sub, int indel);
/* Cost_matrix.Two_D.set_affine (src/cost_matrix.ml:405-416); model 0 =
 * Linnear, 1 = Affine gap_open, 2 = No_Alignment. */
void po_cm2_set_affine(po_cm2 *c, int model, int gap_open);
/* the matrix's gap opening (Cost_matrix.Two_D.gap_opening). */
int po_cm2_gap_open(const po_cm2 *c);
/* Cost_matrix.Three_D.of_two_dim (src/cost_matrix.ml:803-861). */
po_cm3 *po_cm3_of_two_dim(const po_cm2 *c);
void po_cm3_free(po_cm3 *c);
int po_cm3_cost(const po_cm3 *c, int a, int b, int d);
int po_cm3_median(const po_cm3 *c, int a, int b, int d);

/* ---- R3: band width (src/sequence.ml:691-713) ------------------------- */
int po_count_gaps(const uint8_t *s, int len, int gap);
/* deltaw of Sequence.Align.cost_2; has_deltaw = 0 means ?deltaw absent. */
int po_deltaw(const uint8_t *s1, int l1, const uint8_t *s2, int l2, const po_cm2 *c,
              int has_deltaw, int deltaw);

/* ---- R4/R5: banded linear NW (src/algn.c:87-968, 3246-3283) ----------- */
/* algn_nw: s1 must be at least as long as s2.  dm (l1*l2 u16, may be NULL)
 * receives the direction matrix; band_lo/band_hi (l1 ints each, may be
 * NULL) receive the filled column range of every row. */
int po_algn_nw(const uint8_t *s1, int l1, const uint8_t *s2, int l2, const po_cm2 *c,
               int deltawh, uint16_t *dm, int *band_lo, int *band_hi);
/* Number of DP cells algn_nw evaluates (the band-GCUPS numerator). */
int64_t po_algn_nw_cells(int l1, int l2, int deltawh);

/* ---- R6: backtrack_2d linear branch (src/algn.c:3551-3666) ------------ */
/* Returns the aligned length, or -1 if the walk read a cell outside the band. */
int po_backtrack_2d(const uint16_t *dm, const int *band_lo, const int *band_hi,
                    const uint8_t *s1, int l1, const uint8_t *s2, int l2,
                    const po_cm2 *c, int swaped, uint8_t *r1, uint8_t *r2, int cap);

/* ---- Sequence.Align wrappers (src/sequence.ml:716-869) ---------------- */
int po_cost_2(const uint8_t *s1, int l1, const uint8_t *s2, int l2, const po_cm2 *c,
              int has_deltaw, int deltaw);
/* align_2 ~first_gap:true: returns cost, writes the two aligned sequences
 * (capacity cap >= l1+l2) and their common length. */
int po_align_2(const uint8_t *s1, int l1, const uint8_t *s2, int l2, const po_cm2 *c,
               uint8_t *o1, uint8_t *o2, int cap, int *olen);

/* ---- R7: medians (src/algn.c:4024-4184) -------------------------------- */
int po_median_2_no_gaps(const uint8_t *a, const uint8_t *b, int len, const po_cm2 *c,
                        uint8_t *out, int cap);
int po_median_2_with_gaps(const uint8_t *a, const uint8_t *b, int len, const po_cm2 *c,
                          uint8_t *out, int cap);
int po_ancestor_2(const uint8_t *a, const uint8_t *b, int len, const po_cm2 *c,
                  uint8_t *out, int cap);
int po_union(const uint8_t *a, const uint8_t *b, int len, uint8_t *out, int cap);
/* union_merge (src/union.c:102-164): the union of two children's unions
 * over their aligned pair (a, b, lena columns) and its median with gaps.
 * Children left-aligned (items, offset array); output right-aligned in cap
 * (items, offset, union_c1, union_c2).  Returns the length or -1. */
int po_union_make(const uint8_t *a, const uint8_t *b, const uint8_t *median, int lena, int gap,
                  const uint8_t *ua, const int32_t *ua_off, int na,
                  const uint8_t *ub, const int32_t *ub_off, int nb,
                  uint8_t *out, int32_t *out_off, int32_t *out_ca, int32_t *out_cb, int cap);
/* algn_calculate_from_2_aligned with the cost table (src/algn.c:3316-3370). */
int po_verify_2(const uint8_t *a, const uint8_t *b, int len, const po_cm2 *c);
int po_worst_2(const uint8_t *a, const uint8_t *b, int len, const po_cm2 *c);

/* ---- R8/R9: affine band + backtrace (src/algn.c:1695-2680) ------------ */
/* algn_CAML_align_affine_3: outputs are indexed by the caller's argument
 * order (resi belongs to si).  lens[0..3] = median, medianwg, resi, resj.
 * Returns the cost, or -1 if the backtrace read a cell outside the band. */
int po_align_affine_3(const uint8_t *si, int li, const uint8_t *sj, int lj, const po_cm2 *c,
                      uint8_t *median, uint8_t *medianwg, uint8_t *resi, uint8_t *resj,
                      int cap, int *lens);
/* algn_CAML_cost_affine_3 (no backtrace variant). */
int po_cost_affine_3(const uint8_t *si, int li, const uint8_t *sj, int lj, const po_cm2 *c);
/* Number of DP cells the affine band evaluates (interior band cells). */
int64_t po_affine_cells(int li, int lj);

/* ---- R10/R11: three-dimensional alignment (src/algn.c:2812-3055,
 * 3829-3905, 4160-4173) ---------------------------------------------------- */
/* align_3 ~first_gap:true: cost + three aligned sequences (cap >= l1+l2+l3). */
int po_align_3(const uint8_t *s1, int l1, const uint8_t *s2, int l2, const uint8_t *s3, int l3,
               const po_cm3 *c, uint8_t *o1, uint8_t *o2, uint8_t *o3, int cap, int *olen);
int po_cost_3(const uint8_t *s1, int l1, const uint8_t *s2, int l2, const uint8_t *s3, int l3,
              const po_cm3 *c);
int po_median_3(const uint8_t *a, const uint8_t *b, const uint8_t *d, int len, const po_cm3 *c,
                uint8_t *out, int cap);

/* ---- R12/R13: Fitch over packed nonadditive characters
 * (src/nonaddCSc.c:422-544, 1076-1105, 1349-1376) --------------------------- */
int64_t po_nonadd_median(const uint8_t *a, const uint8_t *b, uint8_t *res, int len);
void po_nonadd_median_3(const uint8_t *A, const uint8_t *N, const uint8_t *D1,
                        const uint8_t *D2, uint8_t *F, int len);
int64_t po_nonadd_median_w(int width, const void *a, const void *b, void *res, int len);
void po_nonadd_median_3_w(int width, const void *A, const void *N, const void *D1, const void *D2, void *F,
                          int len);
int64_t po_nonadd_distance(const uint8_t *a, const uint8_t *b, int len);
int64_t po_nonadd_dist_2(const uint8_t *a, const uint8_t *b, const uint8_t *c, int len);
void po_nonadd_reroot_median(const uint8_t *a, const uint8_t *b, uint8_t *res, int len);

/* ---- R15: powell_3D_align (src/ukkCommon.c, src/ukk.checkp.c) ----------
 * Least cost of Powell's three-way affine model (mismatch mm, gap open go,
 * gap extension ge) and one alignment of that cost; inputs with the leading
 * gap.  -1 when a character has no base bit. */
int po_powell_3d(const uint8_t *a, int la, const uint8_t *b, int lb, const uint8_t *c, int lc, int mm,
                 int go, int ge, uint8_t *oa, uint8_t *ob, uint8_t *oc, int cap, int *olen);
/* The reference's check-point Ukkonen search itself (src/ukk.checkp.c):
 * its cost and the alignment it picks.  planes_mode 0 reuses cost planes as
 * the reference does, 1 keeps every plane. */
int po_powell_ukk(const uint8_t *a, int la, const uint8_t *b, int lb, const uint8_t *c, int lc, int mm, int go,
                  int ge, int planes_mode, uint8_t *oa, uint8_t *ob, uint8_t *oc, int cap, int *olen);

#ifdef __cplusplus
}
#endif
#endif
