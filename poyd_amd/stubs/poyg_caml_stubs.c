/*
 * OCaml-side C stubs that route POY's alignment `external`s to libpoyg.
 *
 * This file replaces the bodies of the stubs in POY's src/algn.c (the
 * alignment entry points), src/cm.c (the cost-matrix custom block) and
 * src/nonaddCSc.c (the Fitch medians) inside POY's own source tree: it
 * is built there by POY's build (src/myocamlbuild.ml.in), against the real
 * OCaml runtime headers and POY's seq.h / nonaddCSc.c definitions, with
 * -I<repo>/include -L<repo>/poyd_amd -lpoyg.  Each stub keeps the name,
 * arity, argument order and result of the one it replaces, so the OCaml
 * `external` declarations (src/sequence.ml:455-773, src/cost_matrix.ml:33-74,
 * src/nonaddCS.ml:61-174) do not change.  Errors of the engine become OCaml
 * Failure exceptions at this edge; the library itself never longjmps.
 *
 * Sequences stay in POY's custom block (src/seq.h:49-58): the stubs pass
 * seq->begin / seq->len to the engine and fill caller-allocated results
 * right-aligned, as seq_prepend does.  Cost matrices: the custom block
 * holds a poyg_cm * (Poyg_cm_val below) with the reference's custom-op
 * identifiers and serialized bytes (poyg_cm_serialize).  The Fitch stubs
 * are compiled once per CHARSIZE (8, 16, 32), as nonaddCSc.c is.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <caml/alloc.h>
#include <caml/custom.h>
#include <caml/fail.h>
#include <caml/memory.h>
#include <caml/mlvalues.h>

#include "poyg.h"
#include "seq.h"   /* POY's: struct seq, seqt, Seq_custom_val, seq_prepend */

/* ---- the engine context: one per process (POY enters C from one thread,
 * src/poyd/PoydThread.ml:40-77) ------------------------------------------ */
static poyg_ctx *poyg_context(void) {
    static poyg_ctx *ctx = NULL;
    if (!ctx && poyg_ctx_create(0, &ctx) != POYG_OK) caml_failwith(poyg_last_error());
    return ctx;
}

static void poyg_fail(int rc) {
    if (rc != POYG_OK) caml_failwith(poyg_last_error());
}

/* A result sequence filled right-aligned by the engine: out[cap - n .. cap)
 * of a temporary buffer is prepended to the caller's (empty) sequence. */
static void put_right_aligned(seqt dst, const uint8_t *buf, int cap, int n) {
    for (int i = cap - 1; i >= cap - n; --i) seq_prepend(dst, buf[i]);
}

/* ---- cost matrices (src/cm.c) ------------------------------------------ */
#define Poyg_cm_val(v) (*(poyg_cm **)Data_custom_val(v))
#define Poyg_cm3_val(v) (*(poyg_cm3 **)Data_custom_val(v))

static void cm_finalize(value v) { poyg_cm_destroy(Poyg_cm_val(v)); }
static void cm3_finalize(value v) { poyg_cm3_destroy(Poyg_cm3_val(v)); }

static int cm_custom_compare(value a, value b) { return poyg_cm_compare(Poyg_cm_val(a), Poyg_cm_val(b)); }

/* cm_CAML_serialize / deserialize (src/cm.c:909-932): the reference's bytes */
static void cm_custom_serialize(value v, uintnat *wsize_32, uintnat *wsize_64) {
    poyg_cm *c = Poyg_cm_val(v);
    int64_t n = poyg_cm_serialized_size(c);
    uint8_t *buf = malloc(n);
    if (!buf) caml_failwith("cm_CAML_serialize: out of memory");
    poyg_fail(poyg_cm_serialize(c, buf, n));
    caml_serialize_block_1(buf, n);
    free(buf);
    *wsize_32 = sizeof(void *);
    *wsize_64 = sizeof(void *);
}

static uintnat cm_custom_deserialize(void *dst) {
    uint8_t head[32];
    caml_deserialize_block_1(head, 32);
    const int32_t lcm = (head[4] << 24) | (head[5] << 16) | (head[6] << 8) | head[7];
    const int64_t n = 32 + 17 * ((int64_t)2 << (2 * lcm));
    uint8_t *buf = malloc(n);
    if (!buf) caml_failwith("cm_CAML_deserialize: out of memory");
    memcpy(buf, head, 32);
    caml_deserialize_block_1(buf + 32, n - 32);
    poyg_fail(poyg_cm_deserialize(buf, n, (poyg_cm **)dst));
    free(buf);
    return sizeof(poyg_cm *);
}

/* the custom-op identifier of src/cm.c:1024-1032 */
static struct custom_operations poyg_cm_ops = {
    "http://www.amnh.org/poy/cost_matrix/two_dimensional.0.1",
    cm_finalize, cm_custom_compare, custom_hash_default, cm_custom_serialize, cm_custom_deserialize,
    custom_compare_ext_default};

static struct custom_operations poyg_cm3_ops = {
    "http://www.amnh.org/poy/cost_matrix/three_dimensional.0.1",
    cm3_finalize, custom_compare_default, custom_hash_default, custom_serialize_default,
    custom_deserialize_default, custom_compare_ext_default};

value cm_CAML_register(value u) {
    CAMLparam1(u);
    caml_register_custom_operations(&poyg_cm_ops);
    caml_register_custom_operations(&poyg_cm3_ops);
    CAMLreturn(Val_unit);
}

/* cm_CAML_create (src/cm.c:1389-1398) */
value cm_CAML_create(value a_sz, value combine, value aff, value go, value all) {
    CAMLparam5(a_sz, combine, aff, go, all);
    CAMLlocal1(res);
    poyg_cm *c = NULL;
    poyg_fail(poyg_cm_create(Int_val(a_sz), Bool_val(combine), Int_val(aff), Int_val(go), Int_val(all), &c));
    res = caml_alloc_custom(&poyg_cm_ops, sizeof(poyg_cm *), 1, 1000000);
    Poyg_cm_val(res) = c;
    CAMLreturn(res);
}

/* the setters and getters of src/cm.c:1295-1345, one line each */
value cm_CAML_set_cost(value a, value b, value c, value v) {
    poyg_fail(poyg_cm_set_cost(Poyg_cm_val(c), Int_val(a), Int_val(b), Int_val(v)));
    return Val_unit;
}
value cm_CAML_set_worst(value a, value b, value c, value v) {
    poyg_fail(poyg_cm_set_worst(Poyg_cm_val(c), Int_val(a), Int_val(b), Int_val(v)));
    return Val_unit;
}
value cm_CAML_set_median(value a, value b, value c, value v) {
    poyg_fail(poyg_cm_set_median(Poyg_cm_val(c), Int_val(a), Int_val(b), Int_val(v)));
    return Val_unit;
}
value cm_CAML_set_prepend(value a, value b, value c) {
    poyg_fail(poyg_cm_set_prepend(Poyg_cm_val(c), Int_val(a), Int_val(b)));
    return Val_unit;
}
value cm_CAML_set_tail(value a, value b, value c) {
    poyg_fail(poyg_cm_set_tail(Poyg_cm_val(c), Int_val(a), Int_val(b)));
    return Val_unit;
}
value cm_CAML_set_affine(value c, value do_aff, value go) {
    poyg_fail(poyg_cm_set_affine_flag(Poyg_cm_val(c), Int_val(do_aff), Int_val(go)));
    return Val_unit;
}
value cm_CAML_get_cost(value a, value b, value c) {
    return Val_int(poyg_cm_get_cost(Poyg_cm_val(c), Int_val(a), Int_val(b)));
}
value cm_CAML_get_worst(value a, value b, value c) {
    return Val_int(poyg_cm_get_worst(Poyg_cm_val(c), Int_val(a), Int_val(b)));
}
value cm_CAML_get_median(value a, value b, value c) {
    return Val_int(poyg_cm_get_median(Poyg_cm_val(c), Int_val(a), Int_val(b)));
}
value cm_CAML_get_prepend(value a, value c) { return Val_int(poyg_cm_get_prepend(Poyg_cm_val(c), Int_val(a))); }
value cm_CAML_get_tail(value a, value c) { return Val_int(poyg_cm_get_tail(Poyg_cm_val(c), Int_val(a))); }
value cm_CAML_get_gap(value c) { return Val_int(poyg_cm_get_gap(Poyg_cm_val(c))); }
value cm_CAML_get_affine(value c) { return Val_int(poyg_cm_get_affine(Poyg_cm_val(c))); }
value cm_CAML_get_gap_opening(value c) { return Val_int(poyg_cm_get_gap_opening(Poyg_cm_val(c))); }

/* Cost_matrix.Three_D.of_two_dim's table (src/cost_matrix.ml:853-861) */
value cm_CAML_three_of_two(value c2) {
    CAMLparam1(c2);
    CAMLlocal1(res);
    poyg_cm3 *c3 = NULL;
    poyg_fail(poyg_cm3_of_two_dim(Poyg_cm_val(c2), &c3));
    res = caml_alloc_custom(&poyg_cm3_ops, sizeof(poyg_cm3 *), 1, 1000000);
    Poyg_cm3_val(res) = c3;
    CAMLreturn(res);
}

/* ---- pairwise alignment (src/algn.c) ------------------------------------
 * Sequence.Align.create_edited_2 calls c_cost_2 (algn_CAML_simple_2) and
 * then extract_edited_2 (algn_CAML_backtrack_2d) on the same sequences and
 * the global Matrix.default: the cost stub remembers its band parameter, and
 * the backtrace stub runs the engine's one-call alignment with it. */
static struct {
    const uint8_t *s1, *s2;
    int l1, l2, deltawh;
} last_cost_2;

/* algn_CAML_simple_2 (src/algn.c:3408-3427) */
value algn_CAML_simple_2(value s1, value s2, value c, value a, value deltawh) {
    CAMLparam5(s1, s2, c, a, deltawh);
    seqt x, y;
    int cost;
    Seq_custom_val(x, s1);
    Seq_custom_val(y, s2);
    poyg_fail(poyg_algn_simple_2(poyg_context(), Poyg_cm_val(c), x->begin, x->len, y->begin, y->len,
                                 Int_val(deltawh), &cost));
    last_cost_2.s1 = x->begin;
    last_cost_2.s2 = y->begin;
    last_cost_2.l1 = x->len;
    last_cost_2.l2 = y->len;
    last_cost_2.deltawh = Int_val(deltawh);
    CAMLreturn(Val_int(cost));
}

/* algn_CAML_backtrack_2d (src/algn.c:3907-3930): the aligned pair of the
 * last algn_CAML_simple_2 call on these sequences. */
value algn_CAML_backtrack_2d(value s1, value s2, value s1p, value s2p, value a, value c, value swap) {
    CAMLparam5(s1, s2, s1p, s2p, a);
    CAMLxparam2(c, swap);
    seqt x, y, rx, ry;
    int len, cost;
    Seq_custom_val(x, s1);
    Seq_custom_val(y, s2);
    Seq_custom_val(rx, s1p);
    Seq_custom_val(ry, s2p);
    if (x->begin != last_cost_2.s1 || y->begin != last_cost_2.s2 || x->len != last_cost_2.l1 ||
        y->len != last_cost_2.l2)
        caml_failwith("algn_CAML_backtrack_2d: no preceding algn_CAML_simple_2 on these sequences");
    const int cap = x->len + y->len + 2;
    uint8_t *buf = malloc(2 * (size_t)cap);
    if (!buf) caml_failwith("algn_CAML_backtrack_2d: out of memory");
    int rc = poyg_algn_align_2d(poyg_context(), Poyg_cm_val(c), x->begin, x->len, y->begin, y->len,
                                last_cost_2.deltawh, Bool_val(swap), buf, buf + cap, cap, &len, &cost);
    if (rc == POYG_OK) {
        put_right_aligned(rx, buf, cap, len);
        put_right_aligned(ry, buf + cap, cap, len);
    }
    free(buf);
    poyg_fail(rc);
    CAMLreturn(Val_unit);
}

value algn_CAML_backtrack_2d_bc(value *argv, int argn) {
    return algn_CAML_backtrack_2d(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6]);
}

/* algn_CAML_align_2d (src/algn.c:3986-3997) */
value algn_CAML_align_2d(value s1, value s2, value c, value a, value s1p, value s2p, value deltawh,
                         value swaped) {
    CAMLparam5(s1, s2, c, a, s1p);
    CAMLxparam3(s2p, deltawh, swaped);
    seqt x, y, rx, ry;
    int len, cost;
    Seq_custom_val(x, s1);
    Seq_custom_val(y, s2);
    Seq_custom_val(rx, s1p);
    Seq_custom_val(ry, s2p);
    const int cap = x->len + y->len + 2;
    uint8_t *buf = malloc(2 * (size_t)cap);
    if (!buf) caml_failwith("algn_CAML_align_2d: out of memory");
    int rc = poyg_algn_align_2d(poyg_context(), Poyg_cm_val(c), x->begin, x->len, y->begin, y->len,
                                Int_val(deltawh), Bool_val(swaped), buf, buf + cap, cap, &len, &cost);
    if (rc == POYG_OK) {
        put_right_aligned(rx, buf, cap, len);
        put_right_aligned(ry, buf + cap, cap, len);
    }
    free(buf);
    poyg_fail(rc);
    CAMLreturn(Val_int(cost));
}

value algn_CAML_align_2d_bc(value *argv, int argn) {
    return algn_CAML_align_2d(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7]);
}

/* algn_CAML_align_affine_3 (src/algn.c:2550-2614): median, medianwg, resi,
 * resj; the results were created with capacity |si| + |sj| + 2
 * (src/sequence.ml:469-478). */
value algn_CAML_align_affine_3(value si, value sj, value cm, value am, value resi, value resj, value median,
                               value medianwg) {
    CAMLparam4(si, sj, cm, am);
    CAMLxparam4(resi, resj, median, medianwg);
    seqt a, b, ri, rj, m, mwg;
    int lens[4], cost;
    Seq_custom_val(a, si);
    Seq_custom_val(b, sj);
    Seq_custom_val(ri, resi);
    Seq_custom_val(rj, resj);
    Seq_custom_val(m, median);
    Seq_custom_val(mwg, medianwg);
    const int cap = a->len + b->len + 2;
    uint8_t *buf = malloc(4 * (size_t)cap);
    if (!buf) caml_failwith("algn_CAML_align_affine_3: out of memory");
    int rc = poyg_algn_align_affine_3(poyg_context(), Poyg_cm_val(cm), a->begin, a->len, b->begin, b->len, buf,
                                      buf + cap, buf + 2 * cap, buf + 3 * cap, cap, lens, &cost);
    if (rc == POYG_OK) {
        put_right_aligned(m, buf, cap, lens[0]);
        put_right_aligned(mwg, buf + cap, cap, lens[1]);
        put_right_aligned(ri, buf + 2 * cap, cap, lens[2]);
        put_right_aligned(rj, buf + 3 * cap, cap, lens[3]);
    }
    free(buf);
    poyg_fail(rc);
    CAMLreturn(Val_int(cost));
}

value algn_CAML_align_affine_3_bc(value *argv, int argn) {
    return algn_CAML_align_affine_3(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7]);
}

/* algn_CAML_cost_affine_3 (src/algn.c:2627-2680) */
value algn_CAML_cost_affine_3(value si, value sj, value cm, value am) {
    CAMLparam4(si, sj, cm, am);
    seqt a, b;
    int cost;
    Seq_custom_val(a, si);
    Seq_custom_val(b, sj);
    poyg_fail(poyg_algn_cost_affine_3(poyg_context(), Poyg_cm_val(cm), a->begin, a->len, b->begin, b->len, &cost));
    CAMLreturn(Val_int(cost));
}

/* algn_CAML_align_3d (src/algn.c:4004-4012) with the cost of
 * algn_CAML_simple_3 (:3457-3471); uk (the unused Ukkonen barrier) is
 * ignored as in the reference. */
value algn_CAML_align_3d(value s1, value s2, value s3, value c, value a, value s1p, value s2p, value s3p,
                         value uk) {
    CAMLparam5(s1, s2, s3, c, a);
    CAMLxparam4(s1p, s2p, s3p, uk);
    seqt x, y, z, rx, ry, rz;
    int len, cost;
    Seq_custom_val(x, s1);
    Seq_custom_val(y, s2);
    Seq_custom_val(z, s3);
    Seq_custom_val(rx, s1p);
    Seq_custom_val(ry, s2p);
    Seq_custom_val(rz, s3p);
    const int cap = x->len + y->len + z->len + 1;
    uint8_t *buf = malloc(3 * (size_t)cap);
    if (!buf) caml_failwith("algn_CAML_align_3d: out of memory");
    int rc = poyg_algn_align_3d(poyg_context(), Poyg_cm3_val(c), x->begin, x->len, y->begin, y->len, z->begin,
                                z->len, buf, buf + cap, buf + 2 * cap, cap, &len, &cost);
    if (rc == POYG_OK) {
        put_right_aligned(rx, buf, cap, len);
        put_right_aligned(ry, buf + cap, cap, len);
        put_right_aligned(rz, buf + 2 * cap, cap, len);
    }
    free(buf);
    poyg_fail(rc);
    CAMLreturn(Val_int(cost));
}

value algn_CAML_align_3d_bc(value *argv, int argn) {
    return algn_CAML_align_3d(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8]);
}

/* Column rules on aligned sequences (src/algn.c:4186-4235, 4287-4299),
 * on the host. */
#define COLUMN_STUB(NAME, CALL)                                                      \
    value NAME(value s1, value s2, value m, value sm) {                              \
        CAMLparam4(s1, s2, m, sm);                                                   \
        seqt x, y, r;                                                                \
        int n;                                                                       \
        Seq_custom_val(x, s1);                                                       \
        Seq_custom_val(y, s2);                                                       \
        Seq_custom_val(r, sm);                                                       \
        const int cap = x->len + 1;                                                  \
        uint8_t *buf = malloc(cap);                                                  \
        if (!buf) caml_failwith(#NAME ": out of memory");                            \
        int rc = CALL;                                                               \
        if (rc == POYG_OK) put_right_aligned(r, buf, cap, n);                        \
        free(buf);                                                                   \
        poyg_fail(rc);                                                               \
        CAMLreturn(Val_unit);                                                        \
    }

COLUMN_STUB(algn_CAML_median_2_no_gaps,
            poyg_algn_median_2_no_gaps(Poyg_cm_val(m), x->begin, y->begin, x->len, buf, cap, &n))
COLUMN_STUB(algn_CAML_median_2_with_gaps,
            poyg_algn_median_2_with_gaps(Poyg_cm_val(m), x->begin, y->begin, x->len, buf, cap, &n))
COLUMN_STUB(algn_CAML_ancestor_2,
            poyg_algn_ancestor_2(Poyg_cm_val(m), x->begin, y->begin, x->len, buf, cap, &n))

value algn_CAML_union(value s1, value s2, value su) {
    CAMLparam3(s1, s2, su);
    seqt x, y, r;
    int n;
    Seq_custom_val(x, s1);
    Seq_custom_val(y, s2);
    Seq_custom_val(r, su);
    uint8_t *buf = malloc(x->len + 1);
    if (!buf) caml_failwith("algn_CAML_union: out of memory");
    int rc = poyg_algn_union(x->begin, y->begin, x->len, buf, x->len + 1, &n);
    if (rc == POYG_OK) put_right_aligned(r, buf, x->len + 1, n);
    free(buf);
    poyg_fail(rc);
    CAMLreturn(Val_unit);
}

value algn_CAML_median_3(value s1, value s2, value s3, value m, value sm) {
    CAMLparam5(s1, s2, s3, m, sm);
    seqt x, y, z, r;
    int n;
    Seq_custom_val(x, s1);
    Seq_custom_val(y, s2);
    Seq_custom_val(z, s3);
    Seq_custom_val(r, sm);
    uint8_t *buf = malloc(x->len + 1);
    if (!buf) caml_failwith("algn_CAML_median_3: out of memory");
    int rc = poyg_algn_median_3(Poyg_cm3_val(m), x->begin, y->begin, z->begin, x->len, buf, x->len + 1, &n);
    if (rc == POYG_OK) put_right_aligned(r, buf, x->len + 1, n);
    free(buf);
    poyg_fail(rc);
    CAMLreturn(Val_unit);
}

value algn_CAML_worst_2(value s1, value s2, value c) {
    CAMLparam3(s1, s2, c);
    seqt x, y;
    int cost;
    Seq_custom_val(x, s1);
    Seq_custom_val(y, s2);
    poyg_fail(poyg_algn_worst_2(Poyg_cm_val(c), x->begin, y->begin, x->len, &cost));
    CAMLreturn(Val_int(cost));
}

value algn_CAML_verify_2(value s1, value s2, value c) {
    CAMLparam3(s1, s2, c);
    seqt x, y;
    int cost;
    Seq_custom_val(x, s1);
    Seq_custom_val(y, s2);
    poyg_fail(poyg_algn_verify_2(Poyg_cm_val(c), x->begin, y->begin, x->len, &cost));
    CAMLreturn(Val_int(cost));
}

/* powell_3D_align (src/ukkCommon.c:110-151) */
value powell_3D_align(value sa, value sb, value sc, value ra, value rb, value rc_, value mm, value go,
                      value ge) {
    CAMLparam5(sa, sb, sc, ra, rb);
    CAMLxparam4(rc_, mm, go, ge);
    seqt a, b, c, x, y, z;
    int len, cost;
    Seq_custom_val(a, sa);
    Seq_custom_val(b, sb);
    Seq_custom_val(c, sc);
    Seq_custom_val(x, ra);
    Seq_custom_val(y, rb);
    Seq_custom_val(z, rc_);
    const int cap = a->len + b->len + c->len;
    uint8_t *o = malloc(3 * (size_t)cap);
    if (!o) caml_failwith("powell_3D_align: out of memory");
    int rc = poyg_powell_3d(poyg_context(), a->begin, a->len, b->begin, b->len, c->begin, c->len, Int_val(mm),
                            Int_val(go), Int_val(ge), o, o + cap, o + 2 * cap, cap, &len, &cost);
    if (rc == POYG_OK) {
        put_right_aligned(x, o, cap, len);
        put_right_aligned(y, o + cap, cap, len);
        put_right_aligned(z, o + 2 * cap, cap, len);
    }
    free(o);
    poyg_fail(rc);
    CAMLreturn(Val_int(cost));
}

value powell_3D_align_bc(value *argv, int argn) {
    return powell_3D_align(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8]);
}

/* ---- sequence unions (src/union.c) ---------------------------------------
 * Sequence.Unions.u keeps its sequence and three UNION_OFFT bigarrays
 * (offset, union_c1, union_c2; int16 by default) right-aligned in their
 * capacity (union_CAML_unwrap, src/union.c:189-207).  The engine reads the
 * children left-aligned in int32 and writes the result right-aligned; the
 * stub widens and narrows. */
#include <caml/bigarray.h>
#include "union.h"   /* POY's: UNION_OFFT */

static int32_t *union_offsets_i32(value u, int len) {
    const UNION_OFFT *off = (const UNION_OFFT *)Caml_ba_data_val(Field(u, 1));
    int cap = (int)Caml_ba_array_val(Field(u, 1))->dim[0];
    int32_t *w = malloc(sizeof(int32_t) * (len > 0 ? len : 1));
    for (int k = 0; k < len; ++k) w[k] = off[cap - len + k];
    return w;
}

/* union_CAML_make (src/union.c:209-226) */
value union_CAML_make(value s1, value s2, value smedian, value a, value b, value c, value cm) {
    CAMLparam5(s1, s2, a, b, c);
    CAMLxparam2(smedian, cm);
    seqt x, y, m, ua, ub, uc;
    Seq_custom_val(x, s1);
    Seq_custom_val(y, s2);
    Seq_custom_val(m, smedian);
    Seq_custom_val(ua, Field(a, 0));
    Seq_custom_val(ub, Field(b, 0));
    Seq_custom_val(uc, Field(c, 0));
    int32_t *ea = union_offsets_i32(a, ua->len), *eb = union_offsets_i32(b, ub->len);
    int cap = uc->cap, n = 0;
    uint8_t *items = malloc(cap > 0 ? cap : 1);
    int32_t *oe = malloc(sizeof(int32_t) * (cap > 0 ? cap : 1));
    int32_t *o1 = malloc(sizeof(int32_t) * (cap > 0 ? cap : 1));
    int32_t *o2 = malloc(sizeof(int32_t) * (cap > 0 ? cap : 1));
    int rc = poyg_union_make(Poyg_cm_val(cm), x->begin, y->begin, m->begin, x->len, ua->begin, ea, ua->len,
                             ub->begin, eb, ub->len, items, oe, o1, o2, cap, &n);
    if (rc == POYG_OK) {
        UNION_OFFT *de = (UNION_OFFT *)Caml_ba_data_val(Field(c, 1));
        UNION_OFFT *d1 = (UNION_OFFT *)Caml_ba_data_val(Field(c, 2));
        UNION_OFFT *d2 = (UNION_OFFT *)Caml_ba_data_val(Field(c, 3));
        put_right_aligned(uc, items, cap, n);
        for (int k = cap - n; k < cap; ++k) {
            de[k] = (UNION_OFFT)oe[k];
            d1[k] = (UNION_OFFT)o1[k];
            d2[k] = (UNION_OFFT)o2[k];
        }
    }
    free(ea); free(eb); free(items); free(oe); free(o1); free(o2);
    poyg_fail(rc);
    CAMLreturn(Val_unit);
}

value union_CAML_make_b(value *argv, int argn) {
    return union_CAML_make(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6]);
}

/* ---- Fitch (src/nonaddCSc.c), compiled once per CHARSIZE ---------------
 * The character sets stay in POY's custom block (struct _naca_t,
 * src/nonaddCSc.c:292-300: len, heur, code, added_cost, data in vectors of
 * BLOCK_LEN characters after the header); the stubs run the engine on the
 * len characters and clear the padding characters of the last vector, which
 * the reference's vector loop leaves at zero. */
#ifdef CHARSIZE
#define POYG_W (CHARSIZE / 8)
#define NAC_DATA(n) ((uint8_t *)((n) + 1))
#define NAC_PADDED(n) (((n)->len + BLOCK_LEN - 1) / BLOCK_LEN * BLOCK_LEN)

/* char_nonadd_CAML_basic_median (src/nonaddCSc.c:922-945) */
value char_nonadd_CAML_basic_median(value v, value a, value b) {
    CAMLparam3(v, a, b);
    nacat na, nb, nres;
    int64_t added = 0;
    Nonadd_Custom_val(a, na);
    Nonadd_Custom_val(b, nb);
    Nonadd_Custom_val(v, nres);
    nres->heur = na->heur;
    poyg_fail(poyg_nonadd_median_w(poyg_context(), POYG_W, NAC_DATA(na), NAC_DATA(nb), NAC_DATA(nres),
                                   (int)na->len, &added));
    memset(NAC_DATA(nres) + na->len * POYG_W, 0, (NAC_PADDED(na) - na->len) * POYG_W);
    nres->added_cost = (long)added;
    CAMLreturn(Val_unit);
}

/* char_nonadd_CAML_median_3 (src/nonaddCSc.c:998-1026): F = uppass(A, N, D1, D2) */
value char_nonadd_CAML_median_3(value v, value vA, value vN, value vD1, value vD2) {
    CAMLparam5(v, vA, vN, vD1, vD2);
    nacat nA, nN, n1, n2, nres;
    Nonadd_Custom_val(vA, nA);
    Nonadd_Custom_val(vN, nN);
    Nonadd_Custom_val(vD1, n1);
    Nonadd_Custom_val(vD2, n2);
    Nonadd_Custom_val(v, nres);
    poyg_fail(poyg_nonadd_median_3_w(poyg_context(), POYG_W, NAC_DATA(nA), NAC_DATA(nN), NAC_DATA(n1),
                                     NAC_DATA(n2), NAC_DATA(nres), (int)nA->len));
    memset(NAC_DATA(nres) + nA->len * POYG_W, 0, (NAC_PADDED(nA) - nA->len) * POYG_W);
    CAMLreturn(Val_unit);
}
#endif
