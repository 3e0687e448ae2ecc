/*
 * pipeline.c -- candidate generation and the whole CPU path for the oracle
 * (test infrastructure only; see oracle.h).
 *
 * Candidates are the closed overlaps of exact double boxes, the set that
 * CGAL's box_intersection_d reports for the reference CPU path
 * (src/CPU/TestSuite.cpp:79-102). They are found here with a sort-and-sweep
 * along x, or by brute force for small inputs. The reference GPU path uses
 * padded FP32 boxes (src/common/utils.cu:595-643) and an inclusive overlap
 * test (src/traversal/crossCheck.cu:5-9), which yields a superset.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "obig.h"
#include "oracle.h"

typedef struct { double lo[3], hi[3]; } dbox;

static void tri_box(const oracle_mesh* M, uint32_t t, dbox* b) {
    const uint32_t* v = M->tri + 3 * (size_t)t;
    for (int k = 0; k < 3; ++k) {
        double x0 = M->xyz[3 * (size_t)v[0] + k], x1 = M->xyz[3 * (size_t)v[1] + k],
               x2 = M->xyz[3 * (size_t)v[2] + k];
        double lo = x0 < x1 ? x0 : x1, hi = x0 > x1 ? x0 : x1;
        b->lo[k] = lo < x2 ? lo : x2;
        b->hi[k] = hi > x2 ? hi : x2;
    }
}

static int overlap_yz(const dbox* a, const dbox* b) {
    return a->lo[1] <= b->hi[1] && b->lo[1] <= a->hi[1] && a->lo[2] <= b->hi[2] && b->lo[2] <= a->hi[2];
}

static int overlap3(const dbox* a, const dbox* b) {
    return a->lo[0] <= b->hi[0] && b->lo[0] <= a->hi[0] && overlap_yz(a, b);
}

/* ---------------------------------------------------------- pair buffers */

typedef struct {
    uint32_t* p;
    int64_t n, cap;
} pairbuf;

static void pb_push(pairbuf* b, uint32_t a, uint32_t c) {
    if (b->n == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 4096;
        b->p = (uint32_t*)realloc(b->p, (size_t)b->cap * 2 * sizeof(uint32_t));
        if (!b->p) abort();
    }
    b->p[2 * b->n] = a;
    b->p[2 * b->n + 1] = c;
    b->n++;
}

/* ------------------------------------------------------------ sweep state */

typedef struct {
    const oracle_mesh *A, *B;
    uint32_t nA;
    dbox *ba, *bb;
    uint32_t *sa, *sb; /* indices sorted by lo.x */
    double *ka, *kb;   /* sorted lo.x keys */
    int method;
    int nthreads;
    pairbuf* out; /* per thread */
} cand_ctx;

static const dbox* g_sort_boxes;
static int cmp_lox(const void* x, const void* y) {
    uint32_t i = *(const uint32_t*)x, j = *(const uint32_t*)y;
    double a = g_sort_boxes[i].lo[0], b = g_sort_boxes[j].lo[0];
    if (a < b) return -1;
    if (a > b) return 1;
    return i < j ? -1 : (i > j);
}

/* first index with key >= v (strict: key > v) */
static uint32_t lower_bound(const double* k, uint32_t n, double v, int strict) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t m = lo + (hi - lo) / 2;
        int go_right = strict ? (k[m] <= v) : (k[m] < v);
        if (go_right) lo = m + 1;
        else hi = m;
    }
    return lo;
}

typedef struct { cand_ctx* c; int tid; } cand_arg;

static void* cand_worker(void* p) {
    cand_arg* arg = (cand_arg*)p;
    cand_ctx* c = arg->c;
    pairbuf* out = &c->out[arg->tid];
    uint32_t nB = c->B->nt;
    if (c->method == 1) {
        for (uint32_t a = (uint32_t)arg->tid; a < c->nA; a += (uint32_t)c->nthreads)
            for (uint32_t b = 0; b < nB; ++b)
                if (overlap3(&c->ba[a], &c->bb[b])) pb_push(out, a, b);
        return NULL;
    }
    /* A boxes report the B boxes whose lo.x lies in [A.lo.x, A.hi.x] */
    for (uint32_t i = (uint32_t)arg->tid; i < c->nA; i += (uint32_t)c->nthreads) {
        const dbox* a = &c->ba[i];
        for (uint32_t j = lower_bound(c->kb, nB, a->lo[0], 0); j < nB && c->kb[j] <= a->hi[0]; ++j) {
            uint32_t b = c->sb[j];
            if (overlap_yz(a, &c->bb[b])) pb_push(out, i, b);
        }
    }
    /* B boxes report the A boxes whose lo.x lies in (B.lo.x, B.hi.x] */
    for (uint32_t j = (uint32_t)arg->tid; j < nB; j += (uint32_t)c->nthreads) {
        const dbox* b = &c->bb[j];
        for (uint32_t i = lower_bound(c->ka, c->nA, b->lo[0], 1); i < c->nA && c->ka[i] <= b->hi[0]; ++i) {
            uint32_t a = c->sa[i];
            if (overlap_yz(&c->ba[a], b)) pb_push(out, a, j);
        }
    }
    return NULL;
}

/* Runs candidate generation; returns per-thread buffers (caller frees). */
static pairbuf* run_candidates(const oracle_mesh* A, const oracle_mesh* B, uint32_t nA, int method, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    cand_ctx c;
    memset(&c, 0, sizeof(c));
    c.A = A;
    c.B = B;
    c.nA = nA;
    c.method = method;
    c.nthreads = nthreads;
    c.ba = (dbox*)malloc(sizeof(dbox) * (nA ? nA : 1));
    c.bb = (dbox*)malloc(sizeof(dbox) * (B->nt ? B->nt : 1));
    for (uint32_t i = 0; i < nA; ++i) tri_box(A, i, &c.ba[i]);
    for (uint32_t i = 0; i < B->nt; ++i) tri_box(B, i, &c.bb[i]);
    if (method == 0) {
        c.sa = (uint32_t*)malloc(sizeof(uint32_t) * (nA ? nA : 1));
        c.sb = (uint32_t*)malloc(sizeof(uint32_t) * (B->nt ? B->nt : 1));
        c.ka = (double*)malloc(sizeof(double) * (nA ? nA : 1));
        c.kb = (double*)malloc(sizeof(double) * (B->nt ? B->nt : 1));
        for (uint32_t i = 0; i < nA; ++i) c.sa[i] = i;
        for (uint32_t i = 0; i < B->nt; ++i) c.sb[i] = i;
        g_sort_boxes = c.ba;
        qsort(c.sa, nA, sizeof(uint32_t), cmp_lox);
        g_sort_boxes = c.bb;
        qsort(c.sb, B->nt, sizeof(uint32_t), cmp_lox);
        for (uint32_t i = 0; i < nA; ++i) c.ka[i] = c.ba[c.sa[i]].lo[0];
        for (uint32_t i = 0; i < B->nt; ++i) c.kb[i] = c.bb[c.sb[i]].lo[0];
    }
    c.out = (pairbuf*)calloc((size_t)nthreads, sizeof(pairbuf));
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    cand_arg* args = (cand_arg*)malloc(sizeof(cand_arg) * (size_t)nthreads);
    for (int t = 0; t < nthreads; ++t) {
        args[t].c = &c;
        args[t].tid = t;
        pthread_create(&th[t], NULL, cand_worker, &args[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(args);
    free(c.ba);
    free(c.bb);
    free(c.sa);
    free(c.sb);
    free(c.ka);
    free(c.kb);
    return c.out;
}

int64_t oracle_candidates(const oracle_mesh* A, const oracle_mesh* B, int method, int nthreads, uint32_t* out,
                          int64_t cap) {
    if (nthreads < 1) nthreads = 1;
    pairbuf* bufs = run_candidates(A, B, A->nt, method, nthreads);
    int64_t total = 0;
    for (int t = 0; t < nthreads; ++t) {
        for (int64_t i = 0; i < bufs[t].n; ++i) {
            if (out && total < cap) {
                out[2 * total] = bufs[t].p[2 * i];
                out[2 * total + 1] = bufs[t].p[2 * i + 1];
            }
            total++;
        }
        free(bufs[t].p);
    }
    free(bufs);
    return total;
}

/* ------------------------------------------------------------ full path */

typedef struct {
    uint32_t a, b;
    int32_t exp2, nseg, limbs;
    int64_t data; /* offset into the thread's limb pool */
} hitrec;

typedef struct {
    hitrec* h;
    int64_t n, cap;
    uint64_t* pool;
    int64_t pn, pcap;
    int64_t n_no, n_green, n_yellow, n_exact, n_cop, n_deg, n_seg;
    int err;
} hitbuf;

typedef struct {
    const oracle_mesh *A, *B;
    pairbuf* cands;
    int ncbuf;
    int nthreads, use_filter, want_segments;
    hitbuf* out;
} run_ctx;

typedef struct { run_ctx* c; int tid; } run_arg;

static void load_tri(const oracle_mesh* M, uint32_t t, double o[9]) {
    const uint32_t* v = M->tri + 3 * (size_t)t;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) o[3 * i + k] = M->xyz[3 * (size_t)v[i] + k];
}

static void* run_worker(void* p) {
    run_arg* arg = (run_arg*)p;
    run_ctx* c = arg->c;
    hitbuf* hb = &c->out[arg->tid];
    oracle_tt_result* r = (oracle_tt_result*)malloc(sizeof(oracle_tt_result));
    int64_t k = 0;
    for (int bi = 0; bi < c->ncbuf; ++bi) {
        pairbuf* cb = &c->cands[bi];
        for (int64_t i = 0; i < cb->n; ++i, ++k) {
            if (k % c->nthreads != arg->tid) continue;
            uint32_t a = cb->p[2 * i], b = cb->p[2 * i + 1];
            double ta[9], tb[9];
            load_tri(c->A, a, ta);
            load_tri(c->B, b, tb);
            if (oracle_tritri(ta, tb, c->use_filter, r) != 0) { hb->err = 1; continue; }
            int fs = r->info.filter_status;
            if (fs == 0) hb->n_no++;
            else if (fs == 1) hb->n_green++;
            else if (fs == 2) hb->n_yellow++;
            if (fs != 0) hb->n_exact++;
            if (!r->info.intersect) continue;
            if (hb->n == hb->cap) {
                hb->cap = hb->cap ? 2 * hb->cap : 1024;
                hb->h = (hitrec*)realloc(hb->h, sizeof(hitrec) * (size_t)hb->cap);
            }
            hitrec* h = &hb->h[hb->n++];
            h->a = a;
            h->b = b;
            h->exp2 = r->info.exp2;
            h->nseg = c->want_segments ? r->info.nseg : 0;
            h->limbs = r->info.limbs;
            h->data = hb->pn;
            hb->n_cop += r->info.coplanar;
            hb->n_deg += r->info.degenerate;
            if (!c->want_segments) continue;
            hb->n_seg += r->info.nseg;
            int64_t need = (int64_t)r->info.nseg * 8 * r->info.limbs;
            while (hb->pn + need > hb->pcap) {
                hb->pcap = hb->pcap ? 2 * hb->pcap : 1 << 16;
                hb->pool = (uint64_t*)realloc(hb->pool, sizeof(uint64_t) * (size_t)hb->pcap);
            }
            for (int s = 0; s < r->info.nseg; ++s)
                for (int e = 0; e < 8; ++e)
                    memcpy(hb->pool + hb->pn + ((int64_t)s * 8 + e) * r->info.limbs, r->seg[s][e],
                           sizeof(uint64_t) * (size_t)r->info.limbs);
            hb->pn += need;
        }
    }
    free(r);
    return NULL;
}

typedef struct { uint32_t a, b; int t; int64_t i; } sortrec;

static int cmp_sortrec(const void* x, const void* y) {
    const sortrec *p = (const sortrec*)x, *q = (const sortrec*)y;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    if (p->b != q->b) return p->b < q->b ? -1 : 1;
    return 0;
}

int oracle_intersect(const oracle_mesh* A, const oracle_mesh* B, int nthreads, int use_filter, int want_segments,
                     uint32_t a_limit, oracle_result* out) {
    memset(out, 0, sizeof(*out));
    if (nthreads < 1) nthreads = 1;
    uint32_t nA = (a_limit && a_limit < A->nt) ? a_limit : A->nt;
    out->seg_offsets = (uint32_t*)calloc(1, sizeof(uint32_t));
    if (nA == 0 || B->nt == 0) return 0;
    pairbuf* cands = run_candidates(A, B, nA, 0, nthreads);
    for (int t = 0; t < nthreads; ++t) out->n_candidates += cands[t].n;

    run_ctx c;
    c.A = A;
    c.B = B;
    c.cands = cands;
    c.ncbuf = nthreads;
    c.nthreads = nthreads;
    c.use_filter = use_filter;
    c.want_segments = want_segments;
    c.out = (hitbuf*)calloc((size_t)nthreads, sizeof(hitbuf));
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    run_arg* args = (run_arg*)malloc(sizeof(run_arg) * (size_t)nthreads);
    for (int t = 0; t < nthreads; ++t) {
        args[t].c = &c;
        args[t].tid = t;
        pthread_create(&th[t], NULL, run_worker, &args[t]);
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(args);
    for (int t = 0; t < nthreads; ++t) free(cands[t].p);
    free(cands);

    int err = 0;
    int64_t n = 0;
    int limbs = 1;
    for (int t = 0; t < nthreads; ++t) {
        hitbuf* hb = &c.out[t];
        err |= hb->err;
        n += hb->n;
        out->n_filter_no += hb->n_no;
        out->n_filter_green += hb->n_green;
        out->n_filter_yellow += hb->n_yellow;
        out->n_exact += hb->n_exact;
        out->n_coplanar += hb->n_cop;
        out->n_degenerate += hb->n_deg;
        out->n_segments += hb->n_seg;
        for (int64_t i = 0; i < hb->n; ++i)
            if (hb->h[i].limbs > limbs) limbs = hb->h[i].limbs;
    }
    if (!use_filter) out->n_exact = out->n_candidates;
    sortrec* sr = (sortrec*)malloc(sizeof(sortrec) * (size_t)(n ? n : 1));
    int64_t m = 0;
    for (int t = 0; t < nthreads; ++t)
        for (int64_t i = 0; i < c.out[t].n; ++i) {
            sr[m].a = c.out[t].h[i].a;
            sr[m].b = c.out[t].h[i].b;
            sr[m].t = t;
            sr[m].i = i;
            m++;
        }
    qsort(sr, (size_t)n, sizeof(sortrec), cmp_sortrec);
    out->n_pairs = n;
    out->limbs = limbs;
    out->pairs = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (size_t)(n ? n : 1));
    free(out->seg_offsets);
    out->seg_offsets = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n + 1));
    out->seg_exp2 = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    out->seg_data = (uint64_t*)malloc(sizeof(uint64_t) * 8 * (size_t)limbs * (size_t)(out->n_segments ? out->n_segments : 1));
    uint32_t so = 0;
    for (int64_t i = 0; i < n; ++i) {
        const hitrec* h = &c.out[sr[i].t].h[sr[i].i];
        out->pairs[2 * i] = h->a;
        out->pairs[2 * i + 1] = h->b;
        out->seg_offsets[i] = so;
        out->seg_exp2[i] = h->exp2;
        const uint64_t* src = c.out[sr[i].t].pool + h->data;
        for (int s = 0; s < h->nseg; ++s)
            for (int e = 0; e < 8; ++e) {
                const uint64_t* num = src + ((int64_t)s * 8 + e) * h->limbs;
                uint64_t* dst = out->seg_data + (((int64_t)so + s) * 8 + e) * limbs;
                uint64_t ext = (num[h->limbs - 1] >> 63) ? ~(uint64_t)0 : 0;
                for (int l = 0; l < limbs; ++l) dst[l] = l < h->limbs ? num[l] : ext;
            }
        so += (uint32_t)h->nseg;
    }
    out->seg_offsets[n] = so;
    free(sr);
    for (int t = 0; t < nthreads; ++t) {
        free(c.out[t].h);
        free(c.out[t].pool);
    }
    free(c.out);
    return err ? -1 : 0;
}

void oracle_result_free(oracle_result* r) {
    free(r->pairs);
    free(r->seg_offsets);
    free(r->seg_exp2);
    free(r->seg_data);
    memset(r, 0, sizeof(*r));
}
