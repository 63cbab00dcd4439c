/*
 * Serving loop with continuous batching.
 *
 * Clients submit transform requests (kind, precision, rank, sizes, howmany,
 * device input) at any time.  A scheduler thread
 *   - keeps one plan per request shape resident (plan cache; twiddle tables
 *     stay cached on the device),
 *   - admits queued requests in arrival order while their output and scratch
 *     fit the HBM budget (a first-fit arena carved from one allocation made
 *     at creation, so admission never calls the allocator),
 *   - dispatches every admitted group as a "round" on its own stream: the
 *     passes of all requests whose kernels belong to the persistent menu go
 *     into per-phase work tables and run as ONE persistent launch per phase
 *     and precision; other requests run their own passes on the same stream,
 *   - keeps several rounds in flight, so requests arriving while others run
 *     are admitted as soon as slots and memory free up and overlap on the
 *     device,
 *   - retires rounds when their event completes, stamping each request's
 *     completion time (latency = completion - submission).
 * The client releases a finished request's output, which returns its bytes
 * to the budget.
 *
 * FFTW has no serving layer; the per-request semantics are those of a
 * plan_many_dft / plan_many_dft_r2c / plan_many_dft_c2r execute on
 * contiguous batches (api/plan-many-dft.c:26-51).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "engine.h"

#define SV_MAXPH 4
#define SV_ALIGN 256

int fk_menu_kind(const fk_menu_entry *menu, int nm, const fk_kernel *k);

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ---------------- first-fit arena ---------------- */

typedef struct blk {
    size_t off, size;
    struct blk *next;
} blk;

typedef struct {
    char *base;
    size_t size, used;
    blk *free;
} arena;

static size_t rup(size_t n) { return (n + SV_ALIGN - 1) / SV_ALIGN * SV_ALIGN; }

static int arena_alloc(arena *a, size_t n, size_t *off)
{
    blk **pp, *b;
    n = rup(n ? n : 1);
    for (pp = &a->free; (b = *pp); pp = &b->next) {
        if (b->size >= n) {
            *off = b->off;
            b->off += n;
            b->size -= n;
            if (!b->size) {
                *pp = b->next;
                free(b);
            }
            a->used += n;
            return 0;
        }
    }
    return -1;
}

static void arena_release(arena *a, size_t off, size_t n)
{
    blk **pp = &a->free, *b, *nb;
    n = rup(n ? n : 1);
    a->used -= n;
    while (*pp && (*pp)->off < off) pp = &(*pp)->next;
    nb = (blk *)malloc(sizeof(blk));
    nb->off = off;
    nb->size = n;
    nb->next = *pp;
    *pp = nb;
    /* coalesce with the successor, then with the predecessor */
    if (nb->next && nb->off + nb->size == nb->next->off) {
        b = nb->next;
        nb->size += b->size;
        nb->next = b->next;
        free(b);
    }
    for (b = a->free; b && b->next; b = b->next) {
        if (b->next == nb && b->off + b->size == nb->off) {
            b->size += nb->size;
            b->next = nb->next;
            free(nb);
            break;
        }
    }
}

/* ---------------- requests, plans, rounds ---------------- */

typedef struct {
    fk_request d;
    long long id, transforms;
    fk_plan *plan;
    size_t foot, out_bytes, out_off, scr_off[FK_MAX_SCRATCH];
    void *out, *scr[FK_MAX_SCRATCH];
    double t_sub, t_done;
    int state; /* 0 queued, 1 in flight, 2 done, 3 released */
} sreq;

typedef struct {
    fk_request key;
    fk_plan *plan;
} pcache;

typedef struct {
    int busy;
    hipStream_t st;
    hipEvent_t ev;
    sreq **reqs;
    int nreq;
    fk_bpass *htab[2][SV_MAXPH];
    fk_bpass *dtab[2][SV_MAXPH];
    unsigned long long *dcount;
} sround;

struct fk_server {
    pthread_t thr;
    pthread_mutex_t mu;
    pthread_cond_t cv_sub, cv_done;
    int stop;
    arena ar;
    size_t budget;
    sreq **reqs;
    long long nreq, cap;
    long long qhead; /* first id possibly still queued */
    pcache *pc;
    int npc, cpc;
    sround *rounds;
    int nrounds, per_round;
    const fk_menu_entry *menu[2];
    int nmenu[2], nt[2], grid[2];
    const void *fn[2];
    unsigned lds[2];
    double *lat;
    long long nlat, clat, transforms_done;
    double t_first, t_last;
};

static size_t csz_(int prec) { return prec == 0 ? 16 : 8; }
static size_t rsz_(int prec) { return prec == 0 ? 8 : 4; }

static int req_valid(const fk_request *r)
{
    int i;
    if (r->kind < 0 || r->kind > 2 || (r->prec != 0 && r->prec != 1)) return 0;
    if (r->rank < 1 || r->rank > 3 || r->howmany < 1) return 0;
    for (i = 0; i < r->rank; ++i)
        if (r->n[i] < 1) return 0;
    if (!r->in || ((uintptr_t)r->in % (2 * rsz_(r->prec)))) return 0;
    return 1;
}

static int same_shape(const fk_request *a, const fk_request *b)
{
    int i;
    if (a->kind != b->kind || a->prec != b->prec || a->rank != b->rank || a->howmany != b->howmany)
        return 0;
    if (a->kind == FK_C2C && a->sign != b->sign) return 0;
    for (i = 0; i < a->rank; ++i)
        if (a->n[i] != b->n[i]) return 0;
    return 1;
}

/* plan of a contiguous batch: in and out are distinct placeholders (never
   dereferenced by FFTW_ESTIMATE planning); execution binds real buffers */
static fk_plan *make_plan(const fk_request *r)
{
    fk_problem pb;
    long long nl = r->n[r->rank - 1], nh = nl / 2 + 1;
    int i;
    memset(&pb, 0, sizeof(pb));
    pb.kind = r->kind;
    pb.prec = r->prec;
    pb.sign = r->kind == FK_C2C ? (r->sign == -1 ? -1 : 1) : (r->kind == FK_R2C ? -1 : 1);
    pb.flags = FK_ESTIMATE;
    pb.rank = r->rank;
    /* row-major strides; the real side of r2c/c2r is unpadded, the
       complex side has n/2+1 along the last dim */
    {
        long long si = 1, so = 1;
        for (i = r->rank - 1; i >= 0; --i) {
            long long ni = r->n[i], no = r->n[i];
            if (i == r->rank - 1 && r->kind == FK_R2C) no = nh;
            if (i == r->rank - 1 && r->kind == FK_C2R) ni = nh;
            pb.dims[i].n = r->n[i];
            pb.dims[i].is = si;
            pb.dims[i].os = so;
            si *= ni;
            so *= no;
        }
        pb.vrank = 1;
        pb.vdims[0].n = r->howmany;
        pb.vdims[0].is = si;
        pb.vdims[0].os = so;
    }
    pb.in = (void *)(uintptr_t)4096;
    pb.out = (void *)(uintptr_t)8192;
    return fk_plan_create(&pb);
}

static size_t out_bytes_of(const fk_request *r)
{
    long long N = 1, nl = r->n[r->rank - 1];
    int i;
    for (i = 0; i < r->rank; ++i) N *= r->n[i];
    if (r->kind == FK_C2C) return (size_t)N * r->howmany * csz_(r->prec);
    if (r->kind == FK_R2C) return (size_t)(N / nl) * (nl / 2 + 1) * r->howmany * csz_(r->prec);
    return (size_t)N * r->howmany * rsz_(r->prec);
}

/* plan cache lookup/creation; called without the server lock held */
static fk_plan *get_plan(fk_server *s, const fk_request *r)
{
    fk_plan *p = 0;
    int i;
    pthread_mutex_lock(&s->mu);
    for (i = 0; i < s->npc; ++i)
        if (same_shape(&s->pc[i].key, r)) {
            p = s->pc[i].plan;
            break;
        }
    pthread_mutex_unlock(&s->mu);
    if (p) return p;
    p = make_plan(r);
    if (!p) return 0;
    pthread_mutex_lock(&s->mu);
    for (i = 0; i < s->npc; ++i)
        if (same_shape(&s->pc[i].key, r)) { /* lost a race: keep the first */
            fk_plan *q = s->pc[i].plan;
            pthread_mutex_unlock(&s->mu);
            fk_plan_destroy(p);
            return q;
        }
    if (s->npc == s->cpc) {
        s->cpc = s->cpc ? 2 * s->cpc : 32;
        s->pc = (pcache *)realloc(s->pc, sizeof(pcache) * s->cpc);
    }
    s->pc[s->npc].key = *r;
    s->pc[s->npc].key.in = 0;
    s->pc[s->npc].plan = p;
    ++s->npc;
    pthread_mutex_unlock(&s->mu);
    return p;
}

static int in_menu(fk_server *s, const fk_plan *P)
{
    int i;
    if (P->npass < 1 || P->npass > SV_MAXPH) return 0;
    for (i = 0; i < P->npass; ++i) {
        const fk_pass *ps = P->pass + i;
        if (ps->type != FK_PASS_STOCKHAM) return 0;
        if (fk_menu_kind(s->menu[P->prec], s->nmenu[P->prec], ps->k) < 0) return 0;
    }
    return 1;
}

/* build the round's work tables and enqueue it (server lock NOT held) */
static int launch_round(fk_server *s, sround *r)
{
    int np[2][SV_MAXPH];
    long long tot[2][SV_MAXPH];
    int i, j, p, ph, rc = 0;
    memset(np, 0, sizeof(np));
    memset(tot, 0, sizeof(tot));
    for (i = 0; i < r->nreq; ++i) {
        sreq *q = r->reqs[i];
        fk_plan *P = q->plan;
        void *buf[FK_NBUF];
        if (!in_menu(s, P)) continue;
        buf[FK_BUF_IN] = (void *)q->d.in;
        buf[FK_BUF_OUT] = q->out;
        buf[FK_BUF_IN2] = buf[FK_BUF_OUT2] = 0;
        for (j = 0; j < FK_MAX_SCRATCH; ++j) buf[FK_BUF_S0 + j] = q->scr[j];
        for (j = 0; j < P->npass; ++j) {
            const fk_pass *ps = P->pass + j;
            int kind = fk_menu_kind(s->menu[P->prec], s->nmenu[P->prec], ps->k);
            fk_bpass *e = r->htab[P->prec][j] + np[P->prec][j]++;
            memset(e, 0, sizeof(*e));
            e->a = ps->a;
            e->a.in = (const char *)buf[ps->ibuf] + ps->ioff;
            e->a.out = (char *)buf[ps->obuf] + ps->ooff;
            e->kind = kind;
            e->tiles = (ps->a.nlines + s->menu[P->prec][kind].nl - 1) / s->menu[P->prec][kind].nl;
            e->first = tot[P->prec][j];
            tot[P->prec][j] += e->tiles;
        }
    }
    for (ph = 0; ph < SV_MAXPH && !rc; ++ph)
        for (p = 0; p < 2 && !rc; ++p)
            if (np[p][ph])
                rc = (int)hipMemcpyAsync(r->dtab[p][ph], r->htab[p][ph],
                                         sizeof(fk_bpass) * np[p][ph], hipMemcpyHostToDevice, r->st);
    for (ph = 0; ph < SV_MAXPH && !rc; ++ph)
        for (p = 0; p < 2 && !rc; ++p) {
            void *args[4];
            unsigned long long *cnt = r->dcount + 2 * ph + p;
            int n = np[p][ph];
            long long t = tot[p][ph];
            fk_bpass *tab = r->dtab[p][ph];
            if (!t) continue;
            rc = fkm_reset_counters(cnt, 1, r->st);
            args[0] = &tab;
            args[1] = &n;
            args[2] = &t;
            args[3] = &cnt;
            if (!rc)
                rc = (int)hipLaunchKernel(s->fn[p], (dim3){(unsigned)s->grid[p], 1, 1},
                                          (dim3){(unsigned)s->nt[p], 1, 1}, args, s->lds[p], r->st);
        }
    for (i = 0; i < r->nreq && !rc; ++i) {
        sreq *q = r->reqs[i];
        if (in_menu(s, q->plan)) continue;
        rc = fk_enqueue_bound(q->plan, (void *)q->d.in, q->out, q->scr, r->st);
    }
    if (!rc) rc = (int)hipEventRecord(r->ev, r->st);
    return rc;
}

static void free_scratch_locked(fk_server *s, sreq *q)
{
    int j;
    for (j = 0; j < FK_MAX_SCRATCH; ++j)
        if (q->scr[j]) {
            arena_release(&s->ar, q->scr_off[j], q->plan->scratch_bytes[j]);
            q->scr[j] = 0;
        }
}

static void *sched_main(void *arg)
{
    fk_server *s = (fk_server *)arg;
    int rc;
    pthread_mutex_lock(&s->mu);
    while (!s->stop) {
        int progressed = 0, busy = 0, i;
        /* 1. retire finished rounds */
        for (i = 0; i < s->nrounds; ++i) {
            sround *r = s->rounds + i;
            hipError_t e;
            if (!r->busy) continue;
            e = hipEventQuery(r->ev);
            if (e == hipErrorNotReady) {
                busy = 1;
                continue;
            }
            {
                double t = now_s();
                int k;
                for (k = 0; k < r->nreq; ++k) {
                    sreq *q = r->reqs[k];
                    q->state = e == hipSuccess ? 2 : -1;
                    q->t_done = t;
                    free_scratch_locked(s, q);
                    if (s->nlat == s->clat) {
                        s->clat = s->clat ? 2 * s->clat : 1024;
                        s->lat = (double *)realloc(s->lat, sizeof(double) * s->clat);
                    }
                    s->lat[s->nlat++] = (t - q->t_sub) * 1e3;
                    s->transforms_done += q->transforms;
                    s->t_last = t;
                }
            }
            r->busy = 0;
            r->nreq = 0;
            progressed = 1;
        }
        if (progressed) pthread_cond_broadcast(&s->cv_done);
        /* 2. admit queued requests into a free round, in arrival order */
        for (i = 0; i < s->nrounds; ++i)
            if (!s->rounds[i].busy) break;
        if (i < s->nrounds) {
            sround *r = s->rounds + i;
            long long id;
            while (s->qhead < s->nreq && s->reqs[s->qhead]->state != 0) ++s->qhead;
            for (id = s->qhead; id < s->nreq && r->nreq < s->per_round; ++id) {
                sreq *q = s->reqs[id];
                int j, ok = 1;
                if (q->state != 0) continue;
                if (s->ar.size - s->ar.used < q->foot) break; /* head of line waits */
                if (arena_alloc(&s->ar, q->out_bytes, &q->out_off)) break;
                q->out = s->ar.base + q->out_off;
                for (j = 0; j < FK_MAX_SCRATCH && ok; ++j) {
                    q->scr[j] = 0;
                    if (!q->plan->scratch_bytes[j]) continue;
                    if (arena_alloc(&s->ar, q->plan->scratch_bytes[j], &q->scr_off[j])) ok = 0;
                    else q->scr[j] = s->ar.base + q->scr_off[j];
                }
                if (!ok) { /* fragmentation: undo and wait for releases */
                    free_scratch_locked(s, q);
                    arena_release(&s->ar, q->out_off, q->out_bytes);
                    q->out = 0;
                    break;
                }
                q->state = 1;
                r->reqs[r->nreq++] = q;
            }
            if (r->nreq) {
                r->busy = 1;
                pthread_mutex_unlock(&s->mu);
                rc = launch_round(s, r);
                pthread_mutex_lock(&s->mu);
                if (rc) {
                    fprintf(stderr, "fftw3_amd server: launch failed: %s\n",
                            hipGetErrorString((hipError_t)rc));
                    (void)hipGetLastError();
                }
                progressed = 1;
                busy = 1;
            }
        }
        if (!progressed) {
            if (busy) {
                struct timespec ts = {0, 20000};
                pthread_mutex_unlock(&s->mu);
                nanosleep(&ts, 0);
                pthread_mutex_lock(&s->mu);
            } else {
                struct timespec ts;
                clock_gettime(CLOCK_REALTIME, &ts);
                ts.tv_nsec += 1000000;
                if (ts.tv_nsec >= 1000000000) {
                    ts.tv_sec++;
                    ts.tv_nsec -= 1000000000;
                }
                pthread_cond_timedwait(&s->cv_sub, &s->mu, &ts);
            }
        }
    }
    pthread_mutex_unlock(&s->mu);
    return 0;
}

/* ---------------- public entry points ---------------- */

void fk_server_destroy(fk_server *s)
{
    int i, p, ph;
    long long k;
    blk *b, *nb;
    if (!s) return;
    if (s->thr) {
        pthread_mutex_lock(&s->mu);
        s->stop = 1;
        pthread_cond_broadcast(&s->cv_sub);
        pthread_mutex_unlock(&s->mu);
        pthread_join(s->thr, 0);
    }
    for (i = 0; i < s->nrounds; ++i) {
        sround *r = s->rounds + i;
        if (r->st) (void)hipStreamSynchronize(r->st);
        for (p = 0; p < 2; ++p)
            for (ph = 0; ph < SV_MAXPH; ++ph) {
                if (r->htab[p][ph]) (void)hipHostFree(r->htab[p][ph]);
                fk_free(r->dtab[p][ph]);
            }
        fk_free(r->dcount);
        if (r->ev) (void)hipEventDestroy(r->ev);
        if (r->st) (void)hipStreamDestroy(r->st);
        free(r->reqs);
    }
    free(s->rounds);
    for (i = 0; i < s->npc; ++i) fk_plan_destroy(s->pc[i].plan);
    free(s->pc);
    for (k = 0; k < s->nreq; ++k) free(s->reqs[k]);
    free(s->reqs);
    free(s->lat);
    fk_free(s->ar.base);
    for (b = s->ar.free; b; b = nb) {
        nb = b->next;
        free(b);
    }
    pthread_mutex_destroy(&s->mu);
    pthread_cond_destroy(&s->cv_sub);
    pthread_cond_destroy(&s->cv_done);
    free(s);
}

fk_server *fk_server_create(size_t budget, int max_rounds, int per_round)
{
    fk_server *s;
    int i, p, ph, ncu = 256;
    if (fk_init() || budget < SV_ALIGN) return 0;
    if (max_rounds < 1) max_rounds = 4;
    if (per_round < 1) per_round = 64;
    s = (fk_server *)calloc(1, sizeof(fk_server));
    pthread_mutex_init(&s->mu, 0);
    pthread_cond_init(&s->cv_sub, 0);
    pthread_cond_init(&s->cv_done, 0);
    s->budget = rup(budget);
    s->ar.base = (char *)fk_malloc(s->budget);
    if (!s->ar.base) goto fail;
    s->ar.size = s->budget;
    s->ar.free = (blk *)calloc(1, sizeof(blk));
    s->ar.free->size = s->budget;
    s->nrounds = max_rounds;
    s->per_round = per_round;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess ||
        ncu < 1)
        ncu = 256;
    for (p = 0; p < 2; ++p) {
        int per_cu = 0;
        s->menu[p] = fk_persist_menu(p, &s->nmenu[p], &s->fn[p], &s->lds[p], &s->nt[p]);
        if (s->lds[p] > 65536 &&
            hipFuncSetAttribute(s->fn[p], hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)s->lds[p]) != hipSuccess)
            goto fail;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, s->fn[p], s->nt[p], s->lds[p]) !=
                hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        s->grid[p] = ncu * per_cu;
    }
    s->rounds = (sround *)calloc(max_rounds, sizeof(sround));
    for (i = 0; i < max_rounds; ++i) {
        sround *r = s->rounds + i;
        if (hipStreamCreateWithFlags(&r->st, hipStreamNonBlocking) != hipSuccess) goto fail;
        if (hipEventCreateWithFlags(&r->ev, hipEventDisableTiming) != hipSuccess) goto fail;
        r->reqs = (sreq **)calloc(per_round, sizeof(sreq *));
        r->dcount = (unsigned long long *)fk_malloc(sizeof(unsigned long long) * 2 * SV_MAXPH);
        if (!r->dcount) goto fail;
        for (p = 0; p < 2; ++p)
            for (ph = 0; ph < SV_MAXPH; ++ph) {
                if (hipHostMalloc((void **)&r->htab[p][ph], sizeof(fk_bpass) * per_round,
                                  hipHostMallocDefault) != hipSuccess)
                    goto fail;
                r->dtab[p][ph] = (fk_bpass *)fk_malloc(sizeof(fk_bpass) * per_round);
                if (!r->dtab[p][ph]) goto fail;
            }
    }
    if (pthread_create(&s->thr, 0, sched_main, s)) {
        s->thr = 0;
        goto fail;
    }
    return s;
fail:
    (void)hipGetLastError();
    fk_server_destroy(s);
    return 0;
}

int fk_server_prepare(fk_server *s, const fk_request *r)
{
    if (!s || !req_valid(r)) return -1;
    return get_plan(s, r) ? 0 : -1;
}

long long fk_server_submit(fk_server *s, const fk_request *r)
{
    sreq *q;
    fk_plan *P;
    long long id;
    int j;
    if (!s || !req_valid(r)) return -1;
    P = get_plan(s, r);
    if (!P) return -1;
    q = (sreq *)calloc(1, sizeof(sreq));
    q->d = *r;
    q->plan = P;
    q->out_bytes = out_bytes_of(r);
    q->foot = rup(q->out_bytes);
    for (j = 0; j < FK_MAX_SCRATCH; ++j)
        if (P->scratch_bytes[j]) q->foot += rup(P->scratch_bytes[j]);
    q->transforms = r->howmany;
    if (q->foot > s->budget) { /* can never be admitted */
        free(q);
        return -1;
    }
    pthread_mutex_lock(&s->mu);
    if (s->nreq == s->cap) {
        s->cap = s->cap ? 2 * s->cap : 1024;
        s->reqs = (sreq **)realloc(s->reqs, sizeof(sreq *) * s->cap);
    }
    id = s->nreq++;
    q->id = id;
    q->t_sub = now_s();
    if (!s->t_first) s->t_first = q->t_sub;
    s->reqs[id] = q;
    pthread_cond_signal(&s->cv_sub);
    pthread_mutex_unlock(&s->mu);
    return id;
}

int fk_server_poll(fk_server *s, long long id)
{
    int st;
    if (!s) return -1;
    pthread_mutex_lock(&s->mu);
    st = id >= 0 && id < s->nreq ? s->reqs[id]->state : -2;
    pthread_mutex_unlock(&s->mu);
    if (st == -2 || st == 3) return -1;
    return st == 2 ? 1 : (st < 0 ? -1 : 0);
}

int fk_server_wait(fk_server *s, long long id, void **out, double *lat_ms)
{
    sreq *q;
    int rc = 0;
    if (!s) return -1;
    pthread_mutex_lock(&s->mu);
    if (id < 0 || id >= s->nreq) {
        pthread_mutex_unlock(&s->mu);
        return -1;
    }
    q = s->reqs[id];
    while (q->state == 0 || q->state == 1) pthread_cond_wait(&s->cv_done, &s->mu);
    if (q->state != 2) rc = -1;
    if (out) *out = q->out;
    if (lat_ms) *lat_ms = (q->t_done - q->t_sub) * 1e3;
    pthread_mutex_unlock(&s->mu);
    return rc;
}

int fk_server_release(fk_server *s, long long id)
{
    sreq *q;
    if (!s) return -1;
    pthread_mutex_lock(&s->mu);
    if (id < 0 || id >= s->nreq || (s->reqs[id]->state != 2 && s->reqs[id]->state != -1)) {
        pthread_mutex_unlock(&s->mu);
        return -1;
    }
    q = s->reqs[id];
    if (q->out) arena_release(&s->ar, q->out_off, q->out_bytes);
    q->out = 0;
    q->state = 3;
    pthread_cond_signal(&s->cv_sub);
    pthread_mutex_unlock(&s->mu);
    return 0;
}

static int cmpd(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y;
}

int fk_server_stats(fk_server *s, double *p50, double *p99, long long *done, long long *transforms,
                    double *elapsed_s, size_t *hbm_in_use)
{
    double *c;
    long long n;
    if (!s) return -1;
    pthread_mutex_lock(&s->mu);
    n = s->nlat;
    c = (double *)malloc(sizeof(double) * (n ? n : 1));
    memcpy(c, s->lat, sizeof(double) * n);
    if (done) *done = n;
    if (transforms) *transforms = s->transforms_done;
    if (elapsed_s) *elapsed_s = s->t_last > s->t_first ? s->t_last - s->t_first : 0;
    if (hbm_in_use) *hbm_in_use = s->ar.used;
    pthread_mutex_unlock(&s->mu);
    qsort(c, n, sizeof(double), cmpd);
    if (p50) *p50 = n ? c[(n - 1) / 2] : 0;
    if (p99) *p99 = n ? c[(long long)ceil(0.99 * n) - 1] : 0;
    free(c);
    return 0;
}
