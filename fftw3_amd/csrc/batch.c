/*
 * Batched execution of uneven transform sizes in persistent launches.
 *
 * A batch holds groups of 1-D c2c transforms of different lengths.  Every
 * group is planned with the ordinary planner (so FFTW_MEASURE picks its
 * geometry on the device, and wisdom applies); pass i of every group whose
 * passes all belong to the persistent kernel's menu becomes an entry of the
 * phase-i work table.  Each phase is ONE launch of the persistent kernel:
 * workgroups pull tiles from a device counter until the table is drained,
 * so transforms of different sizes share the machine without per-size
 * launches or tails.  Phase i+1 starts after phase i (multi-pass lengths
 * need their column pass complete).  Groups outside the menu run their own
 * passes after the phases.  The whole sequence (counter resets, phase
 * launches, remaining passes) is captured once into a hipGraph and replayed
 * by execute.
 *
 * FFTW's counterpart is the vector loop over howmany
 * (dft/vrank-geq1.c:54-65, threads/dft-vrank-geq1.c:59-70), which only
 * batches equal sizes.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "engine.h"

#define MAXPH 4

int fkm_reset_counters(unsigned long long *c, int n, hipStream_t st);
int fk_enqueue(fk_plan *p, void *in, void *out, hipStream_t st);

typedef struct fk_batch {
    int prec, ngroups, nphase;
    fk_plan **plans;
    void **in, **out;
    int *persist;
    fk_bpass *htab[MAXPH];
    int npass[MAXPH];
    long long total[MAXPH];
    fk_bpass *dtab[MAXPH];
    unsigned long long *dcount;
    const void *fn;
    unsigned lds;
    int nt, grid;
    hipStream_t st, st2;
    hipEvent_t ev_fork, ev_join;
    hipGraph_t graph;
    hipGraphExec_t exec;
    long long ntransforms;
    double bytes;
} fk_batch;

static int same_rad(const fk_menu_entry *m, const fk_kernel *k)
{
    int i;
    if (m->nrad != k->nrad) return 0;
    for (i = 0; i < k->nrad; ++i)
        if (m->rad[i] != k->rad[i]) return 0;
    return 1;
}

int fk_menu_kind(const fk_menu_entry *menu, int nm, const fk_kernel *k)
{
    int i;
    for (i = 0; i < nm; ++i)
        if (menu[i].prec == k->prec && menu[i].L == k->L && menu[i].mode == k->mode &&
            menu[i].io == k->io && same_rad(menu + i, k))
            return i;
    return -1;
}

static int enqueue_all(fk_batch *b)
{
    int p, g, rc = 0, nside = 0;
    for (g = 0; g < b->ngroups; ++g) nside += !b->persist[g];
    /* fork: groups with their own launches run on the side stream,
       concurrently with the persistent phases */
    if (nside) {
        rc = (int)hipEventRecord(b->ev_fork, b->st);
        if (!rc) rc = (int)hipStreamWaitEvent(b->st2, b->ev_fork, 0);
        for (g = 0; g < b->ngroups && !rc; ++g)
            if (!b->persist[g]) rc = fk_enqueue(b->plans[g], b->in[g], b->out[g], b->st2);
        if (!rc) rc = (int)hipEventRecord(b->ev_join, b->st2);
    }
    for (p = 0; p < b->nphase && !rc; ++p) {
        void *args[4];
        int np = b->npass[p];
        if (!b->total[p]) continue;
        rc = fkm_reset_counters(b->dcount + p, 1, b->st);
        args[0] = &b->dtab[p];
        args[1] = &np;
        args[2] = &b->total[p];
        {
            unsigned long long *cp = b->dcount + p;
            args[3] = &cp;
            if (!rc)
                rc = (int)hipLaunchKernel(b->fn, (dim3){(unsigned)b->grid, 1, 1},
                                          (dim3){(unsigned)b->nt, 1, 1}, args, b->lds, b->st);
        }
    }
    /* join */
    if (nside && !rc) rc = (int)hipStreamWaitEvent(b->st, b->ev_join, 0);
    return rc;
}

void fk_batch_destroy(fk_batch *b)
{
    int i;
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (i = 0; i < MAXPH; ++i) {
        free(b->htab[i]);
        fk_free(b->dtab[i]);
    }
    fk_free(b->dcount);
    for (i = 0; i < b->ngroups; ++i) fk_plan_destroy(b->plans ? b->plans[i] : 0);
    free(b->plans);
    free(b->persist);
    free(b->in);
    free(b->out);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_join) (void)hipEventDestroy(b->ev_join);
    if (b->st2) (void)hipStreamDestroy(b->st2);
    if (b->st) (void)hipStreamDestroy(b->st);
    free(b);
}

fk_batch *fk_batch_create(int prec, int ngroups, const int *n, const int *howmany,
                          void *const *in, void *const *out, int sign, unsigned flags)
{
    fk_batch *b;
    const fk_menu_entry *menu;
    int nm, g, i, p, ncu = 0;
    if (ngroups <= 0 || fk_init()) return 0;
    b = (fk_batch *)calloc(1, sizeof(fk_batch));
    b->prec = prec;
    b->ngroups = ngroups;
    b->plans = (fk_plan **)calloc(ngroups, sizeof(fk_plan *));
    b->persist = (int *)calloc(ngroups, sizeof(int));
    b->in = (void **)calloc(ngroups, sizeof(void *));
    b->out = (void **)calloc(ngroups, sizeof(void *));
    menu = fk_persist_menu(prec, &nm, &b->fn, &b->lds, &b->nt);
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) goto fail;
    if (hipStreamCreateWithFlags(&b->st2, hipStreamNonBlocking) != hipSuccess) goto fail;
    if (hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess) goto fail;
    if (hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess) goto fail;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess ||
        ncu <= 0)
        ncu = 256;
    if (b->lds > 65536 &&
        hipFuncSetAttribute(b->fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->lds) !=
            hipSuccess)
        goto fail;
    for (g = 0; g < ngroups; ++g) {
        fk_problem pb;
        if (n[g] <= 0 || howmany[g] < 0) goto fail;
        memset(&pb, 0, sizeof(pb));
        pb.kind = FK_C2C;
        pb.prec = prec;
        pb.sign = sign == -1 ? -1 : 1;
        pb.flags = flags;
        pb.rank = 1;
        pb.dims[0].n = n[g];
        pb.dims[0].is = pb.dims[0].os = 1;
        pb.vrank = 1;
        pb.vdims[0].n = howmany[g];
        pb.vdims[0].is = pb.vdims[0].os = n[g];
        pb.in = in[g];
        pb.out = out[g];
        b->plans[g] = fk_plan_create(&pb);
        if (!b->plans[g]) goto fail;
        b->plans[g]->stream = b->st;
        b->in[g] = in[g];
        b->out[g] = out[g];
        b->ntransforms += howmany[g];
        b->bytes += 2.0 * (double)n[g] * howmany[g] * (prec == 0 ? 16 : 8);
        /* all passes in the menu, at most MAXPH of them, and the group too
           small to fill the GPU with its own tuned launches?  Groups that
           fill it alone (every pass >= 4 tiles per CU) keep their own
           kernels on a parallel branch. */
        b->persist[g] = b->plans[g]->npass > 0 && b->plans[g]->npass <= MAXPH;
        {
            int big = 1;
            for (i = 0; i < b->plans[g]->npass; ++i) {
                const fk_pass *ps = b->plans[g]->pass + i;
                if (ps->type != FK_PASS_STOCKHAM || fk_menu_kind(menu, nm, ps->k) < 0)
                    b->persist[g] = 0;
                if (ps->type != FK_PASS_STOCKHAM ||
                    (ps->a.nlines + ps->k->nl - 1) / ps->k->nl < 4LL * ncu)
                    big = 0;
            }
            if (big) b->persist[g] = 0;
        }
    }
    /* work tables, one per phase */
    for (g = 0; g < ngroups; ++g) {
        fk_plan *P = b->plans[g];
        void *buf[FK_NBUF];
        if (!b->persist[g]) continue;
        buf[FK_BUF_IN] = in[g];
        buf[FK_BUF_OUT] = out[g];
        buf[FK_BUF_IN2] = buf[FK_BUF_OUT2] = 0;
        for (i = 0; i < FK_MAX_SCRATCH; ++i) buf[FK_BUF_S0 + i] = P->scratch[i];
        for (i = 0; i < P->npass; ++i) {
            const fk_pass *ps = P->pass + i;
            int kind = fk_menu_kind(menu, nm, ps->k);
            fk_bpass *e;
            p = i;
            b->htab[p] = (fk_bpass *)realloc(b->htab[p], sizeof(fk_bpass) * (b->npass[p] + 1));
            e = b->htab[p] + b->npass[p]++;
            memset(e, 0, sizeof(*e));
            e->a = ps->a;
            e->a.in = (const char *)buf[ps->ibuf] + ps->ioff;
            e->a.out = (char *)buf[ps->obuf] + ps->ooff;
            e->kind = kind;
            e->tiles = (ps->a.nlines + menu[kind].nl - 1) / menu[kind].nl;
            e->first = b->total[p];
            b->total[p] += e->tiles;
            if (b->nphase < p + 1) b->nphase = p + 1;
        }
    }
    for (p = 0; p < b->nphase; ++p) {
        size_t sz = sizeof(fk_bpass) * (size_t)b->npass[p];
        b->dtab[p] = (fk_bpass *)fk_malloc(sz ? sz : 1);
        if (!b->dtab[p] || (sz && hipMemcpy(b->dtab[p], b->htab[p], sz, hipMemcpyHostToDevice) !=
                                      hipSuccess))
            goto fail;
    }
    b->dcount = (unsigned long long *)fk_malloc(sizeof(unsigned long long) * MAXPH);
    if (!b->dcount) goto fail;
    {
        /* one resident wave of persistent workgroups */
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, b->fn, b->nt, b->lds) !=
                hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        b->grid = ncu * per_cu;
    }
    /* record the launch sequence once */
    if (hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal) != hipSuccess) goto fail;
    if (enqueue_all(b)) {
        hipGraph_t gtmp = 0;
        (void)hipStreamEndCapture(b->st, &gtmp);
        if (gtmp) (void)hipGraphDestroy(gtmp);
        goto fail;
    }
    if (hipStreamEndCapture(b->st, &b->graph) != hipSuccess) goto fail;
    if (hipGraphInstantiate(&b->exec, b->graph, 0, 0, 0) != hipSuccess) goto fail;
    return b;
fail:
    (void)hipGetLastError();
    fk_batch_destroy(b);
    return 0;
}

int fk_batch_execute(fk_batch *b, int use_graph)
{
    int rc;
    if (!b) return -1;
    rc = use_graph ? (int)hipGraphLaunch(b->exec, b->st) : enqueue_all(b);
    if (!rc && !fk_get_async()) rc = (int)hipStreamSynchronize(b->st);
    return rc;
}

hipStream_t fk_batch_stream(fk_batch *b) { return b ? b->st : 0; }

int fk_batch_info(fk_batch *b, char *buf, int len)
{
    int g, np = 0, off = 0;
    if (!b) return -1;
    for (g = 0; g < b->ngroups; ++g) np += b->persist[g];
    off += snprintf(buf + off, len - off,
                    "batch: %d groups (%d in persistent phases, %d on the parallel branch), "
                    "%lld transforms, %d phases",
                    b->ngroups, np, b->ngroups - np, b->ntransforms, b->nphase);
    for (g = 0; g < b->nphase && off < len; ++g)
        off += snprintf(buf + off, len - off, "; phase %d: %d passes %lld tiles", g, b->npass[g],
                        b->total[g]);
    if (off < len)
        snprintf(buf + off, len - off, "; grid %d x %d threads, lds %u", b->grid, b->nt, b->lds);
    return 0;
}
