/*
 * Wisdom: the autotuner's decisions, keyed by a 64-bit FNV-1a hash of the
 * canonical problem description, persisted as text.
 *
 * Same contract as FFTW's wisdom (kernel/planner.c:786-906,
 * api/import-wisdom*.c, api/export-wisdom*.c):
 *   - a preamble identifies what the entries are valid for: here the
 *     fingerprint of the kernel catalog (FFTW uses an md5 of its solver
 *     configuration), and wisdom for another catalog is rejected;
 *   - import is atomic: a string that does not parse completely changes
 *     nothing and returns 0;
 *   - forget_wisdom clears every entry.
 * Format:
 *   (fftw3-amd-gfx950-wisdom #x<catalog>
 *     (#x<problem> <patience> <ndecisions> d0 d1 ...)
 *     ...)
 */
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "engine.h"

#define MAXD 32
/* largest decision value accepted at import: a decision indexes the
   alternatives of one planner choice (at most 16 kernel variants) */
#define MAXC 63

typedef struct {
    unsigned long long h;
    int level, n;
    int c[MAXD];
} went;

static went *g_tab = 0;
static int g_n = 0, g_cap = 0;
static pthread_mutex_t g_wlock = PTHREAD_MUTEX_INITIALIZER;

unsigned long long fk_fnv1a(const void *p, size_t n, unsigned long long h)
{
    const unsigned char *s = (const unsigned char *)p;
    size_t i;
    if (!h) h = 0xcbf29ce484222325ULL;
    for (i = 0; i < n; ++i) {
        h ^= s[i];
        h *= 0x100000001b3ULL;
    }
    return h;
}

/* fingerprint of the kernel catalog (no device needed) */
static unsigned long long catalog_hash(void)
{
    static unsigned long long h = 0;
    if (!h) {
        int n, i;
        const fk_kernel *k = fk_registry(&n);
        /* the planner's decision sequence is part of what wisdom means:
           the tag changes whenever decisions are added or reordered */
        h = fk_fnv1a("fk-catalog-3/planner-chunk", 26, 0);
        for (i = 0; i < n; ++i) {
            int v[21], j, m = 0;
            v[m++] = k[i].prec;
            v[m++] = k[i].L;
            v[m++] = k[i].mode;
            v[m++] = k[i].io;
            v[m++] = k[i].tpl;
            v[m++] = k[i].nl;
            v[m++] = k[i].variant;
            v[m++] = k[i].nrad;
            for (j = 0; j < k[i].nrad && m < 21; ++j) v[m++] = k[i].rad[j];
            h = fk_fnv1a(v, sizeof(int) * m, h);
        }
    }
    return h;
}

static void store_locked(unsigned long long h, int level, const int *c, int n)
{
    int i;
    went *e = 0;
    if (n > MAXD) n = MAXD;
    for (i = 0; i < g_n; ++i)
        if (g_tab[i].h == h) { e = g_tab + i; break; }
    if (!e) {
        if (g_n == g_cap) {
            g_cap = g_cap ? 2 * g_cap : 64;
            g_tab = (went *)realloc(g_tab, sizeof(went) * g_cap);
        }
        e = g_tab + g_n++;
        e->h = h;
        e->level = -1;
    }
    if (level < e->level) return; /* keep the more patient entry */
    e->level = level;
    e->n = n;
    memcpy(e->c, c, sizeof(int) * n);
}

void fk_wisdom_store(unsigned long long h, int level, const int *c, int n)
{
    pthread_mutex_lock(&g_wlock);
    store_locked(h, level, c, n);
    pthread_mutex_unlock(&g_wlock);
}

int fk_wisdom_lookup(unsigned long long h, int *level, int *c, int *n)
{
    int i, found = 0;
    pthread_mutex_lock(&g_wlock);
    for (i = 0; i < g_n; ++i)
        if (g_tab[i].h == h) {
            *level = g_tab[i].level;
            *n = g_tab[i].n;
            memcpy(c, g_tab[i].c, sizeof(int) * g_tab[i].n);
            found = 1;
            break;
        }
    pthread_mutex_unlock(&g_wlock);
    return found;
}

/* appends to a growable string; returns 0 when memory runs out */
typedef struct {
    char *s;
    size_t len, cap;
} sbuf;

static int sb_printf(sbuf *b, const char *fmt, ...)
{
    va_list ap;
    int need;
    va_start(ap, fmt);
    need = vsnprintf(0, 0, fmt, ap);
    va_end(ap);
    if (need < 0) return 0;
    if (b->len + (size_t)need + 1 > b->cap) {
        size_t cap = 2 * b->cap + (size_t)need + 1;
        char *t = (char *)realloc(b->s, cap);
        if (!t) return 0;
        b->s = t;
        b->cap = cap;
    }
    va_start(ap, fmt);
    vsnprintf(b->s + b->len, b->cap - b->len, fmt, ap);
    va_end(ap);
    b->len += (size_t)need;
    return 1;
}

char *fk_wisdom_export(void)
{
    sbuf b = {0, 0, 0};
    int i, j, ok;
    pthread_mutex_lock(&g_wlock);
    ok = sb_printf(&b, "(fftw3-amd-gfx950-wisdom #x%016llx\n", catalog_hash());
    for (i = 0; ok && i < g_n; ++i) {
        ok = sb_printf(&b, "  (#x%016llx %d %d", g_tab[i].h, g_tab[i].level, g_tab[i].n);
        for (j = 0; ok && j < g_tab[i].n; ++j) ok = sb_printf(&b, " %d", g_tab[i].c[j]);
        if (ok) ok = sb_printf(&b, ")\n");
    }
    if (ok) ok = sb_printf(&b, ")\n");
    pthread_mutex_unlock(&g_wlock);
    if (!ok) {
        free(b.s);
        return 0;
    }
    return b.s;
}

static const char *skip_ws(const char *p)
{
    while (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r') ++p;
    return p;
}

int fk_wisdom_import(const char *s)
{
    static const char pre[] = "(fftw3-amd-gfx950-wisdom";
    unsigned long long cat;
    const char *p;
    went *tmp = 0;
    int nt = 0, cap = 0, i, used;
    if (!s) return 0;
    p = skip_ws(s);
    if (strncmp(p, pre, sizeof(pre) - 1)) return 0;
    p += sizeof(pre) - 1;
    if (sscanf(p, " #x%llx%n", &cat, &used) != 1) return 0;
    if (cat != catalog_hash()) return 0; /* wisdom of another kernel catalog */
    p += used;
    for (;;) {
        went e;
        p = skip_ws(p);
        if (*p == ')') break;
        if (*p != '(') goto bad;
        ++p;
        if (sscanf(p, " #x%llx %d %d%n", &e.h, &e.level, &e.n, &used) != 3) goto bad;
        p += used;
        if (e.n < 0 || e.n > MAXD || e.level < 0 || e.level > 3) goto bad;
        for (i = 0; i < e.n; ++i) {
            if (sscanf(p, " %d%n", &e.c[i], &used) != 1 || e.c[i] < 0 || e.c[i] > MAXC)
                goto bad;
            p += used;
        }
        p = skip_ws(p);
        if (*p != ')') goto bad;
        ++p;
        if (nt == cap) {
            cap = cap ? 2 * cap : 16;
            tmp = (went *)realloc(tmp, sizeof(went) * cap);
        }
        tmp[nt++] = e;
    }
    pthread_mutex_lock(&g_wlock);
    for (i = 0; i < nt; ++i) store_locked(tmp[i].h, tmp[i].level, tmp[i].c, tmp[i].n);
    pthread_mutex_unlock(&g_wlock);
    free(tmp);
    return 1;
bad:
    free(tmp);
    return 0;
}

void fk_forget_wisdom(void)
{
    pthread_mutex_lock(&g_wlock);
    g_n = 0;
    pthread_mutex_unlock(&g_wlock);
}

int fk_wisdom_count(void)
{
    int n;
    pthread_mutex_lock(&g_wlock);
    n = g_n;
    pthread_mutex_unlock(&g_wlock);
    return n;
}
