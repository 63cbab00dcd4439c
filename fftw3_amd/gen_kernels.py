"""Generate the compile-time Stockham kernel catalog for gfx950.

Writes fftw3_amd/csrc/gen/fk_inst_<i>.hip (template instantiations split for
parallel compilation) and fftw3_amd/csrc/gen/fk_registry.hip (the table the
host planner searches).  Each entry fixes: precision, length L, radix
sequence, threads per line (TPL), lines per workgroup (NL), thread mapping
(ROW / COL) and the IO kind (plain, transposed store, r2c, c2r, Bluestein).

The catalog plays the role of FFTW's codelet lists
(dft/scalar/codelets/Makefile.am:19-52, dft/simd/codlist.mk:15-78): it says
which lengths run as one fused pass; every other length is decomposed by the
host planner into several passes or a Bluestein convolution.
"""
import itertools
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "csrc", "gen")

LDS_MAX = 160 * 1024          # per workgroup on gfx950
LDS_TARGET = 64 * 1024        # keep >= 2 workgroups per CU when possible
CSZ = {"double": 16, "float": 8}
PREC = {"double": 0, "float": 1}
RADICES = [16, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2]
XR_ROWS = (8, 4)              # rows per tile of the cross-row real kernels
# row-mapped plain passes of long lines also offered pipelined: each
# workgroup runs several tiles and loads the next tile's first-stage inputs
# during the current tile's stages (needs one first-stage butterfly per
# thread).  The default for FP64 2048 and 4096 (measured 25% and 8% faster
# on batches beyond the memory-side cache), a tuner alternative for FP32
# 4096; FP32 8192 measured slower.
PIPE_DEFAULT = {("double", 2048), ("double", 4096)}
PIPE_SIZES = PIPE_DEFAULT | {("float", 4096)}
# row-mapped plain passes also offered with the split (real, then
# imaginary) LDS exchange, where it measured faster than the other
# variants on a rotated 2048 x 2048 FP64 batch (4096 FP64 and 8192 FP32
# measured slower: their registers, not their LDS, limit residency)
XS_SIZES = {("double", 2048)}


def factorizations(n, maxr=16):
    """All ordered-by-size multisets of radices from RADICES with product n."""
    out = []

    def rec(m, lo, acc):
        if m == 1:
            out.append(list(acc))
            return
        for r in RADICES:
            if r > lo or m % r:
                continue
            rec(m // r, r, acc + [r])

    rec(n, maxr, [])
    return out


def choose_radices(n):
    """Fewest stages, then most balanced (largest smallest radix); larger
    radices first (stage 0 reads HBM with stride L/R0)."""
    fs = factorizations(n)
    if not fs:
        return None
    fs.sort(key=lambda f: (len(f), -min(f), -sum(1 for r in f if r in (16, 8, 4, 2))))
    return fs[0]


def geometry(prec, L, rads, mode, io, tpl_div=1, nl_mul=1.0):
    """Default geometry: one butterfly of the largest radix per thread and
    stage (tpl = L / Rmax); tpl_div > 1 gives each thread tpl_div of them;
    nl_mul scales the lines per workgroup of the default choice."""
    rmax = max(rads)
    tpl = max(1, L // rmax // tpl_div)
    pad = rads[0]
    lp = L + L // pad + 2
    csz = CSZ[prec]
    if mode == "ROW":
        per_line = lp * csz
        nl = max(1, 256 // tpl)
        while nl > 1 and nl * per_line > LDS_TARGET:
            nl //= 2
        if io == "TOUT":
            # transposed stores write runs of nl elements: aim for 128 B
            # runs within the thread and LDS limits
            while nl * csz < 128 and 2 * nl * tpl <= 1024 and 2 * nl * per_line + 16 * nl <= LDS_MAX:
                nl *= 2
        if nl * per_line + 8 * nl > LDS_MAX:
            return None
    else:
        per_line = (L + 2) * csz
        # lines of a tile are adjacent in memory: aim for 256 B runs (of
        # reals for the real-to-real kernels, which load one real per line)
        nl = 256 // (csz // 2 if io in ("R2C", "C2R", "DCT4") else csz)
        while nl > 1 and (nl * per_line > LDS_MAX - 1024 or nl * tpl > 1024):
            nl //= 2
        # small transforms: more lines per workgroup for occupancy
        while nl * tpl < 256 and (nl * 2) * per_line <= LDS_TARGET and nl < 64:
            nl *= 2
    nl = max(1, int(nl * nl_mul))
    lds = (lp * nl if mode == "ROW" else (L + 2) * nl) * csz + 8 * nl
    if lds > LDS_MAX or tpl * nl > 1024:
        return None
    return dict(tpl=tpl, nl=nl, pad=pad, lds=lds)


def smooth(n, primes):
    for p in primes:
        while n % p == 0:
            n //= p
    return n == 1


def size_list(prec):
    lmax = 8192 if prec == "double" else 16384
    s = set()
    for n in range(2, 65):
        if smooth(n, (2, 3, 5, 7, 11, 13)):
            s.add(n)
    k = 1
    while k <= lmax:
        for m in (1, 3, 5):
            if 2 <= k * m <= lmax:
                s.add(k * m)
        k *= 2
    for base in (3, 5, 7):
        p = base
        while p <= lmax:
            s.add(p)
            p *= base
    s.update([100, 1000, 1001, 2187, 81, 729])
    if lmax >= 10000:
        s.add(10000)
    return sorted(x for x in s if x <= lmax)


def variants(prec, L, rads, mode, io, g0):
    """Alternative geometries for the autotuner on power-of-two lengths
    64..8192: half the threads per line (two butterflies each), twice or half
    the lines per workgroup, and the reversed radix order.  The default
    geometry stays first in the registry (what FFTW_ESTIMATE uses)."""
    if L & (L - 1) or L < 64 or L > 8192:
        return []
    out, seen = [], {(tuple(rads), g0["tpl"], g0["nl"])}
    cands = [(rads, 2, 1.0), (rads, 1, 2.0), (rads, 1, 0.5)]
    if io == "TOUT":
        # transposed stores want many lines per tile (long output runs)
        cands += [(rads, 1, 4.0), (rads, 2, 4.0), (rads, 2, 8.0)]
    if list(reversed(rads)) != list(rads):
        cands.append((list(reversed(rads)), 1, 1.0))
    if L >= 1024 and mode == "ROW" and io in ("PLAIN", "R2C", "C2R"):
        # radix-8 stages: one more LDS stage, twice the threads per line
        # (latency hiding when LDS allows only two workgroups per CU)
        k = L.bit_length() - 1
        r8 = [8] * (k // 3) + ([1 << (k % 3)] if k % 3 else [])
        if r8 != list(rads):
            cands.append((r8, 1, 1.0))
    for r, td, nm in cands:
        g = geometry(prec, L, r, mode, io, tpl_div=td, nl_mul=nm)
        if not g or g["tpl"] < 16:
            continue
        key = (tuple(r), g["tpl"], g["nl"])
        if key in seen:
            continue
        seen.add(key)
        out.append((prec, L, list(r), mode, io, g))
    return out


def catalog():
    entries = []
    for prec in ("double", "float"):
        lmax_plain = 8192 if prec == "double" else 16384
        for L in size_list(prec):
            rads = choose_radices(L)
            if rads is None:
                continue
            fam = smooth(L, (2,)) or (L % 3 == 0 and smooth(L // 3, (2,))) or \
                (L % 5 == 0 and smooth(L // 5, (2,)))
            for mode, io in (("ROW", "PLAIN"), ("COL", "PLAIN"), ("ROW", "TOUT")):
                if io == "TOUT" and (L < 16 or not fam):
                    continue
                r0 = rads
                if mode == "ROW" and io == "PLAIN" and L in (2048, 4096):
                    # one line per workgroup: radix-8 stages (twice the
                    # threads per line) measured faster than radix 16 for
                    # both precisions; radix 16 stays a tuner alternative
                    r0 = [8] * ((L.bit_length() - 1) // 3) + ([4] if L == 2048 else [])
                g = geometry(prec, L, r0, mode, io)
                if g:
                    pipe = mode == "ROW" and io == "PLAIN" and (prec, L) in PIPE_SIZES \
                        and L // r0[0] == g["tpl"]
                    if pipe and (prec, L) in PIPE_DEFAULT:
                        entries.append((prec, L, r0, mode, io, dict(g, pipe=1)))
                    entries.append((prec, L, r0, mode, io, g))
                    entries += variants(prec, L, r0, mode, io, g)
                    if pipe and (prec, L) not in PIPE_DEFAULT:
                        entries.append((prec, L, r0, mode, io, dict(g, pipe=1)))
                    if r0 is not rads:
                        g16 = geometry(prec, L, rads, mode, io)
                        if g16:
                            entries.append((prec, L, rads, mode, io, g16))
                    if mode == "ROW" and io == "PLAIN" and (prec, L) in XS_SIZES:
                        k = L.bit_length() - 1
                        r8 = [8] * (k // 3) + ([1 << (k % 3)] if k % 3 else [])
                        gx = geometry(prec, L, r8, mode, io)
                        if gx:
                            lp = L + L // gx["pad"] + 2
                            entries.append((prec, L, r8, mode, io,
                                            dict(gx, xs=1, lds=lp * gx["nl"] * CSZ[prec] // 2)))
            # half-length kernels for even real transforms (n = 2L)
            if (fam or L <= 64) and L <= 8192:
                for io in ("R2C", "C2R"):
                    g = geometry(prec, L, rads, "ROW", io)
                    if g:
                        entries.append((prec, L, rads, "ROW", io, g))
                        entries += variants(prec, L, rads, "ROW", io, g)
                        # the same kernel with a fused real-to-real kind
                        # (runtime rmode; kept out of the plain r2c / c2r
                        # code), row and column mappings: the column one
                        # serves r2r dims whose lines are strided
                        entries.append((prec, L, rads, "ROW", io + "_RR", g))
                        gc = geometry(prec, L, rads, "COL", io)
                        # not for FP64 c2r at one line per workgroup
                        # (L = 5120, 6144, 8192): an execute through the
                        # 8192-point one aborted on the device, cause not
                        # established; strided lines of those lengths take
                        # the row-mapped kernel
                        if gc and not (prec == "double" and io == "C2R" and gc["nl"] == 1):
                            entries.append((prec, L, rads, "COL", io + "_RR", gc))
            # half-length r2c / c2r fused with one radix-8 stage of the next
            # dimension's DFT across the 8 rows of a tile (2-D and higher
            # real transforms): the remaining column pass is 8x shorter, so
            # its tiles hold 8x more lines and read longer runs
            # (8 rows first; 4 rows, half the LDS, as the tuner's alternative)
            if (prec, L) in (("float", 512), ("float", 1024), ("double", 512)):
                g = geometry(prec, L, rads, "ROW", "R2C")
                for rows in XR_ROWS:
                    if not g:
                        break
                    lp = L + L // g["pad"] + 2
                    gx = dict(g, nl=rows, lds=lp * rows * CSZ[prec] + 8 * rows)
                    if gx["lds"] <= LDS_MAX and gx["tpl"] * rows <= 1024:
                        entries.append((prec, L, rads, "ROW", "R2C_XR", gx))
                        entries.append((prec, L, rads, "ROW", "C2R_XR", gx))
            # full-length real kernels for odd lengths (no half-length form)
            if L % 2 == 1 and L <= 8192:
                for io in ("R2C_FULL", "C2R_FULL"):
                    g = geometry(prec, L, rads, "ROW", "PLAIN")
                    if g:
                        entries.append((prec, L, rads, "ROW", io, g))
            if (fam or L <= 64) and L <= 8192:
                # DCT-IV / DST-IV of 2L reals through the L-point DFT
                for mode in ("ROW", "COL"):
                    g = geometry(prec, L, rads, mode, "DCT4")
                    if g:
                        entries.append((prec, L, rads, mode, "DCT4", g))
            # fused Bluestein on power-of-two convolution lengths; below 64
            # every length n with 2n - 1 <= L is 13-smooth and never takes
            # Bluestein
            if L & (L - 1) == 0 and L >= 64:
                g = geometry(prec, L, rads, "ROW", "BLUE")
                if g:
                    entries.append((prec, L, rads, "ROW", "BLUE", g))
    return entries


def cxx_entry(e):
    prec, L, rads, mode, io, g = e
    rl = "fk::RadList<%s>" % ",".join(map(str, rads))
    targs = "%s,%d,FK_%s,FK_IO_%s,%d,%s,%d" % (prec, g["nl"], mode, io, g["pad"], rl, g["tpl"])
    return targs, rl


PERSIST_NT = 512
PERSIST_LDS = 80 * 1024          # two persistent workgroups per CU
PERSIST_SIZES = {
    "float": [(n, "ROW", "PLAIN") for n in (64, 128, 256, 512, 1000, 1001, 1024, 2048, 2187, 4096)]
    + [(n, "COL", "PLAIN") for n in (64, 128, 256, 512, 1024, 2048)]
    + [(n, "ROW", "TOUT") for n in (64, 128, 256, 512, 1024, 2048, 4096)]
    + [(n, "ROW", "BLUE") for n in (2048, 4096, 8192)],
    "double": [(n, "ROW", "PLAIN") for n in (64, 128, 256, 512, 1000, 1001, 1024, 2048, 2187, 4096)]
    + [(n, "COL", "PLAIN") for n in (64, 128, 256, 512, 1024, 2048)]
    + [(n, "ROW", "TOUT") for n in (64, 128, 256, 512, 1024, 2048)]
    + [(n, "ROW", "BLUE") for n in (1024, 2048, 4096)],
}


def persist_menu(prec):
    """Geometries the persistent kernel carries: default radices and threads
    per line of the catalog, lines per workgroup grown to fill PERSIST_NT
    threads within PERSIST_LDS bytes."""
    out = []
    for L, mode, io in PERSIST_SIZES[prec]:
        rads = choose_radices(L)
        tpl = max(1, L // max(rads))
        if tpl > PERSIST_NT:
            continue
        csz = CSZ[prec]
        per = (L + L // rads[0] + 2) * csz if mode == "ROW" else (L + 2) * csz
        nl = max(1, PERSIST_NT // tpl)
        while nl > 1 and nl * per + 8 * nl + 64 * csz > PERSIST_LDS:
            nl //= 2
        if nl * per + 8 * (nl + 1) + 64 * csz > PERSIST_LDS:
            continue
        out.append((L, rads, tpl, nl, mode, io))
    return out


def write_persist():
    """fk_persist_<p>.hip (one per precision) + fk_persist.hip (menu lookup)."""
    for prec in ("float", "double"):
        menu = persist_menu(prec)
        p = PREC[prec]
        lines = ['// generated by fftw3_amd/gen_kernels.py -- do not edit',
                 '#include "../fk_stockham.h"', '',
                 'namespace {', 'constexpr int NT = %d;' % PERSIST_NT, '']
        # one separately register-allocated function per geometry
        for i, (L, rads, tpl, nl, mode, io) in enumerate(menu):
            lines.append('__device__ __noinline__ void body_%d(const fk_args *a, long long tile, unsigned char *smem)' % i)
            lines.append('{')
            lines.append('    fk::stockham_body<%s, %d, FK_%s, FK_IO_%s, %d, fk::RadList<%s>, %d, NT>(*a, tile, smem);'
                         % (prec, nl, mode, io, rads[0], ",".join(map(str, rads)), tpl))
            lines += ['}', '']
        lines.append('constexpr unsigned lds_max()')
        lines.append('{')
        lines.append('    unsigned m = 0;')
        for (L, rads, tpl, nl, mode, io) in menu:
            t = '%s, %d, FK_%s, FK_IO_%s, %d, fk::RadList<%s>, %d, NT' % (
                prec, nl, mode, io, rads[0], ",".join(map(str, rads)), tpl)
            lines.append('    m = m > fk::stockham_lds<%s>() ? m : fk::stockham_lds<%s>();' % (t, t))
        lines += ['    return m;', '}', '']
        lines.append('const fk_menu_entry menu[] = {')
        for (L, rads, tpl, nl, mode, io) in menu:
            lines.append('    {%d, %d, %d, {%s}, %d, %d, FK_%s, FK_IO_%s},'
                         % (p, L, len(rads), ",".join(map(str, rads)), tpl, nl, mode, io))
        lines += ['};', '']
        lines += [
            '/* Work-table driven persistent kernel: every workgroup takes the next',
            '   tile from the global counter until the table is drained (the exit',
            '   every wave reaches), finds its pass by binary search over the tile',
            '   prefix sums, and runs that geometry on the tile. */',
            '__global__ void __launch_bounds__(NT) persist(const fk_bpass *tab, int npass, long long total,',
            '                                              unsigned long long *counter)',
            '{',
            '    extern __shared__ unsigned char smem[];',
            '    __shared__ long long s_item;',
            '    for (;;) {',
            '        if (threadIdx.x == 0) s_item = (long long)atomicAdd(counter, 1ULL);',
            '        __syncthreads();',
            '        const long long t = s_item;',
            '        __syncthreads();',
            '        if (t >= total) break;',
            '        int lo = 0, hi = npass - 1;',
            '        while (lo < hi) {',
            '            const int mid = (lo + hi + 1) >> 1;',
            '            if (tab[mid].first <= t) lo = mid; else hi = mid - 1;',
            '        }',
            '        const fk_bpass *bp = tab + lo;',
            '        const long long tile = t - bp->first;',
            '        switch (bp->kind) {']
        for i in range(len(menu)):
            lines.append('        case %d: body_%d(&bp->a, tile, smem); break;' % (i, i))
        lines += [
            '        default: break;',
            '        }',
            '        __syncthreads();',
            '    }',
            '}',
            '} // namespace',
            '',
            'extern "C" const fk_menu_entry *fk_persist_menu_%d(int *n, const void **fn, unsigned *lds)' % p,
            '{',
            '    *n = (int)(sizeof(menu) / sizeof(menu[0]));',
            '    *fn = (const void *)&persist;',
            '    *lds = lds_max();',
            '    return menu;',
            '}',
            '']
        path = os.path.join(OUT, "fk_persist_%d.hip" % p)
        txt = "\n".join(lines)
        if not os.path.exists(path) or open(path).read() != txt:
            open(path, "w").write(txt)
    lines = ['// generated by fftw3_amd/gen_kernels.py -- do not edit',
             '#include <hip/hip_runtime.h>',
             '#include "../fk_args.h"', '',
             'extern "C" const fk_menu_entry *fk_persist_menu_0(int *n, const void **fn, unsigned *lds);',
             'extern "C" const fk_menu_entry *fk_persist_menu_1(int *n, const void **fn, unsigned *lds);', '',
             'namespace {',
             '__global__ void reset_counters(unsigned long long *c, int n)',
             '{',
             '    if ((int)threadIdx.x < n) c[threadIdx.x] = 0ULL;',
             '}',
             '} // namespace', '',
             'extern "C" const fk_menu_entry *fk_persist_menu(int prec, int *n, const void **fn, unsigned *lds, int *nt)',
             '{',
             '    *nt = %d;' % PERSIST_NT,
             '    return prec == 0 ? fk_persist_menu_0(n, fn, lds) : fk_persist_menu_1(n, fn, lds);',
             '}', '',
             'extern "C" int fkm_reset_counters(unsigned long long *c, int n, hipStream_t st)',
             '{',
             '    hipLaunchKernelGGL(reset_counters, dim3(1), dim3(64), 0, st, c, n);',
             '    return (int)hipGetLastError();',
             '}', '']
    path = os.path.join(OUT, "fk_persist.hip")
    txt = "\n".join(lines)
    if not os.path.exists(path) or open(path).read() != txt:
        open(path, "w").write(txt)


def main():
    write_persist()
    os.makedirs(OUT, exist_ok=True)
    entries = catalog()
    # variant index within (precision, L, mapping, io): 0 is the default
    vcount = {}
    numbered = []
    for e in entries:
        key = (e[0], e[1], e[3], e[4])
        v = vcount.get(key, 0)
        vcount[key] = v + 1
        numbered.append(e + (v,))
    entries = numbered
    nparts = 8
    parts = [[] for _ in range(nparts)]
    # round-robin by estimated cost so parts compile in similar time
    for i, e in enumerate(sorted(entries, key=lambda e: -e[1])):
        parts[i % nparts].append(e)
    written = set()
    for p, es in enumerate(parts):
        lines = ['// generated by fftw3_amd/gen_kernels.py -- do not edit',
                 '#include "../fk_stockham.h"', '',
                 'extern "C" const fk_kernel *fk_registry_part%d(int *n)' % p, '{',
                 '    static const fk_kernel tab[] = {']
        for e in es:
            prec, L, rads, mode, io, g, var = e
            targs, rl = cxx_entry(e[:6])
            radarr = ",".join(map(str, rads))
            if g.get("xs"):
                xargs = "%s,%d,FK_%s,%d,%s,%d" % (prec, g["nl"], mode, g["pad"], rl, g["tpl"])
                lines.append(
                    '        {%d, %d, %d, {%s}, %d, %d, FK_%s, FK_IO_%s, fk::stockham_xs_lds<%s>(), '
                    '%s::twsize(), (const void *)&fk::stockham_xs<%s>, %d},'
                    % (PREC[prec], L, len(rads), radarr, g["tpl"], g["nl"], mode, io, xargs, rl, xargs, var))
                continue
            if g.get("pipe"):
                pargs = "%s,%d,%d,%s,%d" % (prec, g["nl"], g["pad"], rl, g["tpl"])
                lines.append(
                    '        {%d, %d, %d, {%s}, %d, %d, FK_%s, FK_IO_%s, fk::stockham_lds<%s>(), '
                    '%s::twsize(), (const void *)&fk::stockham_pipe<%s>, %d, 1},'
                    % (PREC[prec], L, len(rads), radarr, g["tpl"], g["nl"], mode, io, targs, rl, pargs, var))
                continue
            lines.append(
                '        {%d, %d, %d, {%s}, %d, %d, FK_%s, FK_IO_%s, fk::stockham_lds<%s>(), '
                '%s::twsize(), (const void *)&fk::stockham<%s>, %d},'
                % (PREC[prec], L, len(rads), radarr, g["tpl"], g["nl"], mode, io, targs, rl, targs, var))
        lines += ['    };', '    *n = (int)(sizeof(tab) / sizeof(tab[0]));', '    return tab;', '}', '']
        path = os.path.join(OUT, "fk_inst_%d.hip" % p)
        written.add(path)
        txt = "\n".join(lines)
        if not os.path.exists(path) or open(path).read() != txt:
            open(path, "w").write(txt)
    reg = ['// generated by fftw3_amd/gen_kernels.py -- do not edit',
           '#include <stdlib.h>', '#include <string.h>', '#include "../fk_args.h"', '']
    for p in range(nparts):
        reg.append('extern "C" const fk_kernel *fk_registry_part%d(int *n);' % p)
    reg += ['', 'extern "C" const fk_kernel *fk_registry(int *count)', '{',
            '    static fk_kernel *all = 0;', '    static int total = 0;',
            '    if (!all) {', '        int n, off = 0;', '        const fk_kernel *t;',
            '        fk_kernel *buf = (fk_kernel *)malloc(sizeof(fk_kernel) * %d);' % len(entries)]
    for p in range(nparts):
        reg.append('        t = fk_registry_part%d(&n); memcpy(buf + off, t, sizeof(fk_kernel) * n); off += n;' % p)
    reg += ['        total = off;', '        all = buf;', '    }', '    *count = total;', '    return all;', '}', '']
    path = os.path.join(OUT, "fk_registry.hip")
    txt = "\n".join(reg)
    if not os.path.exists(path) or open(path).read() != txt:
        open(path, "w").write(txt)
    print("kernels:", len(entries), file=sys.stderr)


if __name__ == "__main__":
    main()
