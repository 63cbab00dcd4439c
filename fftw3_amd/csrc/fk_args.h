/*
 * Kernel argument block and kernel registry shared by the C host library
 * and the HIP kernels.  Plain C types only: the same struct is filled by
 * the host and passed by value to every transform kernel.
 */
#ifndef FK_ARGS_H
#define FK_ARGS_H

#ifdef __cplusplus
extern "C" {
#endif

/* compute mapping of threads onto (line, butterfly) */
enum { FK_ROW = 0,   /* threads of a line are consecutive: unit-stride lines */
       FK_COL = 1 }; /* lines of a tile are consecutive: unit-stride batch */

/* what the kernel does around the Stockham stages */
enum {
    FK_IO_PLAIN = 0,  /* load line, FFT, store line */
    FK_IO_TOUT = 1,   /* store through LDS with lines-fastest mapping */
    FK_IO_R2C = 2,    /* real input viewed as L complex; store n/2+1 bins */
    FK_IO_C2R = 3,    /* load n/2+1 bins, half-length inverse, real output */
    FK_IO_BLUE = 4,   /* fused Bluestein convolution of length L */
    FK_IO_R2C_RR = 5, /* FK_IO_R2C with a fused real-to-real kind (rmode) */
    FK_IO_C2R_RR = 6, /* FK_IO_C2R with a fused real-to-real kind (rmode) */
    FK_IO_DCT4 = 7,   /* REDFT11 / RODFT11 of n = 2L reals through an L-point DFT */
    FK_IO_R2C_FULL = 8, /* real line of n = L samples, L-point DFT, bins k <= L/2 stored */
    FK_IO_C2R_FULL = 9, /* bins k <= L/2, Hermitian L-point inverse, real part stored */
    FK_IO_R2C_XR = 10,  /* FK_IO_R2C on the NL rows of a tile, then the first (radix-NL,
                           decimation in frequency) stage of the column DFT across them */
    FK_IO_C2R_XR = 11   /* the last (radix-NL, decimation in time) stage of the inverse
                           column DFT across the NL rows of a tile, then FK_IO_C2R */
};

/* flags */
#define FK_SWAP_IN   1u   /* swap re/im on load  (backward transform) */
#define FK_SWAP_OUT  2u   /* swap re/im on store (backward transform) */
#define FK_TW_OUT    4u   /* multiply output bin k of line with batch index
                             i0 by W_N^(k*i0) (four-step twiddle) */
#define FK_SCALE     8u   /* multiply output by scale */
#define FK_REAL_UNIT 16u  /* R2C: real input stride is 1 / C2R: output stride 1 */
#define FK_REAL_QUAD 32u  /* FP32 real input stride 1, lines 16-byte aligned: the
                             reordered loads of REDFT10 / RODFT10 read quads */

/* real-to-real kinds fused into the FK_IO_R2C_RR / FK_IO_C2R_RR kernels (rmode):
   the real line is read or written in the kind's order, with the kind's
   twiddle exp(-i pi m / 4n) from rtw */
enum { FK_RM_NONE = 0,
       FK_RM_HC = 1,      /* R2C: halfcomplex output (R2HC) */
       FK_RM_DHT = 2,     /* R2C: Hartley output */
       FK_RM_DCT2 = 3,    /* R2C: REDFT10 */
       FK_RM_DST2 = 4,    /* R2C: RODFT10 */
       FK_RM_HC2R = 5,    /* C2R: halfcomplex input (HC2R) */
       FK_RM_DCT3 = 6,    /* C2R: REDFT01 */
       FK_RM_DST3 = 7,    /* C2R: RODFT01 */
       FK_RM_E00 = 8,     /* R2C of the even extension (length 2L = 2(n-1)): REDFT00 */
       FK_RM_O00 = 9,     /* R2C of the odd extension (length 2L = 2(n+1)): RODFT00 */
       FK_RM_DCT4 = 10,   /* FK_IO_DCT4: REDFT11 */
       FK_RM_DST4 = 11 }; /* FK_IO_DCT4: RODFT11 */

typedef struct {
    const void *in;
    void *out;
    long long is, os;         /* element strides within a line */
    long long ids[3], ods[3]; /* batch strides, dim 0 fastest */
    long long nb[3];          /* batch counts */
    long long nlines;         /* nb[0]*nb[1]*nb[2] */
    const void *tw;           /* per-stage twiddles, working precision */
    /* four-step output twiddle W_N^e, e = k*i0 < N, from an FP64
       two-level table: W_N^e = tlo[e % tb] * thi[e / tb] */
    const double *tlo, *thi;
    long long tN, tb;
    /* FK_IO_R2C / FK_IO_C2R: W_n^k, k <= L (working precision)
       FK_IO_BLUE: chirp c_m (m < nvalid), then FFT_L(conj chirp)/L */
    const void *aux, *aux2;
    long long nvalid;         /* BLUE: logical length n; else unused */
    double scale;
    unsigned flags;
    unsigned xcd_remap;       /* nonzero: spread tiles XCD-contiguously */
    unsigned rmode;           /* FK_RM_*: fused real-to-real kind */
    unsigned pad_;
    const void *rtw;          /* FK_RM_DCT2/3, FK_RM_DST2/3: exp(-i pi m / 4n), m < 2n */
} fk_args;

typedef struct {
    int prec;          /* 0 = double, 1 = float */
    int L;             /* FFT length computed by the kernel */
    int nrad;
    int rad[12];
    int tpl;           /* threads per line */
    int nl;            /* lines per workgroup */
    int mode;          /* FK_ROW / FK_COL */
    int io;            /* FK_IO_* */
    unsigned lds;      /* dynamic LDS bytes per workgroup */
    int twsz;          /* complex entries in the per-stage twiddle table */
    const void *fn;    /* kernel host stub for hipLaunchKernel */
    int variant;       /* 0 = default geometry; >0 autotuner alternatives */
    int loop;          /* 1: each workgroup runs several tiles (grid sized to
                          the workgroups the device holds at once) */
} fk_kernel;

/* elementwise passes around the Stockham kernels (multi-pass paths) */
enum {
    FK_EW_COPY = 0,     /* complex copy, optional swaps */
    FK_EW_COPY_R,       /* real copy */
    FK_EW_BS_PREP,      /* Bluestein: a_j = x_j c_j (j < p0), 0 up to p1 */
    FK_EW_BS_MUL,       /* Bluestein: swap(A_j * B_j) */
    FK_EW_BS_POST,      /* Bluestein: y_k = swap(v_k) c_k, k < p0 */
    FK_EW_R2C_POST,     /* Z (L = p0 complex) -> X[k], k <= L */
    FK_EW_C2R_PRE,      /* X[k], k <= L -> Z (L = p0 complex) */
    FK_EW_C2R_OUT,      /* z (p0 complex) -> interleaved real output */
    FK_EW_WIDEN,        /* real -> complex (re, 0) */
    FK_EW_HERM,         /* n/2+1 bins -> full Hermitian line of n = p0 */
    FK_EW_REAL,         /* complex -> real part */
    FK_EW_SPLIT_GATHER, /* (in[j], in2[j]) reals -> interleaved complex */
    FK_EW_SPLIT_SCATTER,/* interleaved complex -> out[j] (re), out2[j] (im) */
    /* r2r pre / post passes around an r2c, c2r or c2c of a staging line.
       p0 = n (the r2r length), p1 = 1 for the sine kinds; aux = the table
       exp(-i pi m / 4n), m < 2n */
    FK_EW_HC_POST,      /* n/2+1 bins -> halfcomplex (p1 = 0) or DHT (p1 = 1) */
    FK_EW_HC_PRE,       /* halfcomplex -> n/2+1 Hermitian bins */
    FK_EW_DCT2_PRE,     /* x -> v (even samples up, odd samples down) */
    FK_EW_DCT2_POST,    /* V (n/2+1 bins) -> 2 Re(w^k V_k) */
    FK_EW_DCT3_PRE,     /* x -> conj(w^k) (x_k - i x_{n-k}), k <= n/2 */
    FK_EW_DCT3_POST,    /* v (n reals) -> interleave v_j, v_{n-1-j} */
    FK_EW_DCT4_PRE,     /* even n: (x_2m + i x_{n-1-2m}) w^{4m+1} */
    FK_EW_DCT4_POST,    /* even n: Y_2p = 2 Re(c_p), Y_{n-1-2p} = -2 Im(c_p) */
    FK_EW_DCT4O_PRE,    /* odd n: x_j w^{2j}, zero up to 2n */
    FK_EW_DCT4O_POST,   /* odd n: Y_k = 2 Re(w^{2k+1} A_k) */
    FK_EW_SYM_PRE,      /* even (p1 = 0, length 2(n-1)) / odd (p1 = 1, 2(n+1)) extension */
    FK_EW_SYM_POST      /* Re(B_k) / -Im(B_{k+1}) */
};

typedef struct {
    const void *in;
    void *out;
    long long n;              /* elements per line */
    long long is, os;
    long long ids[3], ods[3];
    long long nb[3];
    long long nlines;
    const void *aux, *aux2;
    long long p0, p1;
    double scale;
    unsigned flags;           /* FK_SWAP_IN / FK_SWAP_OUT / FK_SCALE */
    int op;
    const void *in2;          /* FK_EW_SPLIT_GATHER: imaginary input */
    void *out2;               /* FK_EW_SPLIT_SCATTER: imaginary output */
} fk_ew_args;

/* registry, defined by the generated instantiation files */
extern const fk_kernel *fk_registry(int *count);

/* One pass of a batched set executed by the persistent kernel: tiles
   [first, first + tiles) of the global work counter run geometry `kind` of
   the persistent menu on the lines described by `a`. */
typedef struct {
    fk_args a;
    int kind;
    int pad_;
    long long tiles;
    long long first;
} fk_bpass;

typedef struct {
    int prec, L, nrad;
    int rad[12];
    int tpl, nl, mode, io;
} fk_menu_entry;

/* persistent-kernel menu of a precision: entries, kernel stub, dynamic LDS
   bytes and workgroup size (generated) */
extern const fk_menu_entry *fk_persist_menu(int prec, int *n, const void **fn, unsigned *lds,
                                            int *nt);

#ifdef __cplusplus
}
#endif
#endif
