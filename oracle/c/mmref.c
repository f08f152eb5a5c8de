/*
 * mmref.c -- FP64 C/OpenMP restatement of the 2-D spiral CG-SENSE path,
 * used as the CPU baseline (TEST INFRASTRUCTURE ONLY: the product path is
 * the HIP library; nothing there links this).
 *
 * It follows the NumPy oracle (oracle/nufft.py, oracle/rcvr.py,
 * oracle/mrsamp.py, oracle/solvers.py), which restates:
 *   nufftOp plan (doPrep)      operators/@nufftOp/nufftOp.m:138-274
 *   nufftOp forward / adjoint  operators/@nufftOp/nufftOp.m:378-394, :354-377
 *   fftnc / ifftnc             general/fftnc.m, general/ifftnc.m
 *   rcvrOp maps path           operators/@rcvrOp/rcvrOp.m:145-157
 *   mrSampFunc (E = S o R)     operators/mrSampFunc.m:82-159
 *   cgne                       iterativeSolvers/cgne.m:136-335
 *   sampHighOrder direct       operators/@sampHighOrder/sampHighOrder.m:361-391,
 *                              phase of prepForDirect :496-588 (oracle/highorder.py)
 * and is pinned against it by tests/test_oracle_c.py (1e-12 per
 * application, identical CG histories to 1e-10).
 *
 * Arrays are column-major and complex values interleaved (re, im), like
 * MATLAB and like NumPy complex128 arrays in Fortran order.
 */
#define _GNU_SOURCE
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ---- plan constants (nufftOp.m:146-186) --------------------------------- */
static int max_prime(int n) {
    int best = 1;
    for (int p = 2; p * p <= n; ++p)
        while (n % p == 0) { best = p; n /= p; }
    return n > 1 ? (n > best ? n : best) : best;
}

static int grid_size(int n, double os) {
    int g = (int)ceil(n * os / 2.0) * 2;
    while (max_prime(g) > 7) g += 2;
    return g;
}

static double bessel_i0(double x) {
    const double q = 0.25 * x * x;
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 500; ++k) {
        term *= q / ((double)k * k);
        sum += term;
        if (term < 1e-17 * sum) break;
    }
    return sum;
}

/* deapodisation of one dimension, normalised by its smallest magnitude and
 * scaled by sqrt(G/N) */
static void deapod(int N, int G, int W, double beta, double* c) {
    double cmin = 1e300;
    int imin = 0;
    for (int i = 0; i < N; ++i) {
        const double j = -N / 2.0 + i;
        const double x = j * M_PI * W / G;
        double v;
        if (beta * beta > x * x) {
            const double a = sqrt(beta * beta - x * x);
            v = a < 1e-12 ? 1.0 : a / sinh(a);
        } else {
            const double a = sqrt(x * x - beta * beta);
            v = a < 1e-12 ? 1.0 : a / sin(a);
        }
        c[i] = v;
        if (fabs(v) < cmin) { cmin = fabs(v); imin = i; }
    }
    const double nrm = c[imin];
    for (int i = 0; i < N; ++i) c[i] = c[i] / nrm * sqrt((double)G / N);
}

/* ---- FFT: Stockham, radices 4, 2, 3, 5, 7 ------------------------------ */
typedef struct {
    int n, nf, r[32];
    cplx* tw;    /* exp(+2 pi i t / n), t < n */
} fftplan;

static int fft_init(fftplan* f, int n) {
    f->n = n;
    f->nf = 0;
    int m = n;
    while (m % 4 == 0) { f->r[f->nf++] = 4; m /= 4; }
    const int pr[4] = {2, 3, 5, 7};
    for (int i = 0; i < 4; ++i)
        while (m % pr[i] == 0) { f->r[f->nf++] = pr[i]; m /= pr[i]; }
    if (m != 1) return -1;
    f->tw = (cplx*)malloc(sizeof(cplx) * n);
    for (int t = 0; t < n; ++t) f->tw[t] = cexp(2.0 * M_PI * I * t / n);
    return 0;
}

/* unnormalised transform of a[0..n-1]: X_q = sum_j a_j exp(sgn 2 pi i j q / n) */
static void fft_run(const fftplan* f, cplx* a, cplx* tmp, int sgn) {
    const int n = f->n;
    cplx *src = a, *dst = tmp;
    int p = 1;
    for (int s = 0; s < f->nf; ++s) {
        const int R = f->r[s];
        const int nb = n / R;
        const int step = n / (p * R);
        cplx wr[7];
        for (int j = 0; j < R; ++j) {
            const cplx w = f->tw[(j * (n / R)) % n];
            wr[j] = sgn > 0 ? w : conj(w);
        }
        for (int i = 0; i < nb; ++i) {
            const int k = i % p;
            cplx v[7], o[7];
            for (int j = 0; j < R; ++j) {
                v[j] = src[i + j * nb];
                if (j > 0) {
                    const cplx t = f->tw[(j * k * step) % n];
                    v[j] *= sgn > 0 ? t : conj(t);
                }
            }
            if (R == 2) {
                o[0] = v[0] + v[1];
                o[1] = v[0] - v[1];
            } else if (R == 4) {
                const cplx a0 = v[0] + v[2], a1 = v[0] - v[2];
                const cplx b0 = v[1] + v[3], b1 = (v[1] - v[3]) * (sgn > 0 ? I : -I);
                o[0] = a0 + b0;
                o[2] = a0 - b0;
                o[1] = a1 + b1;
                o[3] = a1 - b1;
            } else {
                for (int q = 0; q < R; ++q) {
                    cplx acc = 0;
                    for (int j = 0; j < R; ++j) acc += v[j] * wr[(j * q) % R];
                    o[q] = acc;
                }
            }
            const int ob = (i - k) * R + k;
            for (int j = 0; j < R; ++j) dst[ob + j * p] = o[j];
        }
        cplx* t = src;
        src = dst;
        dst = t;
        p *= R;
    }
    if (src != a) memcpy(a, src, sizeof(cplx) * n);
}

/* ---- SENSE operator ------------------------------------------------------ */
typedef struct {
    int N1, N2, G1, G2, M, W, C, ntap;
    double* c1;       /* deapodisation, dim 1 [N1] and dim 2 [N2] */
    double* c2;
    double* wt;       /* [M][ntap] interpolation weights */
    int* lin;         /* [M][ntap] grid index (column-major) */
    int* gptr;        /* transposed CSR: per grid cell, its (sample, tap) entries */
    int* gent;
    double* dcf;      /* [M] or NULL */
    cplx* shift;      /* fftShiftArray [G1*G2] */
    cplx* maps;       /* [V][C] */
    fftplan f1, f2;
    int nthr;
} mmref_sense;

int mmref_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void mmref_sense_destroy(mmref_sense* s) {
    if (!s) return;
    free(s->c1); free(s->c2); free(s->wt); free(s->lin); free(s->gptr); free(s->gent);
    free(s->dcf); free(s->shift); free(s->maps); free(s->f1.tw); free(s->f2.tw);
    free(s);
}

mmref_sense* mmref_sense_create(int N1, int N2, int M, const double* kloc, const double* dcf, double os,
                                int W, int C, const double* maps) {
    mmref_sense* s = (mmref_sense*)calloc(1, sizeof(mmref_sense));
    if (!s) return NULL;
    s->N1 = N1; s->N2 = N2; s->M = M; s->W = W; s->C = C;
    s->G1 = grid_size(N1, os);
    s->G2 = grid_size(N2, os);
    s->ntap = W * W;
    const double beta = M_PI * sqrt((double)W * W / (os * os) * (os - 0.5) * (os - 0.5) - 0.8);
    const int nk = (int)ceil(0.5 * (W + 1) * sqrt(0.37 / (os * os) / 1e-6)) * 2;
    /* LUT with the zero sentinel, normalised (nufftOp.m:157-165) */
    double* kern = (double*)malloc(sizeof(double) * (nk + 1));
    double ksum = 0;
    for (int i = 0; i < nk; ++i) {
        const double kx = -1.0 + 2.0 * i / (nk - 1);
        kern[i] = bessel_i0(beta * sqrt(fmax(0.0, 1.0 - kx * kx)));
        ksum += kern[i];
    }
    kern[nk] = 0.0;
    for (int i = 0; i <= nk; ++i) kern[i] = kern[i] / ksum * (nk - 1) / W;
    s->c1 = (double*)malloc(sizeof(double) * N1);
    s->c2 = (double*)malloc(sizeof(double) * N2);
    deapod(N1, s->G1, W, beta, s->c1);
    deapod(N2, s->G2, W, beta, s->c2);
    /* per-sample taps, dim 1 fastest (nufftOp.m:187-246) */
    const int nt = s->ntap;
    s->wt = (double*)malloc(sizeof(double) * (size_t)M * nt);
    s->lin = (int*)malloc(sizeof(int) * (size_t)M * nt);
    for (int m = 0; m < M; ++m) {
        double w[2][8];
        int g[2][8];
        for (int d = 0; d < 2; ++d) {
            const int G = d == 0 ? s->G1 : s->G2;
            const double ka = (kloc[(size_t)d * M + m] + 0.5) * G;
            const double fl = floor(ka) - ka;
            for (int q = 0; q < W; ++q) {
                const int off = -W / 2 + 1 + q;
                double ki = (fl + off) * (nk - 1) / W + nk / 2.0 + 0.5;
                const double fk = floor(ki);
                const double frac = ki - fk;
                const int ii = (int)fk;   /* 1-based */
                w[d][q] = (1 - frac) * kern[ii - 1] + frac * kern[ii];
                long ci = (long)floor(ka) + off + G / 2;   /* 0-based before the wrap */
                ci %= G;
                if (ci < 0) ci += G;
                g[d][q] = (int)ci;
            }
        }
        for (int q2 = 0; q2 < W; ++q2)
            for (int q1 = 0; q1 < W; ++q1) {
                const int t = q2 * W + q1;
                s->wt[(size_t)m * nt + t] = w[0][q1] * w[1][q2];
                s->lin[(size_t)m * nt + t] = g[0][q1] + s->G1 * g[1][q2];
            }
    }
    free(kern);
    /* transposed CSR for the adjoint (gather per grid cell, entries in
     * (sample, tap) order) */
    const int G = s->G1 * s->G2;
    s->gptr = (int*)calloc((size_t)G + 1, sizeof(int));
    s->gent = (int*)malloc(sizeof(int) * (size_t)M * nt);
    for (size_t e = 0; e < (size_t)M * nt; ++e) s->gptr[s->lin[e] + 1]++;
    for (int i = 0; i < G; ++i) s->gptr[i + 1] += s->gptr[i];
    int* pos = (int*)malloc(sizeof(int) * G);
    memcpy(pos, s->gptr, sizeof(int) * G);
    for (size_t e = 0; e < (size_t)M * nt; ++e) s->gent[pos[s->lin[e]]++] = (int)e;
    free(pos);
    if (dcf) {
        s->dcf = (double*)malloc(sizeof(double) * M);
        memcpy(s->dcf, dcf, sizeof(double) * M);
    }
    /* fftShiftArray (nufftOp.m:258-273) */
    s->shift = (cplx*)malloc(sizeof(cplx) * G);
    for (int b = 0; b < s->G2; ++b)
        for (int a = 0; a < s->G1; ++a) {
            const double v = -(a - s->G1 / 2.0) * N1 / s->G1 - (b - s->G2 / 2.0) * N2 / s->G2;
            s->shift[a + s->G1 * b] = cexp(I * M_PI * v);
        }
    const size_t V = (size_t)N1 * N2;
    s->maps = (cplx*)malloc(sizeof(cplx) * V * C);
    memcpy(s->maps, maps, sizeof(cplx) * V * C);
    if (fft_init(&s->f1, s->G1) || fft_init(&s->f2, s->G2)) {
        mmref_sense_destroy(s);
        return NULL;
    }
    s->nthr = mmref_num_threads();
    return s;
}

/* unitary inverse (sgn = +1) or forward (sgn = -1) 2-D FFT of a G1 x G2
 * grid; rows with index >= nrow1 are zero (forward: zero-padded image) */
static void fft2(const mmref_sense* s, cplx* z, int sgn, int nrow_in) {
    const int G1 = s->G1, G2 = s->G2;
    const double scl = 1.0 / sqrt((double)G1 * G2);
    if (sgn > 0) {
        /* dim 1 over the nonzero columns i2 < nrow_in, then dim 2 */
#pragma omp parallel
        {
            cplx* tmp = (cplx*)malloc(sizeof(cplx) * (G1 > G2 ? G1 : G2) * 2);
            cplx* buf = tmp + (G1 > G2 ? G1 : G2);
#pragma omp for schedule(static)
            for (int b = 0; b < nrow_in; ++b) fft_run(&s->f1, z + (size_t)b * G1, tmp, sgn);
#pragma omp for schedule(static)
            for (int a = 0; a < G1; ++a) {
                for (int b = 0; b < G2; ++b) buf[b] = z[a + (size_t)G1 * b];
                fft_run(&s->f2, buf, tmp, sgn);
                for (int b = 0; b < G2; ++b) z[a + (size_t)G1 * b] = buf[b] * scl;
            }
            free(tmp);
        }
    } else {
        /* dim 2 over every column, then dim 1 over the rows kept by the crop */
#pragma omp parallel
        {
            cplx* tmp = (cplx*)malloc(sizeof(cplx) * (G1 > G2 ? G1 : G2) * 2);
            cplx* buf = tmp + (G1 > G2 ? G1 : G2);
#pragma omp for schedule(static)
            for (int a = 0; a < G1; ++a) {
                for (int b = 0; b < G2; ++b) buf[b] = z[a + (size_t)G1 * b];
                fft_run(&s->f2, buf, tmp, sgn);
                for (int b = 0; b < G2; ++b) z[a + (size_t)G1 * b] = buf[b];
            }
#pragma omp for schedule(static)
            for (int b = 0; b < nrow_in; ++b) {
                fft_run(&s->f1, z + (size_t)b * G1, tmp, sgn);
                for (int a = 0; a < G1; ++a) z[a + (size_t)G1 * b] *= scl;
            }
            free(tmp);
        }
    }
}

/* y [M x C] = E x (SENSE forward: coils, deapodise, pad, IFFT, ramp, degrid, dcf) */
void mmref_forward(mmref_sense* s, const double* xin, double* yout) {
    const cplx* x = (const cplx*)xin;
    cplx* y = (cplx*)yout;
    const int N1 = s->N1, N2 = s->N2, G1 = s->G1, M = s->M, nt = s->ntap;
    const size_t V = (size_t)N1 * N2, G = (size_t)G1 * s->G2;
    cplx* z = (cplx*)malloc(sizeof(cplx) * G);
    for (int c = 0; c < s->C; ++c) {
        memset(z, 0, sizeof(cplx) * G);
        const cplx* mp = s->maps + V * c;
#pragma omp parallel for schedule(static)
        for (int b = 0; b < N2; ++b)
            for (int a = 0; a < N1; ++a) {
                const size_t v = a + (size_t)N1 * b;
                z[a + (size_t)G1 * b] = x[v] * mp[v] * s->c1[a] * s->c2[b];
            }
        fft2(s, z, +1, N2);
#pragma omp parallel for schedule(static)
        for (size_t g = 0; g < G; ++g) z[g] *= s->shift[g];
        cplx* yc = y + (size_t)M * c;
#pragma omp parallel for schedule(static)
        for (int m = 0; m < M; ++m) {
            cplx acc = 0;
            for (int t = 0; t < nt; ++t) acc += s->wt[(size_t)m * nt + t] * z[s->lin[(size_t)m * nt + t]];
            yc[m] = s->dcf ? acc * s->dcf[m] : acc;
        }
    }
    free(z);
}

/* x = E' y (dcf, gridding, conj ramp, FFT, crop, deapodise, conj-coil sum) */
void mmref_adjoint(mmref_sense* s, const double* yin, double* xout) {
    const cplx* y = (const cplx*)yin;
    cplx* x = (cplx*)xout;
    const int N1 = s->N1, N2 = s->N2, G1 = s->G1, M = s->M, nt = s->ntap;
    const size_t V = (size_t)N1 * N2, G = (size_t)G1 * s->G2;
    cplx* z = (cplx*)malloc(sizeof(cplx) * G);
    cplx* yd = (cplx*)malloc(sizeof(cplx) * M);
    memset(x, 0, sizeof(cplx) * V);
    for (int c = 0; c < s->C; ++c) {
        const cplx* yc = y + (size_t)M * c;
#pragma omp parallel for schedule(static)
        for (int m = 0; m < M; ++m) yd[m] = s->dcf ? yc[m] * s->dcf[m] : yc[m];
#pragma omp parallel for schedule(dynamic, 64)
        for (size_t g = 0; g < G; ++g) {
            cplx acc = 0;
            for (int e = s->gptr[g]; e < s->gptr[g + 1]; ++e) {
                const int en = s->gent[e];
                acc += s->wt[en] * yd[en / nt];
            }
            z[g] = acc * conj(s->shift[g]);
        }
        fft2(s, z, -1, N2);
        const cplx* mp = s->maps + V * c;
#pragma omp parallel for schedule(static)
        for (int b = 0; b < N2; ++b)
            for (int a = 0; a < N1; ++a) {
                const size_t v = a + (size_t)N1 * b;
                x[v] += conj(mp[v]) * (z[a + (size_t)G1 * b] * s->c1[a] * s->c2[b]);
            }
    }
    free(z);
    free(yd);
}

static double dot_re(const cplx* a, const cplx* b, size_t n) {
    double sr = 0;
#pragma omp parallel for reduction(+ : sr) schedule(static)
    for (size_t i = 0; i < n; ++i) sr += creal(conj(a[i]) * b[i]);
    return sr;
}

static cplx dot_c(const cplx* a, const cplx* b, size_t n) {
    double sr = 0, si = 0;
#pragma omp parallel for reduction(+ : sr, si) schedule(static)
    for (size_t i = 0; i < n; ++i) {
        const cplx t = conj(a[i]) * b[i];
        sr += creal(t);
        si += cimag(t);
    }
    return sr + I * si;
}

static int g_verbose = 0;

/* v != 0: cgne prints one line per iteration on stderr */
void mmref_set_verbose(int v) { g_verbose = v; }

/* an encoding operator E: image [V] -> data [D] and back */
typedef struct {
    void (*fwd)(void* s, const double* x, double* y);
    void (*adj)(void* s, const double* y, double* x);
    void* s;
    size_t V, D;
} refop;

static void sense_fwd(void* s, const double* x, double* y) { mmref_forward((mmref_sense*)s, x, y); }
static void sense_adj(void* s, const double* y, double* x) { mmref_adjoint((mmref_sense*)s, y, x); }

/* cgne on E'E x = E'y (cgne.m:136-335): x0 = NULL means x0 = b; thresholds
 * < 0 are off; resThresh_rel (if >= 0) adds |r|^2 < resThresh_rel*|b|^2.
 * resSqAll (NitMax+1 doubles) may be NULL.  Returns the iteration count. */
static int cgne_core(const refop* op, const double* yin, const double* x0in, double* xout, int NitMax,
                     double resThresh, double resThresh_rel, double xChangeThresh, int expResN, double* resSqAll) {
    void* s = op->s;
    const size_t V = op->V;
    const size_t D = op->D;
    cplx* b = (cplx*)malloc(sizeof(cplx) * V);
    cplx* r = (cplx*)malloc(sizeof(cplx) * V);
    cplx* p = (cplx*)malloc(sizeof(cplx) * V);
    cplx* q = (cplx*)malloc(sizeof(cplx) * V);
    cplx* xold = (cplx*)malloc(sizeof(cplx) * V);
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * D);
    cplx* x = (cplx*)xout;
    op->adj(s, yin, (double*)b);
    if (x0in) memcpy(xold, x0in, sizeof(cplx) * V);
    else memcpy(xold, b, sizeof(cplx) * V);
    const double bb = dot_re(b, b, V);
    double thr = resThresh;
    if (resThresh_rel >= 0) thr = thr >= 0 ? fmax(thr, resThresh_rel * bb) : resThresh_rel * bb;
    /* r = b - E'E x0, p = r */
    op->fwd(s, (const double*)xold, (double*)tmp);
    op->adj(s, (const double*)tmp, (double*)q);
    for (size_t i = 0; i < V; ++i) { r[i] = b[i] - q[i]; p[i] = r[i]; }
    double rr_old = dot_re(r, r, V);
    if (resSqAll) resSqAll[0] = rr_old;
    if (expResN <= 0) expResN = 100;
    int nit = 0, finished = NitMax <= 0;
    while (!finished) {
        op->fwd(s, (const double*)p, (double*)tmp);
        op->adj(s, (const double*)tmp, (double*)q);
        const cplx alph = rr_old / dot_c(p, q, V);
        for (size_t i = 0; i < V; ++i) x[i] = xold[i] + alph * p[i];
        double betaFact;
        if (nit % expResN == 0) {
            op->fwd(s, (const double*)x, (double*)tmp);
            op->adj(s, (const double*)tmp, (double*)r);
            for (size_t i = 0; i < V; ++i) r[i] = b[i] - r[i];
            betaFact = 0.0;
        } else {
            for (size_t i = 0; i < V; ++i) r[i] -= alph * q[i];
            betaFact = 1.0;
        }
        const double rr = dot_re(r, r, V);
        double xd = 0;
        for (size_t i = 0; i < V; ++i) {
            const cplx d = x[i] - xold[i];
            xd += creal(conj(d) * d);
        }
        if (resSqAll) resSqAll[nit + 1] = rr;
        if (g_verbose) {
            fprintf(stderr, "[mmref] cgne iteration %d: resSq %.6e\n", nit + 1, rr);
            fflush(stderr);
        }
        if (xChangeThresh >= 0 && xd < xChangeThresh) finished = 1;
        if (thr >= 0 && rr < thr) finished = 1;
        const double be = betaFact * rr / rr_old;
        for (size_t i = 0; i < V; ++i) p[i] = r[i] + be * p[i];
        rr_old = rr;
        memcpy(xold, x, sizeof(cplx) * V);
        ++nit;
        if (nit >= NitMax) finished = 1;
    }
    if (NitMax <= 0) memcpy(x, xold, sizeof(cplx) * V);
    free(b); free(r); free(p); free(q); free(xold); free(tmp);
    return nit;
}

int mmref_cgne(mmref_sense* s, const double* yin, const double* x0in, double* xout, int NitMax,
               double resThresh, double resThresh_rel, double xChangeThresh, int expResN, double* resSqAll) {
    const refop op = {sense_fwd, sense_adj, s, (size_t)s->N1 * s->N2, (size_t)s->M * s->C};
    return cgne_core(&op, yin, x0in, xout, NitMax, resThresh, resThresh_rel, xChangeThresh, expResN, resSqAll);
}

/* ---- expanded-encoding (high-order) SENSE operator ------------------------
 * sampHighOrder.m:361-391 (direct worker) composed with rcvrOp.m:145-157:
 *   Phi[v, k] = sum_l B[v, l] H[k, l]     (prepForDirect :496-588; the caller
 *                                          passes the basis values with the
 *                                          mask applied and the B0 map as
 *                                          columns of B, the coefficients and
 *                                          the sample times as columns of H)
 *   y[k, c]  = sum_v exp(+i Phi[v, k]) maps[v, c] x[v] / sqrt(V)
 *   x[v]     = sum_c conj(maps[v, c]) sum_k exp(-i Phi[v, k]) y[k, c] / sqrt(V)
 * Optionally exp(i Phi) is kept ([K][V], 16 bytes a value) between
 * applications, as repeated solves with one operator allow. */
enum { HIO_CB = 8, HIO_T1 = 32, HIO_T2 = 256 };

typedef struct {
    int V, K, Lt, C, Cp;  /* Cp: C rounded up to a multiple of 8 */
    double* B;        /* [Lt][V] (column-major V x Lt) */
    double* H;        /* [Lt][K] (column-major K x Lt) */
    cplx* maps;       /* [V][C] */
    double *er, *ei;  /* exp(i Phi) [K][V], or NULL */
    double *zr, *zi;  /* [V][Cp] maps .* x */
    double *wr, *wi;  /* [K][Cp] data transposed */
    int round_io;     /* inputs and outputs of every application rounded to complex64 */
} mmref_hio;

/* v != 0: every application rounds its input and its output to single
 * precision (an FP64 operator behind complex64 I/O) */
void mmref_hio_set_round_io(mmref_hio* s, int v) { s->round_io = v; }

static inline cplx round_c64(cplx z) { return (double)(float)creal(z) + I * (double)(float)cimag(z); }

void mmref_hio_destroy(mmref_hio* s) {
    if (!s) return;
    free(s->B); free(s->H); free(s->maps); free(s->er); free(s->ei);
    free(s->zr); free(s->zi); free(s->wr); free(s->wi);
    free(s);
}

/* exp(i*sgn*Phi) along a run of n voxels at one time point (dir 0: v0..,
 * k fixed) or of n time points at one voxel (dir 1: k0.., v fixed); Phi
 * summed over l in ascending order, as prepForDirect does */
static void hio_exp_run(const mmref_hio* s, int dir, int k, int v, int n, double sgn, double* re, double* im) {
    double phi[4096];
    for (int j0 = 0; j0 < n; j0 += 4096) {
        const int m = n - j0 < 4096 ? n - j0 : 4096;
        if (s->er) {
            for (int j = 0; j < m; ++j) {
                const size_t i = dir == 0 ? (size_t)k * s->V + v + j0 + j : (size_t)(k + j0 + j) * s->V + v;
                re[j0 + j] = s->er[i];
                im[j0 + j] = sgn * s->ei[i];
            }
            continue;
        }
        for (int j = 0; j < m; ++j) phi[j] = 0.0;
        for (int l = 0; l < s->Lt; ++l) {
            if (dir == 0) {
                const double h = s->H[(size_t)l * s->K + k];
                const double* b = s->B + (size_t)l * s->V + v + j0;
                for (int j = 0; j < m; ++j) phi[j] += b[j] * h;
            } else {
                const double b = s->B[(size_t)l * s->V + v];
                const double* h = s->H + (size_t)l * s->K + k + j0;
                for (int j = 0; j < m; ++j) phi[j] += b * h[j];
            }
        }
        for (int j = 0; j < m; ++j) {
            double sn, cs;
            sincos(phi[j], &sn, &cs);
            re[j0 + j] = cs;
            im[j0 + j] = sgn * sn;
        }
    }
}

/* B: column-major V x Lt, H: column-major K x Lt, maps: column-major V x C
 * complex (interleaved) */
mmref_hio* mmref_hio_create(int V, int K, int Lt, const double* B, const double* H, int C, const double* maps,
                            int cache_exp) {
    mmref_hio* s = (mmref_hio*)calloc(1, sizeof(mmref_hio));
    if (!s) return NULL;
    s->V = V; s->K = K; s->Lt = Lt; s->C = C;
    s->B = (double*)malloc(sizeof(double) * (size_t)V * Lt);
    s->H = (double*)malloc(sizeof(double) * (size_t)K * Lt);
    s->maps = (cplx*)malloc(sizeof(cplx) * (size_t)V * C);
    s->Cp = (C + HIO_CB - 1) / HIO_CB * HIO_CB;
    s->zr = (double*)calloc((size_t)V * s->Cp, sizeof(double));
    s->zi = (double*)calloc((size_t)V * s->Cp, sizeof(double));
    s->wr = (double*)calloc((size_t)K * s->Cp, sizeof(double));
    s->wi = (double*)calloc((size_t)K * s->Cp, sizeof(double));
    if (!s->B || !s->H || !s->maps || !s->zr || !s->zi || !s->wr || !s->wi) { mmref_hio_destroy(s); return NULL; }
    memcpy(s->B, B, sizeof(double) * (size_t)V * Lt);
    memcpy(s->H, H, sizeof(double) * (size_t)K * Lt);
    const cplx* mp = (const cplx*)maps;
    for (int v = 0; v < V; ++v)
        for (int c = 0; c < C; ++c) s->maps[(size_t)v * C + c] = mp[v + (size_t)V * c];
    if (cache_exp) {
        s->er = (double*)malloc(sizeof(double) * (size_t)K * V);
        s->ei = (double*)malloc(sizeof(double) * (size_t)K * V);
        if (!s->er || !s->ei) { mmref_hio_destroy(s); return NULL; }
        double *er = s->er, *ei = s->ei;
        s->er = s->ei = NULL;   /* computed while the cache fills */
#pragma omp parallel for schedule(static)
        for (int k = 0; k < K; ++k) hio_exp_run(s, 0, k, 0, V, 1.0, er + (size_t)k * V, ei + (size_t)k * V);
        s->er = er;
        s->ei = ei;
    }
    return s;
}

/* Both directions are complex products of a tile of exp(+-i Phi) values
 * [rows x n] with a [n x C] operand, coils in blocks of 8 (padded with
 * zeros), two rows at a time so the 32 accumulators stay in registers. */

typedef double v4d __attribute__((vector_size(32)));

static inline v4d ld4(const double* p) { v4d v; memcpy(&v, p, sizeof(v)); return v; }
static inline void st4(double* p, v4d v) { memcpy(p, &v, sizeof(v)); }

/* acc0 (row 0) and acc1 (row 1), each [re 8][im 8], += sum_j e_row[j] z[j][0..7] */
static inline void mac2x8(int n, const double* er0, const double* ei0, const double* er1, const double* ei1,
                          const double* zr, const double* zi, size_t ldz, double* acc0, double* acc1) {
    v4d a0r0 = ld4(acc0), a0r1 = ld4(acc0 + 4), a0i0 = ld4(acc0 + 8), a0i1 = ld4(acc0 + 12);
    v4d a1r0 = ld4(acc1), a1r1 = ld4(acc1 + 4), a1i0 = ld4(acc1 + 8), a1i1 = ld4(acc1 + 12);
    for (int j = 0; j < n; ++j) {
        const double* r = zr + (size_t)j * ldz;
        const double* i = zi + (size_t)j * ldz;
        const v4d r0 = ld4(r), r1 = ld4(r + 4), i0 = ld4(i), i1 = ld4(i + 4);
        const v4d e0r = {er0[j], er0[j], er0[j], er0[j]}, e0i = {ei0[j], ei0[j], ei0[j], ei0[j]};
        const v4d e1r = {er1[j], er1[j], er1[j], er1[j]}, e1i = {ei1[j], ei1[j], ei1[j], ei1[j]};
        a0r0 += e0r * r0 - e0i * i0;
        a0r1 += e0r * r1 - e0i * i1;
        a0i0 += e0r * i0 + e0i * r0;
        a0i1 += e0r * i1 + e0i * r1;
        a1r0 += e1r * r0 - e1i * i0;
        a1r1 += e1r * r1 - e1i * i1;
        a1i0 += e1r * i0 + e1i * r0;
        a1i1 += e1r * i1 + e1i * r1;
    }
    st4(acc0, a0r0); st4(acc0 + 4, a0r1); st4(acc0 + 8, a0i0); st4(acc0 + 12, a0i1);
    st4(acc1, a1r0); st4(acc1 + 4, a1r1); st4(acc1 + 8, a1i0); st4(acc1 + 12, a1i1);
}

/* acc[rows][Cp/8][2][8] += E[rows x n] * Z[n x Cp]; E rows at stride lde */
static void tile_mac(int rows, int n, const double* er, const double* ei, int lde, const double* zr,
                     const double* zi, int Cp, double* acc) {
    static const double zero[HIO_T2] = {0};
    const int nb = Cp / HIO_CB;
    for (int cb = 0; cb < nb; ++cb)
        for (int r = 0; r < rows; r += 2) {
            const int r1 = r + 1 < rows ? r + 1 : r;   /* odd count: the last row twice, second result dropped */
            double dummy[2 * HIO_CB];
            double* a0 = acc + ((size_t)r * nb + cb) * 2 * HIO_CB;
            double* a1 = r + 1 < rows ? acc + ((size_t)r1 * nb + cb) * 2 * HIO_CB : dummy;
            if (a1 == dummy) memset(dummy, 0, sizeof(dummy));
            mac2x8(n, er + (size_t)r * lde, ei + (size_t)r * lde, r + 1 < rows ? er + (size_t)r1 * lde : zero,
                   r + 1 < rows ? ei + (size_t)r1 * lde : zero, zr + cb * HIO_CB, zi + cb * HIO_CB, Cp, a0, a1);
        }
}

void mmref_hio_forward(mmref_hio* s, const double* xin, double* yout) {
    const int V = s->V, K = s->K, C = s->C, Cp = s->Cp;
    const cplx* x = (const cplx*)xin;
    cplx* y = (cplx*)yout;
    const double scale = 1.0 / sqrt((double)V);
#pragma omp parallel for schedule(static)
    for (int v = 0; v < V; ++v)
        for (int c = 0; c < C; ++c) {
            const cplx z = s->maps[(size_t)v * C + c] * (s->round_io ? round_c64(x[v]) : x[v]);
            s->zr[(size_t)v * Cp + c] = creal(z);
            s->zi[(size_t)v * Cp + c] = cimag(z);
        }
#pragma omp parallel
    {
        double* acc = (double*)malloc(sizeof(double) * HIO_T1 * 2 * Cp);
        double* er = (double*)malloc(sizeof(double) * HIO_T1 * HIO_T2);
        double* ei = (double*)malloc(sizeof(double) * HIO_T1 * HIO_T2);
#pragma omp for schedule(dynamic)
        for (int k0 = 0; k0 < K; k0 += HIO_T1) {
            const int nk = k0 + HIO_T1 < K ? HIO_T1 : K - k0;
            memset(acc, 0, sizeof(double) * HIO_T1 * 2 * Cp);
            for (int v0 = 0; v0 < V; v0 += HIO_T2) {
                const int nv = v0 + HIO_T2 < V ? HIO_T2 : V - v0;
                for (int k = 0; k < nk; ++k) hio_exp_run(s, 0, k0 + k, v0, nv, 1.0, er + k * HIO_T2, ei + k * HIO_T2);
                tile_mac(nk, nv, er, ei, HIO_T2, s->zr + (size_t)v0 * Cp, s->zi + (size_t)v0 * Cp, Cp, acc);
            }
            const int nb = Cp / HIO_CB;
            for (int k = 0; k < nk; ++k)
                for (int c = 0; c < C; ++c) {
                    const double* a = acc + ((size_t)k * nb + c / HIO_CB) * 2 * HIO_CB;
                    const cplx r = scale * (a[c % HIO_CB] + I * a[HIO_CB + c % HIO_CB]);
                    y[k0 + k + (size_t)K * c] = s->round_io ? round_c64(r) : r;
                }
        }
        free(acc); free(er); free(ei);
    }
}

void mmref_hio_adjoint(mmref_hio* s, const double* yin, double* xout) {
    const int V = s->V, K = s->K, C = s->C, Cp = s->Cp;
    const cplx* y = (const cplx*)yin;
    cplx* x = (cplx*)xout;
    const double scale = 1.0 / sqrt((double)V);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < K; ++k)
        for (int c = 0; c < C; ++c) {
            const cplx u = s->round_io ? round_c64(y[k + (size_t)K * c]) : y[k + (size_t)K * c];
            s->wr[(size_t)k * Cp + c] = creal(u);
            s->wi[(size_t)k * Cp + c] = cimag(u);
        }
#pragma omp parallel
    {
        double* acc = (double*)malloc(sizeof(double) * HIO_T1 * 2 * Cp);
        double* er = (double*)malloc(sizeof(double) * HIO_T1 * HIO_T2);
        double* ei = (double*)malloc(sizeof(double) * HIO_T1 * HIO_T2);
#pragma omp for schedule(dynamic)
        for (int v0 = 0; v0 < V; v0 += HIO_T1) {
            const int nv = v0 + HIO_T1 < V ? HIO_T1 : V - v0;
            memset(acc, 0, sizeof(double) * HIO_T1 * 2 * Cp);
            for (int k0 = 0; k0 < K; k0 += HIO_T2) {
                const int nk = k0 + HIO_T2 < K ? HIO_T2 : K - k0;
                for (int v = 0; v < nv; ++v)
                    hio_exp_run(s, 1, k0, v0 + v, nk, -1.0, er + v * HIO_T2, ei + v * HIO_T2);   /* exp(-i Phi) */
                tile_mac(nv, nk, er, ei, HIO_T2, s->wr + (size_t)k0 * Cp, s->wi + (size_t)k0 * Cp, Cp, acc);
            }
            const int nb = Cp / HIO_CB;
            for (int v = 0; v < nv; ++v) {
                cplx t = 0;
                for (int c = 0; c < C; ++c) {
                    const double* a = acc + ((size_t)v * nb + c / HIO_CB) * 2 * HIO_CB;
                    t += conj(s->maps[(size_t)(v0 + v) * C + c]) * (a[c % HIO_CB] + I * a[HIO_CB + c % HIO_CB]);
                }
                x[v0 + v] = s->round_io ? round_c64(scale * t) : scale * t;
            }
        }
        free(acc); free(er); free(ei);
    }
}

static void hio_fwd(void* s, const double* x, double* y) { mmref_hio_forward((mmref_hio*)s, x, y); }
static void hio_adj(void* s, const double* y, double* x) { mmref_hio_adjoint((mmref_hio*)s, y, x); }

int mmref_hio_cgne(mmref_hio* s, const double* yin, const double* x0in, double* xout, int NitMax,
                   double resThresh, double resThresh_rel, double xChangeThresh, int expResN, double* resSqAll) {
    const refop op = {hio_fwd, hio_adj, s, (size_t)s->V, (size_t)s->K * s->C};
    return cgne_core(&op, yin, x0in, xout, NitMax, resThresh, resThresh_rel, xChangeThresh, expResN, resSqAll);
}
