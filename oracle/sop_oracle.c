/*
 * sop_oracle.c -- CPU restatement of the SOP-GPU hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in the product uses it.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp (see oracle/Makefile).  With
 * contraction off every float expression below rounds exactly as written,
 * which is what makes it a bit-level statement of the kernels.
 *
 * Reference files restated here (relative to BarsegovGroup/SOP-GPU):
 *   src/Potentials/covalent.cu:65-109, covalent_kernel.cu:28-77
 *   src/Potentials/native.cu:58-107,   native_kernel.cu:27-82
 *   src/Potentials/pairs.cu:18-49,     pairs_kernel.cu:33-91
 *   src/Updaters/possiblepairlist_kernel.cu:18-71
 *   src/Updaters/pairlist.cu:23-128,   pairlist_kernel.cu:33-76
 *   src/Integrators/langevin.cu:17-82, langevin_kernel.cu:17-58
 *   src/Integrators/ht.cu:40-106,      ran2.cpp:51-89
 *   src/Potentials/pulling.cu:36-214
 *   src/Potentials/indentation.cu:84-200,414-550, indentation_kernel.cu:71-128
 *   src/Updaters/output_manager.cu:73-183
 *   src/gsop.cu:111-163 (cadence), 255-271 (energy sums)
 */
#include "sop_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static char g_err[512];
const char* or_error(void) { return g_err; }

typedef struct { float x, y, z, w; } f4;

struct or_sim {
    or_params p;
    int N, ntr, n;
    f4* coord;
    f4* force;
    f4* ext;
    f4* tipF;
    float* T;
    uint32_t* rng; /* 4 per bead */
    /* slot-major maps, entry k of bead i at [k*n+i] */
    int *cov_cnt, *cov_j; float* cov_r0;
    int *nat_cnt, *nat_j; float *nat_r02, *nat_m12;
    int *pair_cnt, *pair_j;
    int *pp_cnt, *pp_j;
    int overflow;
    /* topology copies for output (natives in .top order) */
    int nn; int *ni, *nj; float* nr0;
    /* derived constants (computed as the reference's constructors do) */
    float R_limit_sq, a2, minus6elovera2, pairsCutoff2;
    float hOverZeta, tempNorm, var;
    /* pulling */
    int nfixed, npulled; int *fixed, *pulled;
    /* per trajectory, as cantCoord0/cantCoord/pullVector/extForce[traj] (pulling.cu:66-85) */
    float (*pullVec)[3], (*cant0)[3], (*cant)[3], (*extForce)[3];
    int pullingUpdater;
    /* indentation */
    float tip[3], chip[3], chip0[3], dir[3], micaR[3], micaR0[3], micaN[3];
    float tipAp, tipBp, surfAp, surfBp, V, dx;
    /* cadence */
    long long step;
    int stride;
    /* dat rows */
    double* dat; int ndat, capdat;
    /* pull .dat rows (traj, step, e, ex, f, fx, fy, fz) and indentation .dat
     * rows (step, dx, |f| avg, f mod, Ks dr, Ks dr avg, f xyz, tip xyz,
     * chip xyz), as the reference prints them */
    double* prow; int nprow, capprow;
    double* irow; int nirow, capirow;
    float chipAv[3], tipAv[3], fav[4], kDeltaXAv;
    int nthreads;
    /* hydrodynamics (bdhitea.cu) */
    f4 *rforce, *mforce, *mforce0, *hcoord, *ci;
    float* eps;
    float* beta;       /* per trajectory */
    double* eps_traj;  /* per trajectory, last TEA update */
    float* L;          /* per trajectory (3N)^2 factor, row-major lower */
};

/* ------------------------------------------------------------------ */
/* RNG: src/Integrators/ran2.cpp:51-89 (Numerical Recipes ran2) with its
 * process-static state made explicit, and ht.cu:40-106.                */
/* ------------------------------------------------------------------ */
#define IM1 2147483563
#define IM2 2147483399
#define AM (1.0 / IM1)
#define IMM1 (IM1 - 1)
#define IA1 40014
#define IA2 40692
#define IQ1 53668
#define IQ2 52774
#define IR1 12211
#define IR2 3791
#define NTAB 32
#define NDIV (1 + IMM1 / NTAB)
#define EPS1 1.2e-7
#define RNMX (1.0 - EPS1)

void or_ran2_init(or_ran2_state* st) {
    st->idum2 = 123456789;
    st->iy = 0;
    memset(st->iv, 0, sizeof(st->iv));
}

double or_ran2(int* idum, or_ran2_state* st) {
    int j, k;
    double t;
    if (*idum <= 0) { /* table load only for non-positive seeds (ran2.cpp:58) */
        if (-(*idum) < 1) *idum = 1;
        else *idum = -(*idum);
        st->idum2 = *idum;
        for (j = NTAB + 7; j >= 0; j--) {
            k = (*idum) / IQ1;
            *idum = IA1 * (*idum - k * IQ1) - k * IR1;
            if (*idum < 0) *idum += IM1;
            if (j < NTAB) st->iv[j] = *idum;
        }
        st->iy = st->iv[0];
    }
    k = (*idum) / IQ1;
    *idum = IA1 * (*idum - k * IQ1) - k * IR1;
    if (*idum < 0) *idum += IM1;
    k = st->idum2 / IQ2;
    st->idum2 = IA2 * (st->idum2 - k * IQ2) - k * IR2;
    if (st->idum2 < 0) st->idum2 += IM2;
    j = st->iy / NDIV;
    st->iy = st->iv[j] - st->idum2;
    st->iv[j] = *idum;
    if (st->iy < 1) st->iy += IMM1;
    if ((t = AM * st->iy) > RNMX) return RNMX;
    return t;
}

/* ht.cu:40-56: four components per bead, each redrawn while < 128. */
void or_generate_seeds(int seed, int n, uint32_t* out) {
    or_ran2_state st;
    or_ran2_init(&st);
    for (int i = 0; i < n; i++) {
        for (int c = 0; c < 4; c++) {
            uint32_t v;
            do {
                v = (uint32_t)(or_ran2(&seed, &st) * 4294967295.0);
            } while (v < 128u);
            out[4 * i + c] = v;
        }
    }
}

static inline uint32_t taus_step(uint32_t* z, int S1, int S2, int S3, uint32_t M) {
    uint32_t b = (((*z << S1) ^ *z) >> S2);
    return *z = (((*z & M) << S3) ^ b);
}

/* ht.cu:80-90 */
uint32_t or_hybrid_taus(uint32_t* s) {
    uint32_t a = taus_step(&s[0], 13, 19, 12, 4294967294u);
    uint32_t b = taus_step(&s[1], 2, 25, 4, 4294967288u);
    uint32_t c = taus_step(&s[2], 3, 11, 17, 4294967280u);
    s[3] = 1664525u * s[3] + 1013904223u;
    return a ^ b ^ c ^ s[3];
}

/* ht.cu:61-69 */
float or_uint_to_float(uint32_t u) {
    uint32_t bits = 0x3f800000u | (0x007fffffu & u);
    float f;
    memcpy(&f, &bits, 4);
    f = f - 1.0f;
    return f == 0.0f ? 1.0e-8f : f;
}

/* Natural log by exponent split and the atanh series
 *   log(m) = 2 (s + s^3/3 + s^5/5 + s^7/7 + s^9/9),  s = (m-1)/(m+1),
 * m in [sqrt(1/2), sqrt(2)], |s| <= 0.1716, plus e*ln2 with ln2 split in
 * a short high part and a correction (Cody & Waite).  The GPU kernel
 * evaluates the same operations in the same order. */
float or_logf(float x) {
    uint32_t bits;
    memcpy(&bits, &x, 4);
    int e = (int)((bits >> 23) & 0xffu) - 127;
    uint32_t mb = (bits & 0x007fffffu) | 0x3f800000u;
    float m;
    memcpy(&m, &mb, 4);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e = e + 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float z = s * s;
    float p = 0.222222222f;
    p = p * z + 0.285714298f;
    p = p * z + 0.400000006f;
    p = p * z + 0.666666687f;
    p = p * z + 2.0f;
    p = p * s;
    float fe = (float)e;
    return fe * 0.693145751953125f + (p + fe * 1.42860677e-06f);
}

/* sin and cos by Cody-Waite reduction to [-pi/4, pi/4] with pi/2 split
 * in three parts, then Taylor polynomials to r^9 / r^10. */
void or_sincosf(float x, float* sp, float* cp) {
    float qf = rintf(x * 0.636619772f);
    int q = (int)qf;
    float r = x - qf * 1.5703125f;
    r = r - qf * 4.83751297e-04f;
    r = r - qf * 7.54978995e-08f;
    float z = r * r;
    float ps = 2.75573192e-06f;
    ps = ps * z - 1.98412698e-04f;
    ps = ps * z + 8.33333377e-03f;
    ps = ps * z - 1.66666672e-01f;
    float sn = r + r * (z * ps);
    float pc = -2.75573188e-07f;
    pc = pc * z + 2.48015876e-05f;
    pc = pc * z - 1.38888892e-03f;
    pc = pc * z + 4.16666679e-02f;
    float cs = (1.0f - 0.5f * z) + (z * z) * pc;
    switch (q & 3) {
    case 0: *sp = sn; *cp = cs; break;
    case 1: *sp = cs; *cp = -sn; break;
    case 2: *sp = -sn; *cp = -cs; break;
    default: *sp = -cs; *cp = sn; break;
    }
}

/* ht.cu:93-106: four draws, Box-Muller pairs (x,y) and (z,w).  theta is
 * formed in double as the reference's 2.0f*M_PI*u expression does. */
void or_rforce(uint32_t* s, float* out) {
    float r = sqrtf(-2.0f * or_logf(or_uint_to_float(or_hybrid_taus(s))));
    float th = (float)(6.283185307179586 * (double)or_uint_to_float(or_hybrid_taus(s)));
    float sn, cs;
    or_sincosf(th, &sn, &cs);
    out[0] = r * sn;
    out[1] = r * cs;
    r = sqrtf(-2.0f * or_logf(or_uint_to_float(or_hybrid_taus(s))));
    th = (float)(6.283185307179586 * (double)or_uint_to_float(or_hybrid_taus(s)));
    or_sincosf(th, &sn, &cs);
    out[2] = r * sn;
    out[3] = r * cs;
}

/* ------------------------------------------------------------------ */
/* construction                                                         */
/* ------------------------------------------------------------------ */
static void* xcalloc(size_t n, size_t sz) {
    void* p = calloc(n ? n : 1, sz);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

static int contains(const int* v, int n, int x) {
    for (int i = 0; i < n; i++) if (v[i] == x) return 1;
    return 0;
}

or_sim* or_create(const or_params* p, int N, int ntr, const float* coords4,
                  int nb, const int* bi, const int* bj, const float* br0,
                  int nn, const int* ni, const int* nj, const float* nr0, const float* neh,
                  int np, const int* pi, const int* pj,
                  int nfixed, const int* fixed, int npulled, const int* pulled) {
    or_sim* s = (or_sim*)xcalloc(1, sizeof(or_sim));
    s->p = *p;
    s->N = N; s->ntr = ntr; s->n = N * ntr;
    int n = s->n;
    s->nthreads = 1;
    s->coord = (f4*)xcalloc(n, sizeof(f4));
    s->force = (f4*)xcalloc(n, sizeof(f4));
    s->ext = (f4*)xcalloc(n, sizeof(f4));
    s->tipF = (f4*)xcalloc(n, sizeof(f4));
    s->T = (float*)xcalloc(n, sizeof(float));
    s->rng = (uint32_t*)xcalloc((size_t)4 * n, sizeof(uint32_t));
    for (int t = 0; t < ntr; t++)
        for (int i = 0; i < N; i++) {
            f4* c = &s->coord[t * N + i];
            c->x = coords4[4 * (t * N + i) + 0];
            c->y = coords4[4 * (t * N + i) + 1];
            c->z = coords4[4 * (t * N + i) + 2];
            c->w = 0.0f; /* gsop.cu:196 */
        }
    /* covalent map, covalent.cu:65-109 */
    int mc = p->max_covalent;
    s->cov_cnt = (int*)xcalloc(n, sizeof(int));
    s->cov_j = (int*)xcalloc((size_t)mc * n, sizeof(int));
    s->cov_r0 = (float*)xcalloc((size_t)mc * n, sizeof(float));
    for (int k = 0; k < nb; k++) {
        int i = bi[k], j = bj[k];
        if (s->cov_cnt[i] >= mc || s->cov_cnt[j] >= mc || (i == j && s->cov_cnt[i] + 1 >= mc)) {
            snprintf(g_err, sizeof g_err, "Maximum number of covalent bonds exceeded the limit of %d.", mc);
            or_destroy(s);
            return NULL;
        }
        s->cov_j[s->cov_cnt[i] * n + i] = j;
        s->cov_r0[s->cov_cnt[i] * n + i] = br0[k];
        s->cov_cnt[i]++;
        s->cov_j[s->cov_cnt[j] * n + j] = i;
        s->cov_r0[s->cov_cnt[j] * n + j] = br0[k];
        s->cov_cnt[j]++;
    }
    /* native map, native.cu:58-107; parameters formed in double */
    int mn = p->max_native;
    s->nat_cnt = (int*)xcalloc(n, sizeof(int));
    s->nat_j = (int*)xcalloc((size_t)mn * n, sizeof(int));
    s->nat_r02 = (float*)xcalloc((size_t)mn * n, sizeof(float));
    s->nat_m12 = (float*)xcalloc((size_t)mn * n, sizeof(float));
    for (int k = 0; k < nn; k++) {
        int i = ni[k], j = nj[k];
        if (s->nat_cnt[i] >= mn || s->nat_cnt[j] >= mn) {
            snprintf(g_err, sizeof g_err, "ERROR: Maximum number of native contacts exceeded the limit of %d.", mn);
            or_destroy(s);
            return NULL;
        }
        float r02 = nr0[k] * nr0[k];
        float m12 = (float)(-12.0 * (double)neh[k] / (double)(nr0[k] * nr0[k]));
        s->nat_j[s->nat_cnt[i] * n + i] = j;
        s->nat_r02[s->nat_cnt[i] * n + i] = r02;
        s->nat_m12[s->nat_cnt[i] * n + i] = m12;
        s->nat_cnt[i]++;
        s->nat_j[s->nat_cnt[j] * n + j] = i;
        s->nat_r02[s->nat_cnt[j] * n + j] = r02;
        s->nat_m12[s->nat_cnt[j] * n + j] = m12;
        s->nat_cnt[j]++;
    }
    /* initial possible list from [ pairs ], pairlist.cu:41-57 */
    int mpp = p->max_possiblePairs;
    s->pp_cnt = (int*)xcalloc(n, sizeof(int));
    s->pp_j = (int*)xcalloc((size_t)mpp * n, sizeof(int));
    for (int k = 0; k < np; k++) {
        int i = pi[k], j = pj[k];
        if (s->pp_cnt[i] >= mpp || s->pp_cnt[j] >= mpp) {
            snprintf(g_err, sizeof g_err, "ERROR: Maximum number of possible pairs exceeded the limit of %d.", mpp);
            or_destroy(s);
            return NULL;
        }
        s->pp_j[s->pp_cnt[i] * n + i] = j; s->pp_cnt[i]++;
        s->pp_j[s->pp_cnt[j] * n + j] = i; s->pp_cnt[j]++;
    }
    /* replicate for many-runs mode */
    for (int t = 1; t < ntr; t++)
        for (int i = 0; i < N; i++) {
            int d = t * N + i;
            s->cov_cnt[d] = s->cov_cnt[i];
            for (int k = 0; k < mc; k++) { s->cov_j[k * n + d] = s->cov_j[k * n + i] + t * N; s->cov_r0[k * n + d] = s->cov_r0[k * n + i]; }
            s->nat_cnt[d] = s->nat_cnt[i];
            for (int k = 0; k < mn; k++) { s->nat_j[k * n + d] = s->nat_j[k * n + i] + t * N; s->nat_r02[k * n + d] = s->nat_r02[k * n + i]; s->nat_m12[k * n + d] = s->nat_m12[k * n + i]; }
            s->pp_cnt[d] = s->pp_cnt[i];
            for (int k = 0; k < mpp; k++) s->pp_j[k * n + d] = s->pp_j[k * n + i] + t * N;
        }
    s->pair_cnt = (int*)xcalloc(n, sizeof(int));
    s->pair_j = (int*)xcalloc((size_t)p->max_pairs * n, sizeof(int));
    s->nn = nn;
    s->ni = (int*)xcalloc(nn, sizeof(int)); s->nj = (int*)xcalloc(nn, sizeof(int)); s->nr0 = (float*)xcalloc(nn, sizeof(float));
    memcpy(s->ni, ni, nn * sizeof(int)); memcpy(s->nj, nj, nn * sizeof(int)); memcpy(s->nr0, nr0, nn * sizeof(float));

    /* derived constants */
    s->R_limit_sq = p->R_limit * p->R_limit;              /* covalent.cu:84 */
    s->a2 = p->a * p->a;                                   /* pairs.cu:24 */
    s->minus6elovera2 = -6.0f * p->el / s->a2;             /* pairs.cu:26 */
    s->pairsCutoff2 = p->pairs_cutoff * p->pairs_cutoff;   /* pairs.cu:29 */
    s->hOverZeta = p->timestep / p->zeta;                  /* langevin.cu:21 */
    s->tempNorm = (float)((double)p->zeta / (6.0 * (double)p->timestep)); /* langevin.cu:22 */
    {
        float temp = p->heating ? p->initialT : p->temperature;
        s->var = sqrtf(2.0f * temp * p->zeta / p->timestep); /* langevin.cu:52 */
    }
    /* RNG seeds: langevin.cu:23-24 */
    or_generate_seeds(p->seed, n, s->rng);

    /* pulling, pulling.cu:36-146 */
    if (p->pulling) {
        if (p->deltax == 0.0f && p->fconst == 0.0f) {
            snprintf(g_err, sizeof g_err, "ERROR: Either 'deltax' or 'fconst' parameter should be specified to initiate pulling");
            or_destroy(s);
            return NULL;
        }
        s->pullingUpdater = p->deltax != 0.0f;
        s->nfixed = nfixed; s->npulled = npulled;
        s->fixed = (int*)xcalloc(nfixed, sizeof(int)); s->pulled = (int*)xcalloc(npulled, sizeof(int));
        memcpy(s->fixed, fixed, nfixed * sizeof(int)); memcpy(s->pulled, pulled, npulled * sizeof(int));
        /* each trajectory's pulling geometry from its own coordinates
         * (pulling.cu:62-92) */
        s->pullVec = (float(*)[3])xcalloc(ntr, sizeof(float[3]));
        s->cant0 = (float(*)[3])xcalloc(ntr, sizeof(float[3]));
        s->cant = (float(*)[3])xcalloc(ntr, sizeof(float[3]));
        s->extForce = (float(*)[3])xcalloc(ntr, sizeof(float[3]));
        for (int t = 0; t < ntr; t++) {
            const f4* c = s->coord + (size_t)t * N;
            if (p->pullDirectionVector) {
                s->pullVec[t][0] = p->pullVector[0]; s->pullVec[t][1] = p->pullVector[1]; s->pullVec[t][2] = p->pullVector[2];
            } else {
                float vx = c[p->pulledEnd].x - c[p->fixedEnd].x;
                float vy = c[p->pulledEnd].y - c[p->fixedEnd].y;
                float vz = c[p->pulledEnd].z - c[p->fixedEnd].z;
                float v = sqrtf(vx * vx + vy * vy + vz * vz);
                s->pullVec[t][0] = vx / v; s->pullVec[t][1] = vy / v; s->pullVec[t][2] = vz / v;
            }
            s->cant0[t][0] = s->cant[t][0] = c[p->pulledEnd].x;
            s->cant0[t][1] = s->cant[t][1] = c[p->pulledEnd].y;
            s->cant0[t][2] = s->cant[t][2] = c[p->pulledEnd].z;
        }
        for (int t = 0; t < ntr; t++) {
            for (int k = 0; k < npulled; k++) {
                int i = pulled[k];
                if (i > N || i < 0) {
                    snprintf(g_err, sizeof g_err, "Invalid pulling bead #%d; should be >=0 and <%d", i, N);
                    or_destroy(s);
                    return NULL;
                }
                f4* e = &s->ext[t * N + i];
                e->x = s->pullVec[t][0] * p->fconst; e->y = s->pullVec[t][1] * p->fconst; e->z = s->pullVec[t][2] * p->fconst; e->w = 2.0f;
            }
            for (int k = 0; k < nfixed; k++) {
                int i = fixed[k];
                if (i > N || i < 0) {
                    snprintf(g_err, sizeof g_err, "Invalid fixed bead #%d; should be >=0 and <%d", i, N);
                    or_destroy(s);
                    return NULL;
                }
                f4* e = &s->ext[t * N + i];
                e->x = 0.0f; e->y = 0.0f; e->z = 0.0f; e->w = 1.0f;
            }
        }
    }
    /* indentation, indentation.cu:84-200 (continuous surface) */
    if (p->indentation) {
        if (ntr > 1) {
            snprintf(g_err, sizeof g_err, "Many-runs-per-GPU mode is not supported by indentation potential.");
            or_destroy(s);
            return NULL;
        }
        for (int d = 0; d < 3; d++) {
            s->chip[d] = s->chip0[d] = p->indChip[d];
            s->tip[d] = p->indTip[d];
            s->dir[d] = p->indDirection[d];
            s->micaR0[d] = s->micaR[d] = p->indSurfR0[d];
            s->micaN[d] = p->indSurfN[d];
        }
        s->V = p->indDeltaX;
        s->dx = 0.0f;
        s->tipAp = 12.0f * p->indTipEl * powf(p->indTipSigma, 12.0f) * p->indTipA;
        s->tipBp = 6.0f * p->indTipEl * powf(p->indTipSigma, 6.0f) * p->indTipB;
        s->surfAp = 12.0f * p->indSurfEl * powf(p->indSurfSigma, 12.0f) * p->indSurfA;
        s->surfBp = 6.0f * p->indSurfEl * powf(p->indSurfSigma, 6.0f) * p->indSurfB;
        float norm = sqrtf(s->dir[0] * s->dir[0] + s->dir[1] * s->dir[1] + s->dir[2] * s->dir[2]);
        s->dir[0] /= norm; s->dir[1] /= norm; s->dir[2] /= norm;
        norm = sqrtf(s->micaN[0] * s->micaN[0] + s->micaN[1] * s->micaN[1] + s->micaN[2] * s->micaN[2]);
        s->micaN[0] /= norm; s->micaN[1] /= norm; s->micaN[2] /= norm;
    }
    /* hydrodynamics, bdhitea.cu:30-100 */
    if (p->hi_on) {
        if (p->minimization) {
            snprintf(g_err, sizeof g_err, "Using minimization with BDHI integrator is not supported and will never be.");
            or_destroy(s);
            return NULL;
        }
        s->rforce = (f4*)xcalloc(n, sizeof(f4));
        s->mforce = (f4*)xcalloc(n, sizeof(f4));
        s->mforce0 = (f4*)xcalloc(n, sizeof(f4));
        s->hcoord = (f4*)xcalloc(n, sizeof(f4));
        s->ci = (f4*)xcalloc(n, sizeof(f4));
        s->eps = (float*)xcalloc(n, sizeof(float));
        s->beta = (float*)xcalloc(ntr, sizeof(float));
        s->eps_traj = (double*)xcalloc(ntr, sizeof(double));
        if (p->hi_exact) s->L = (float*)xcalloc((size_t)ntr * 9 * N * N, sizeof(float));
    }
    /* stride: gsop.cu:116-122 over the updaters initFF creates */
    {
        int st = p->nav;
#define MINF(f) do { if ((f) < st) st = (f); } while (0)
        if (p->indentation) MINF(p->nav);
        if (p->pulling && s->pullingUpdater) MINF(p->pullFreq);
        MINF(p->possiblepairs_freq);
        MINF(p->outputtiming);
        MINF(p->pairs_freq);
        MINF(p->dcdfreq);
        if (p->heating) MINF(p->tempFreq);
        if (p->hi_on && !p->hi_exact) MINF(p->hi_epsilon_freq);
        if (p->updater_freq_min > 0) MINF(p->updater_freq_min); /* updaters added by the host */
#undef MINF
        s->stride = st;
    }
    s->capdat = 0;
    return s;
}

void or_destroy(or_sim* s) {
    if (!s) return;
    free(s->coord); free(s->force); free(s->ext); free(s->tipF); free(s->T); free(s->rng);
    free(s->cov_cnt); free(s->cov_j); free(s->cov_r0);
    free(s->nat_cnt); free(s->nat_j); free(s->nat_r02); free(s->nat_m12);
    free(s->pair_cnt); free(s->pair_j); free(s->pp_cnt); free(s->pp_j);
    free(s->ni); free(s->nj); free(s->nr0); free(s->fixed); free(s->pulled); free(s->dat);
    free(s->prow); free(s->irow);
    free(s->pullVec); free(s->cant0); free(s->cant); free(s->extForce);
    free(s->rforce); free(s->mforce); free(s->mforce0); free(s->hcoord); free(s->ci); free(s->eps); free(s->beta);
    free(s->eps_traj); free(s->L);
    free(s);
}

int or_n(const or_sim* s) { return s->n; }
int or_stride(const or_sim* s) { return s->stride; }
long long or_step(const or_sim* s) { return s->step; }
int or_overflow(const or_sim* s) { return s->overflow; }
void or_get_coords(const or_sim* s, float* o) { memcpy(o, s->coord, sizeof(f4) * s->n); }
void or_set_coords(or_sim* s, const float* in) { memcpy(s->coord, in, sizeof(f4) * s->n); }
void or_get_T(const or_sim* s, float* o) { memcpy(o, s->T, sizeof(float) * s->n); }
void or_get_rng(const or_sim* s, uint32_t* o) { memcpy(o, s->rng, sizeof(uint32_t) * 4 * s->n); }
void or_get_ext(const or_sim* s, float* o) { memcpy(o, s->ext, sizeof(f4) * s->n); }
void or_get_indentation(const or_sim* s, float* o) {
    for (int d = 0; d < 3; d++) { o[d] = s->tip[d]; o[3 + d] = s->chip[d]; o[6 + d] = s->micaR[d]; }
}

int or_list_cap(const or_sim* s, int which) {
    switch (which) {
    case 0: return s->p.max_covalent;
    case 1: return s->p.max_native;
    case 2: return s->p.max_pairs;
    default: return s->p.max_possiblePairs;
    }
}

int or_get_list(const or_sim* s, int which, int* counts, int* entries) {
    const int* c; const int* e;
    switch (which) {
    case 0: c = s->cov_cnt; e = s->cov_j; break;
    case 1: c = s->nat_cnt; e = s->nat_j; break;
    case 2: c = s->pair_cnt; e = s->pair_j; break;
    default: c = s->pp_cnt; e = s->pp_j; break;
    }
    int cap = or_list_cap(s, which);
    memcpy(counts, c, sizeof(int) * s->n);
    memcpy(entries, e, sizeof(int) * (size_t)cap * s->n);
    return cap;
}

/* ------------------------------------------------------------------ */
/* per-bead force terms, FP32, reference expression order              */
/* ------------------------------------------------------------------ */

/* covalent_kernel.cu:36-48 */
static inline void cov_bead(const or_sim* s, int i, f4* f) {
    const int n = s->n;
    const f4 c = s->coord[i];
    for (int k = 0; k < s->cov_cnt[i]; k++) {
        f4 r2 = s->coord[s->cov_j[k * n + i]];
        float r0 = s->cov_r0[k * n + i];
        r2.x -= c.x; r2.y -= c.y; r2.z -= c.z;
        r2.w = sqrtf(r2.x * r2.x + r2.y * r2.y + r2.z * r2.z);
        r2.w -= r0;
        r2.w = s->p.kspring_cov * r2.w / ((r2.w + r0) * (1.0f - r2.w * r2.w / (s->R_limit_sq)));
        f->x += r2.w * r2.x;
        f->y += r2.w * r2.y;
        f->z += r2.w * r2.z;
    }
}

/* native_kernel.cu:36-49 */
static inline void nat_bead(const or_sim* s, int i, f4* f) {
    const int n = s->n;
    const f4 c = s->coord[i];
    for (int k = 0; k < s->nat_cnt[i]; k++) {
        f4 r2 = s->coord[s->nat_j[k * n + i]];
        float r02 = s->nat_r02[k * n + i], m12 = s->nat_m12[k * n + i];
        r2.x -= c.x; r2.y -= c.y; r2.z -= c.z;
        r2.w = r02 / (r2.x * r2.x + r2.y * r2.y + r2.z * r2.z);
        float p6 = r2.w * r2.w * r2.w;
        r2.w = r2.w * (p6 * p6 - p6) * m12;
        f->x += r2.x * r2.w;
        f->y += r2.y * r2.w;
        f->z += r2.z * r2.w;
    }
}

/* pairs_kernel.cu:41-56 */
static inline void pair_bead(const or_sim* s, int i, f4* f) {
    const int n = s->n;
    const f4 c = s->coord[i];
    for (int k = 0; k < s->pair_cnt[i]; k++) {
        f4 r2 = s->coord[s->pair_j[k * n + i]];
        r2.x -= c.x; r2.y -= c.y; r2.z -= c.z;
        r2.w = r2.x * r2.x + r2.y * r2.y + r2.z * r2.z;
        if (r2.w < s->pairsCutoff2) {
            r2.w = s->a2 / r2.w;
            r2.w = r2.w * r2.w;
            r2.w = s->minus6elovera2 * r2.w * r2.w;
            f->x += r2.x * r2.w;
            f->y += r2.y * r2.w;
            f->z += r2.z * r2.w;
        }
    }
}

/* indentation_kernel.cu:71-128 (continuous surface) */
static inline void ind_bead(or_sim* s, int i, f4* f) {
    const f4 c = s->coord[i];
    f4 fin = s->tipF[i];
    f4 df, dr;
    dr.x = c.x - s->tip[0];
    dr.y = c.y - s->tip[1];
    dr.z = c.z - s->tip[2];
    dr.w = sqrtf(dr.x * dr.x + dr.y * dr.y + dr.z * dr.z);
    float r2 = dr.w - s->p.indTipR;
    r2 = 1.0f / r2;
    r2 = r2 * r2;
    float r6 = r2 * r2 * r2;
    df.w = s->tipAp * r6 + s->tipBp;
    df.w = df.w * r6 / (dr.w * (dr.w - s->p.indTipR));
    df.x = dr.x * df.w;
    df.y = dr.y * df.w;
    df.z = dr.z * df.w;
    fin.x += df.x; fin.y += df.y; fin.z += df.z;
    s->tipF[i] = fin;
    dr.x = c.x - s->micaR[0];
    dr.y = c.y - s->micaR[1];
    dr.z = c.z - s->micaR[2];
    dr.w = dr.x * s->micaN[0] + dr.y * s->micaN[1] + dr.z * s->micaN[2];
    dr.w = 1.0f / dr.w;
    dr.w = dr.w * dr.w;
    r6 = dr.w * dr.w * dr.w;
    df.w = s->surfAp * r6 + s->surfBp;
    df.w = df.w * r6 * dr.w;
    f->x += s->micaN[0] * df.w + df.x;
    f->y += s->micaN[1] * df.w + df.y;
    f->z += s->micaN[2] * df.w + df.z;
}

/* langevin_kernel.cu:17-58 */
static inline void langevin_bead(or_sim* s, int i) {
    f4 c = s->coord[i];
    f4 f = s->force[i];
    s->force[i].x = s->force[i].y = s->force[i].z = s->force[i].w = 0.0f;
    if (!s->p.minimization) {
        float var = s->var;
        float df[4];
        or_rforce(&s->rng[4 * i], df);
        f.x += df[0] * var;
        f.y += df[1] * var;
        f.z += df[2] * var;
    }
    if (s->p.pulling) {
        f4 e = s->ext[i];
        if (e.w == 1.0f) { f.x = 0.0f; f.y = 0.0f; f.z = 0.0f; }
        if (e.w == 2.0f) { f.x += e.x; f.y += e.y; f.z += e.z; }
    }
    float mult = s->hOverZeta;
    float dx = mult * f.x, dy = mult * f.y, dz = mult * f.z;
    c.x += dx; c.y += dy; c.z += dz;
    s->T[i] += s->tempNorm * (dx * dx + dy * dy + dz * dz);
    s->coord[i] = c;
}

static void forces_pass(or_sim* s) {
    const int n = s->n;
    /* each kernel is a separate pass that reads coordinates only, so a
     * bead-parallel loop per pass keeps the reference's semantics */
#pragma omp parallel for schedule(static) num_threads(s->nthreads) if (s->nthreads > 1)
    for (int i = 0; i < n; i++) {
        f4 f = s->force[i];
        cov_bead(s, i, &f);
        nat_bead(s, i, &f);
        pair_bead(s, i, &f);
        if (s->p.indentation) ind_bead(s, i, &f);
        s->force[i] = f;
    }
}

static void integrate_pass(or_sim* s) {
    const int n = s->n;
#pragma omp parallel for schedule(static) num_threads(s->nthreads) if (s->nthreads > 1)
    for (int i = 0; i < n; i++) langevin_bead(s, i);
}

void or_forces_f32(or_sim* s, float* out) {
    const int n = s->n;
    f4* save = (f4*)xcalloc(n, sizeof(f4));
    memcpy(save, s->tipF, sizeof(f4) * n);
    for (int i = 0; i < n; i++) {
        f4 f = {0, 0, 0, 0};
        cov_bead(s, i, &f);
        nat_bead(s, i, &f);
        pair_bead(s, i, &f);
        if (s->p.indentation) ind_bead(s, i, &f);
        memcpy(&out[4 * i], &f, sizeof(f4));
    }
    memcpy(s->tipF, save, sizeof(f4) * n);
    free(save);
}

/* FP64 evaluation of the same model (closed forms of the docs,
 * general_sop.rst:19-39, as the kernels implement them). */
void or_forces_f64(or_sim* s, double* out) {
    const int n = s->n;
    const or_params* p = &s->p;
    for (int i = 0; i < n; i++) {
        double fx = 0, fy = 0, fz = 0;
        const f4 c = s->coord[i];
        for (int k = 0; k < s->cov_cnt[i]; k++) {
            f4 o = s->coord[s->cov_j[k * n + i]];
            double x = (double)o.x - c.x, y = (double)o.y - c.y, z = (double)o.z - c.z;
            double r = sqrt(x * x + y * y + z * z), r0 = s->cov_r0[k * n + i], dr = r - r0;
            double R2 = (double)p->R_limit * p->R_limit;
            double w = p->kspring_cov * dr / (r * (1.0 - dr * dr / R2));
            fx += w * x; fy += w * y; fz += w * z;
        }
        for (int k = 0; k < s->nat_cnt[i]; k++) {
            f4 o = s->coord[s->nat_j[k * n + i]];
            double x = (double)o.x - c.x, y = (double)o.y - c.y, z = (double)o.z - c.z;
            double r2 = x * x + y * y + z * z;
            /* r0^2 and eh recovered exactly as the float map stores them */
            double r02 = s->nat_r02[k * n + i];
            double m12 = s->nat_m12[k * n + i];
            double q = r02 / r2, p6 = q * q * q;
            double w = q * (p6 * p6 - p6) * m12;
            fx += w * x; fy += w * y; fz += w * z;
        }
        for (int k = 0; k < s->pair_cnt[i]; k++) {
            f4 o = s->coord[s->pair_j[k * n + i]];
            double x = (double)o.x - c.x, y = (double)o.y - c.y, z = (double)o.z - c.z;
            double r2 = x * x + y * y + z * z;
            double rc = p->pairs_cutoff;
            if (r2 < rc * rc) {
                double a2 = (double)p->a * p->a;
                double q = a2 / r2;
                double w = -6.0 * p->el / a2 * q * q * q * q;
                fx += w * x; fy += w * y; fz += w * z;
            }
        }
        if (p->indentation) {
            double x = (double)c.x - s->tip[0], y = (double)c.y - s->tip[1], z = (double)c.z - s->tip[2];
            double r = sqrt(x * x + y * y + z * z), d = r - p->indTipR;
            double A = 12.0 * p->indTipEl * pow(p->indTipSigma, 12.0) * p->indTipA;
            double B = 6.0 * p->indTipEl * pow(p->indTipSigma, 6.0) * p->indTipB;
            double w = (A / pow(d, 12) + B / pow(d, 6)) / (r * d);
            fx += w * x; fy += w * y; fz += w * z;
            double h = ((double)c.x - s->micaR[0]) * s->micaN[0] + ((double)c.y - s->micaR[1]) * s->micaN[1] +
                       ((double)c.z - s->micaR[2]) * s->micaN[2];
            double As = 12.0 * p->indSurfEl * pow(p->indSurfSigma, 12.0) * p->indSurfA;
            double Bs = 6.0 * p->indSurfEl * pow(p->indSurfSigma, 6.0) * p->indSurfB;
            double ws = (As / pow(h, 12) + Bs / pow(h, 6)) / (h * h);
            fx += s->micaN[0] * ws; fy += s->micaN[1] * ws; fz += s->micaN[2] * ws;
        }
        out[3 * i] = fx; out[3 * i + 1] = fy; out[3 * i + 2] = fz;
    }
}

/* energy kernels: covalent_kernel.cu:57-77, native_kernel.cu:57-82,
 * pairs_kernel.cu:69-91 */
void or_energies_f32(or_sim* s, float* out) {
    const int n = s->n;
    for (int i = 0; i < n; i++) {
        const f4 c = s->coord[i];
        float pot = 0.0f;
        for (int k = 0; k < s->cov_cnt[i]; k++) {
            f4 r2 = s->coord[s->cov_j[k * n + i]];
            float r0 = s->cov_r0[k * n + i];
            r2.x -= c.x; r2.y -= c.y; r2.z -= c.z;
            r2.w = sqrtf(r2.x * r2.x + r2.y * r2.y + r2.z * r2.z);
            r2.w -= r0;
            pot = (float)((double)pot - ((double)s->p.kspring_cov / 2.0) * (double)s->R_limit_sq *
                                            (double)or_logf(1.0f - r2.w * r2.w / (s->R_limit_sq)));
        }
        out[i] = pot;
        pot = 0.0f;
        for (int k = 0; k < s->nat_cnt[i]; k++) {
            f4 r2 = s->coord[s->nat_j[k * n + i]];
            float r02 = s->nat_r02[k * n + i], m12 = s->nat_m12[k * n + i];
            r2.x -= c.x; r2.y -= c.y; r2.z -= c.z;
            r2.w = r02 / (r2.x * r2.x + r2.y * r2.y + r2.z * r2.z);
            float p6 = r2.w * r2.w * r2.w;
            float p12 = p6 * p6;
            p6 = -2.0f * p6;
            float eh = (float)((double)(-m12 * r02) / 12.0);
            pot += eh * (p6 + p12);
        }
        out[n + i] = pot;
        pot = 0.0f;
        for (int k = 0; k < s->pair_cnt[i]; k++) {
            f4 r2 = s->coord[s->pair_j[k * n + i]];
            r2.x -= c.x; r2.y -= c.y; r2.z -= c.z;
            r2.w = r2.x * r2.x + r2.y * r2.y + r2.z * r2.z;
            if (r2.w < s->pairsCutoff2) {
                r2.w = s->a2 / r2.w;
                pot += s->p.el * r2.w * r2.w * r2.w;
            }
        }
        out[2 * n + i] = pot;
    }
}

void or_energies_f64(const or_sim* s, double* out) {
    const int n = s->n;
    const or_params* p = &s->p;
    for (int i = 0; i < n; i++) {
        const f4 c = s->coord[i];
        double e0 = 0, e1 = 0, e2 = 0;
        for (int k = 0; k < s->cov_cnt[i]; k++) {
            f4 o = s->coord[s->cov_j[k * n + i]];
            double x = (double)o.x - c.x, y = (double)o.y - c.y, z = (double)o.z - c.z;
            double dr = sqrt(x * x + y * y + z * z) - s->cov_r0[k * n + i];
            double R2 = (double)p->R_limit * p->R_limit;
            e0 -= 0.5 * p->kspring_cov * R2 * log(1.0 - dr * dr / R2);
        }
        for (int k = 0; k < s->nat_cnt[i]; k++) {
            f4 o = s->coord[s->nat_j[k * n + i]];
            double x = (double)o.x - c.x, y = (double)o.y - c.y, z = (double)o.z - c.z;
            double r02 = s->nat_r02[k * n + i];
            double eh = -(double)s->nat_m12[k * n + i] * r02 / 12.0;
            double q = r02 / (x * x + y * y + z * z), p6 = q * q * q;
            e1 += eh * (p6 * p6 - 2.0 * p6);
        }
        for (int k = 0; k < s->pair_cnt[i]; k++) {
            f4 o = s->coord[s->pair_j[k * n + i]];
            double x = (double)o.x - c.x, y = (double)o.y - c.y, z = (double)o.z - c.z;
            double r2 = x * x + y * y + z * z;
            double rc = p->pairs_cutoff;
            if (r2 < rc * rc) {
                double q = (double)p->a * p->a / r2;
                e2 += p->el * q * q * q;
            }
        }
        out[i] = e0; out[n + i] = e1; out[2 * n + i] = e2;
    }
}

/* SOPPotential::sumEnergies, gsop.cu:255-271 */
void or_sum_energies(or_sim* s, float* out) {
    const int n = s->n, N = s->N;
    float* e = (float*)xcalloc((size_t)3 * n, sizeof(float));
    or_energies_f32(s, e);
    for (int pot = 0; pot < 3; pot++)
        for (int t = 0; t < s->ntr; t++) {
            double acc = 0.0;
            for (int i = 0; i < N; i++) acc += e[pot * n + t * N + i];
            out[pot * s->ntr + t] = (float)(acc * 0.5f);
        }
    free(e);
}

/* ------------------------------------------------------------------ */
/* pairlists                                                            */
/* ------------------------------------------------------------------ */

/* possiblepairlist_kernel.cu:18-71 */
void or_build_possible(or_sim* s) {
    const int n = s->n, N = s->N, cap = s->p.max_possiblePairs;
    const float thr = s->p.pairs_threshold;
    int ovf = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(s->nthreads) if (s->nthreads > 1) reduction(| : ovf)
    for (int d = 0; d < n; d++) {
        const f4 r1 = s->coord[d];
        int cnt = 0;
        int t = d / N;
        for (int i = t * N; i < (t + 1) * N; i++) {
            if (i == d) continue;
            f4 r2 = s->coord[i];
            r2.x -= r1.x; r2.y -= r1.y; r2.z -= r1.z;
            r2.w = sqrtf(r2.x * r2.x + r2.y * r2.y + r2.z * r2.z);
            if (r2.w < thr) {
                int add = 1;
                for (int j = 0; j < s->cov_cnt[d]; j++) if (i == s->cov_j[j * n + d]) add = 0;
                if (add) for (int j = 0; j < s->nat_cnt[d]; j++) if (i == s->nat_j[j * n + d]) add = 0;
                if (add) {
                    if (cnt < cap) s->pp_j[cnt * n + d] = i;
                    else ovf = 1;
                    cnt++;
                }
            }
        }
        s->pp_cnt[d] = cnt < cap ? cnt : cap;
    }
    s->overflow |= ovf;
}

/* pairlist_kernel.cu:33-76 (the -1 padding is not modelled: counts
 * bound every read) */
void or_build_verlet(or_sim* s) {
    const int n = s->n, cap = s->p.max_pairs;
    const float cut = s->p.pairlist_cutoff;
    int ovf = 0;
#pragma omp parallel for schedule(static) num_threads(s->nthreads) if (s->nthreads > 1) reduction(| : ovf)
    for (int d = 0; d < n; d++) {
        const f4 r1 = s->coord[d];
        int cnt = 0;
        for (int k = 0; k < s->pp_cnt[d]; k++) {
            int i2 = s->pp_j[k * n + d];
            f4 r2 = s->coord[i2];
            r2.x -= r1.x; r2.y -= r1.y; r2.z -= r1.z;
            r2.w = sqrtf(r2.x * r2.x + r2.y * r2.y + r2.z * r2.z);
            if (r2.w < cut) {
                if (cnt < cap) s->pair_j[cnt * n + d] = i2;
                else ovf = 1;
                cnt++;
            }
        }
        s->pair_cnt[d] = cnt < cap ? cnt : cap;
    }
    s->overflow |= ovf;
}


/* ------------------------------------------------------------------ */
/* hydrodynamic interactions, bdhitea.cu + bdhitea_kernel.cu            */
/* ------------------------------------------------------------------ */
typedef struct { float xx, xy, xz, yy, yz, zz; } f6;

/* integrateTea_RPY, bdhitea_kernel.cu:67-87: D_ij / D_ii for a normalised
 * separation dr with its length in w. */
static inline f6 rpy(const f4 dr, float a) {
    f6 d;
    const float ra = dr.w / a;
    float crr, cii;
    if (ra > 2.f) {
        crr = 0.75f / ra * (1.f - 2.f / ra / ra);
        cii = 0.75f / ra * (1.f + 2.f / 3.f / ra / ra);
    } else {
        crr = 3.f * ra / 32.f;
        cii = 1.f - 9.f * ra / 32.f;
    }
    d.xx = dr.x * dr.x * crr + cii;
    d.xy = dr.x * dr.y * crr;
    d.xz = dr.x * dr.z * crr;
    d.yy = dr.y * dr.y * crr + cii;
    d.yz = dr.y * dr.z * crr;
    d.zz = dr.z * dr.z * crr + cii;
    return d;
}

static inline f4 unit_sep(const f4 a, const f4 b) { /* b - a, normalised */
    f4 dr = b;
    dr.x -= a.x; dr.y -= a.y; dr.z -= a.z;
    dr.w = sqrtf(dr.x * dr.x + dr.y * dr.y + dr.z * dr.z);
    dr.x /= dr.w; dr.y /= dr.w; dr.z /= dr.w;
    return dr;
}

/* integrateTea_prepare, bdhitea_kernel.cu:33-65 */
static void hi_prepare(or_sim* s) {
    for (int d = 0; d < s->n; d++) {
        float g[4];
        or_rforce(&s->rng[4 * d], g);
        f4 df = {g[0] * s->var, g[1] * s->var, g[2] * s->var, g[3]};
        s->rforce[d] = df;
        f4 f = s->force[d];
        if (s->p.pulling) {
            f4 e = s->ext[d];
            if (e.w == 1.0f) { f.x = 0.0f; f.y = 0.0f; f.z = 0.0f; }
            if (e.w == 2.0f) { f.x += e.x; f.y += e.y; f.z += e.z; }
        }
        s->mforce[d] = f;
        s->force[d].x = s->force[d].y = s->force[d].z = s->force[d].w = 0.0f;
        s->hcoord[d] = s->coord[d];
    }
}

/* integrateTea_epsilon_local, bdhitea_kernel.cu:192-210 */
static inline void eps_local(const or_sim* s, const f4 c1, int j, f4* sum) {
    f4 dr = unit_sep(c1, s->coord[j]);
    f6 d = rpy(dr, s->p.hi_a);
    sum->w += d.xx + 2 * d.xy + 2 * d.xz + d.yy + 2 * d.yz + d.zz;
    sum->x += d.xx * d.xx + d.xy * d.xy + d.xz * d.xz;
    sum->y += d.xy * d.xy + d.yy * d.yy + d.yz * d.yz;
    sum->z += d.xz * d.xz + d.yz * d.yz + d.zz * d.zz;
}

/* integrateTea_epsilon[_unlisted] (bdhitea_kernel.cu:212-249) and
 * TeaUpdater::update (bdhitea.cu:133-176, FP64 on the host). */
static int tea_update(or_sim* s) {
    const int n = s->n, N = s->N;
    for (int d = 0; d < n; d++) {
        const f4 c = s->coord[d];
        f4 sum = {0.f, 0.f, 0.f, 0.f};
        f4 part;
        if (s->p.hi_unlisted) {
            int i0 = (d / N) * N;
            for (int i = i0; i < i0 + N; i++) {
                if (i == d) continue;
                part = (f4){0.f, 0.f, 0.f, 0.f};
                eps_local(s, c, i, &part);
                sum.x += part.x; sum.y += part.y; sum.z += part.z; sum.w += part.w;
            }
        } else {
            for (int k = 0; k < s->cov_cnt[d]; k++) {
                part = (f4){0.f, 0.f, 0.f, 0.f};
                eps_local(s, c, s->cov_j[k * n + d], &part);
                sum.x += part.x; sum.y += part.y; sum.z += part.z; sum.w += part.w;
            }
            for (int k = 0; k < s->nat_cnt[d]; k++) {
                part = (f4){0.f, 0.f, 0.f, 0.f};
                eps_local(s, c, s->nat_j[k * n + d], &part);
                sum.x += part.x; sum.y += part.y; sum.z += part.z; sum.w += part.w;
            }
            for (int k = 0; k < s->pair_cnt[d]; k++) {
                part = (f4){0.f, 0.f, 0.f, 0.f};
                eps_local(s, c, s->pair_j[k * n + d], &part);
                sum.x += part.x; sum.y += part.y; sum.z += part.z; sum.w += part.w;
            }
        }
        s->ci[d] = sum;
        s->eps[d] = sum.w;
    }
    for (int t = 0; t < s->ntr; t++) {
        double epsilon = 0.0;
        for (int i = 0; i < N; i++) epsilon += s->eps[t * N + i];
        if (N > 1) epsilon /= 3. * N * (3. * N - 3.);
        if (epsilon > 1.0) {
            if (s->p.hi_capricious) {
                snprintf(g_err, sizeof g_err, "HI tensor is not diagonal enough for trajectory %d: epsilon = %lf -> 1.0!", t, epsilon);
                return -1;
            }
            epsilon = 1.0;
        }
        if (epsilon > s->p.hi_epsmax) {
            snprintf(g_err, sizeof g_err, "HI tensor is not diagonal enough for trajectory %d: epsilon = %lf > %f = hi_epsmax!", t, epsilon, s->p.hi_epsmax);
            return -1;
        }
        double a = (3. * N - 1.) * epsilon * epsilon - (3. * N - 2.) * epsilon;
        if (fabs(a) < 1e-7) {
            s->beta[t] = .5f;
            if (s->p.hi_capricious && s->p.hi_a > 0.0f) {
                snprintf(g_err, sizeof g_err, "HI tensor is too diagonal for trajectory %d", t);
                return -1;
            }
        } else {
            s->beta[t] = (float)((1. - sqrt(1. - a)) / a);
        }
        s->eps_traj[t] = epsilon;
    }
    return 0;
}

/* integrateTea_force, bdhitea_kernel.cu:252-300; neighbour coordinates
 * from the step-start copy (the reference reads them through a texture of
 * the array the same launch writes, bdhitea_kernel.cu:256 vs :450). */
static inline void tea_force(const or_sim* s, const f4 c1, int j, const float* ci, float* f) {
    f4 dr = unit_sep(c1, s->hcoord[j]);
    f4 g = s->mforce[j];
    const f4 r = s->rforce[j];
    g.x += r.x * ci[0];
    g.y += r.y * ci[1];
    g.z += r.z * ci[2];
    f6 D = rpy(dr, s->p.hi_a);
    f[0] = D.xx * g.x + D.xy * g.y + D.xz * g.z;
    f[1] = D.xy * g.x + D.yy * g.y + D.yz * g.z;
    f[2] = D.xz * g.x + D.yz * g.y + D.zz * g.z;
}

static inline void hi_finish(or_sim* s, int d, f4 f, f4 coord) {
    if (s->p.pulling) {
        f4 e = s->ext[d];
        if (e.w == 1.0f) { f.x = 0.0f; f.y = 0.0f; f.z = 0.0f; }
    }
    const float mult = s->hOverZeta;
    float dx = mult * f.x, dy = mult * f.y, dz = mult * f.z;
    coord.x += dx; coord.y += dy; coord.z += dz;
    s->coord[d] = coord;
    float T = s->T[d];
    T += s->tempNorm * (dx * dx + dy * dy + dz * dz);
    s->T[d] = T;
}

/* integrateTea_kernel[_unlisted], bdhitea_kernel.cu:364-515 */
static void tea_integrate(or_sim* s) {
    const int n = s->n, N = s->N;
    f4* out = (f4*)xcalloc(n, sizeof(f4));
    for (int d = 0; d < n; d++) {
        f4 cs = s->ci[d];
        f4 f = s->mforce[d];
        f4 df = s->rforce[d];
        const float beta = s->beta[d / N];
        float ci[3];
        cs.w = beta * beta;
        ci[0] = 1.f / sqrtf(1.f + cs.w * cs.x);
        ci[1] = 1.f / sqrtf(1.f + cs.w * cs.y);
        ci[2] = 1.f / sqrtf(1.f + cs.w * cs.z);
        const f4 coord = s->hcoord[d];
        f.x += df.x * ci[0];
        f.y += df.y * ci[1];
        f.z += df.z * ci[2];
        ci[0] *= beta; ci[1] *= beta; ci[2] *= beta;
        float t[3];
        if (s->p.hi_unlisted) {
            int i0 = (d / N) * N;
            for (int i = i0; i < i0 + N; i++) {
                if (i == d) continue;
                tea_force(s, coord, i, ci, t);
                f.x += t[0]; f.y += t[1]; f.z += t[2];
            }
        } else {
            for (int k = 0; k < s->cov_cnt[d]; k++) { tea_force(s, coord, s->cov_j[k * n + d], ci, t); f.x += t[0]; f.y += t[1]; f.z += t[2]; }
            for (int k = 0; k < s->nat_cnt[d]; k++) { tea_force(s, coord, s->nat_j[k * n + d], ci, t); f.x += t[0]; f.y += t[1]; f.z += t[2]; }
            for (int k = 0; k < s->pair_cnt[d]; k++) { tea_force(s, coord, s->pair_j[k * n + d], ci, t); f.x += t[0]; f.y += t[1]; f.z += t[2]; }
        }
        out[d] = f;
    }
    for (int d = 0; d < n; d++) hi_finish(s, d, out[d], s->hcoord[d]);
    free(out);
}

/* integrateCholesky_D, bdhitea_kernel.cu:89-152: D/D_ii (identity diagonal
 * blocks) into M when `build`, and F_i + sum_j D_ij F_j from the step-start
 * forces (the reference reads them through a texture of the array it
 * writes in the same launch). */
static void chol_D(or_sim* s, int build) {
    const int n = s->n, N = s->N;
    const int m = 3 * N;
    memcpy(s->mforce0, s->mforce, sizeof(f4) * n);
    for (int d = 0; d < n; d++) {
        const int t = d / N, i0 = t * N;
        float* M = s->L + (size_t)t * m * m;
        const f4 c = s->hcoord[d];
        f4 f = s->mforce0[d];
        for (int i = i0; i < i0 + N; i++) {
            const int r = 3 * (i - i0), col = 3 * (d - i0);
            if (i == d) {
                if (build)
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) M[(size_t)(r + a) * m + col + b] = a == b ? 1.f : 0.f;
                continue;
            }
            f4 dr = unit_sep(c, s->hcoord[i]);
            f6 D = rpy(dr, s->p.hi_a);
            if (build) {
                float blk[9] = {D.xx, D.xy, D.xz, D.xy, D.yy, D.yz, D.xz, D.yz, D.zz};
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) M[(size_t)(r + a) * m + col + b] = blk[3 * a + b];
            }
            const f4 g = s->mforce0[i];
            f.x += D.xx * g.x + D.xy * g.y + D.xz * g.z;
            f.y += D.xy * g.x + D.yy * g.y + D.yz * g.z;
            f.z += D.xz * g.x + D.yz * g.y + D.zz * g.z;
        }
        s->mforce[d] = f;
    }
}

/* integrateCholesky_decompose, bdhitea_kernel.cu:154-190: column-oriented
 * Cholesky, L(t,j) = (A(j,t) - sum_k L(t,k) L(j,k)) / L(j,j), k ascending;
 * the result is kept in the lower triangle (the reference mirrors it). */
static void chol_decompose(or_sim* s) {
    const int m = 3 * s->N;
    float* row = (float*)xcalloc(m, sizeof(float));
    for (int t = 0; t < s->ntr; t++) {
        float* M = s->L + (size_t)t * m * m;
        for (int j = 0; j < m; j++) {
            for (int c = 0; c < m; c++) row[c] = c <= j ? M[(size_t)j * m + c] : M[(size_t)c * m + j];
            float diag = 0.f;
            for (int tid = j; tid < m; tid++) {
                float val = tid == j ? row[j] : M[(size_t)tid * m + j];
                for (int k = 0; k < j; k++) val -= M[(size_t)tid * m + k] * row[k];
                if (tid == j) diag = sqrtf(val);
                row[tid] = val; /* unscaled, consumed below */
            }
            for (int tid = j; tid < m; tid++) M[(size_t)tid * m + j] = row[tid] / diag;
            for (int c = j + 1; c < m; c++) M[(size_t)j * m + c] = 0.f;
        }
    }
    free(row);
}

/* integrateCholesky_L, bdhitea_kernel.cu:303-361 */
static void chol_L(or_sim* s) {
    const int n = s->n, N = s->N, m = 3 * N;
    f4* out = (f4*)xcalloc(n, sizeof(f4));
    for (int d = 0; d < n; d++) {
        const int t = d / N, i0 = t * N, r = 3 * (d - i0);
        const float* M = s->L + (size_t)t * m * m;
        f4 f = s->mforce[d];
        for (int i = i0; i <= d; i++) {
            const f4 g = s->rforce[i];
            const int c = 3 * (i - i0);
            if (s->p.hi_faithful_blocks) {
                /* the six entries the reference reads, applied as a
                 * symmetric block (bdhitea_kernel.cu:322-337) */
                float xx = M[(size_t)r * m + c], xy = M[(size_t)(r + 1) * m + c], xz = M[(size_t)(r + 2) * m + c];
                float yy = M[(size_t)(r + 1) * m + c + 1], yz = M[(size_t)(r + 2) * m + c + 1];
                float zz = M[(size_t)(r + 2) * m + c + 2];
                f.x += xx * g.x + xy * g.y + xz * g.z;
                f.y += xy * g.x + yy * g.y + yz * g.z;
                f.z += xz * g.x + yz * g.y + zz * g.z;
            } else {
                const float* B = M + (size_t)r * m + c;
                f.x += B[0] * g.x + B[1] * g.y + B[2] * g.z;
                f.y += B[m] * g.x + B[m + 1] * g.y + B[m + 2] * g.z;
                f.z += B[2 * m] * g.x + B[2 * m + 1] * g.y + B[2 * m + 2] * g.z;
            }
        }
        out[d] = f;
    }
    for (int d = 0; d < n; d++) hi_finish(s, d, out[d], s->hcoord[d]);
    free(out);
}

/* TeaIntegrator::integrate, bdhitea.cu:76-109 */
static void hi_integrate(or_sim* s) {
    hi_prepare(s);
    if (!s->p.hi_exact) {
        tea_integrate(s);
    } else {
        const int refactor = (s->step % (s->p.hi_chol_freq > 0 ? s->p.hi_chol_freq : 1)) == 0;
        chol_D(s, refactor);
        if (refactor) chol_decompose(s);
        chol_L(s);
    }
}

void or_hi_tea(const or_sim* s, double* eps, float* beta) {
    for (int t = 0; t < s->ntr; t++) {
        eps[t] = s->eps_traj ? s->eps_traj[t] : 0.0;
        beta[t] = s->beta ? s->beta[t] : 0.0f;
    }
}

int or_hi_factor(const or_sim* s, float* L) {
    if (!s->L) return -1;
    memcpy(L, s->L, sizeof(float) * 9 * (size_t)s->N * s->N);
    return 0;
}

void or_rpy_matrix(int n, const float* c4, double a, double* D) {
    const int m = 3 * n;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double blk[9];
            if (i == j) {
                for (int k = 0; k < 9; k++) blk[k] = (k % 4 == 0) ? 1.0 : 0.0;
            } else {
                double x = (double)c4[4 * j] - c4[4 * i], y = (double)c4[4 * j + 1] - c4[4 * i + 1],
                       z = (double)c4[4 * j + 2] - c4[4 * i + 2];
                double r = sqrt(x * x + y * y + z * z), ra = r / a, crr, cii;
                x /= r; y /= r; z /= r;
                if (ra > 2.0) {
                    crr = 0.75 / ra * (1.0 - 2.0 / (ra * ra));
                    cii = 0.75 / ra * (1.0 + 2.0 / (3.0 * ra * ra));
                } else {
                    crr = 3.0 * ra / 32.0;
                    cii = 1.0 - 9.0 * ra / 32.0;
                }
                double u[3] = {x, y, z};
                for (int p = 0; p < 3; p++)
                    for (int q = 0; q < 3; q++) blk[3 * p + q] = u[p] * u[q] * crr + (p == q ? cii : 0.0);
            }
            for (int p = 0; p < 3; p++)
                for (int q = 0; q < 3; q++) D[(size_t)(3 * i + p) * m + 3 * j + q] = blk[3 * p + q];
        }
}

/* ------------------------------------------------------------------ */
/* updaters                                                             */
/* ------------------------------------------------------------------ */

static double* push_row(double** buf, int* nrow, int* cap, int width) {
    if (*nrow == *cap) {
        *cap = *cap ? 2 * *cap : 64;
        *buf = (double*)realloc(*buf, sizeof(double) * width * *cap);
    }
    return &(*buf)[(size_t)width * (*nrow)++];
}

/* PullingUpdater::update + updateForces + computeForce,
 * pulling.cu:152-165,198-214.  The force uses the cantilever position of
 * the previous update; the new position is set afterwards. */
static void pulling_update(or_sim* s) {
    const int N = s->N;
    float xt = s->p.deltax * (float)(s->step / s->p.pullFreq);
    for (int t = 0; t < s->ntr; t++) {
        f4 cn = s->coord[N * t + s->p.pulledEnd];
        float* const F = s->extForce[t];
        float* const cant = s->cant[t];
        const float* const pv = s->pullVec[t];
        F[0] = s->p.k_trans * (cant[0] - cn.x);
        F[1] = s->p.k_trans * (cant[1] - cn.y);
        F[2] = s->p.k_trans * (cant[2] - cn.z);
        cant[0] = s->cant0[t][0] + xt * pv[0];
        cant[1] = s->cant0[t][1] + xt * pv[1];
        cant[2] = s->cant0[t][2] + xt * pv[2];
        for (int k = 0; k < s->npulled; k++) {
            f4* e = &s->ext[t * N + s->pulled[k]];
            e->x = F[0]; e->y = F[1]; e->z = F[2]; e->w = 2.0f;
        }
        /* savePullingData (pulling.cu:166-187): the row of the pass's
         * coordinates and force, every pullFreq steps */
        if (s->step % s->p.pullFreq == 0) {
            const f4 a = s->coord[N * t + s->p.pulledEnd], b = s->coord[N * t + s->p.fixedEnd];
            const float vx = a.x - b.x, vy = a.y - b.y, vz = a.z - b.z;
            double* r = push_row(&s->prow, &s->nprow, &s->capprow, 8);
            r[0] = t;
            r[1] = (double)s->step;
            r[2] = sqrtf(vx * vx + vy * vy + vz * vz);
            r[3] = vx * pv[0] + vy * pv[1] + vz * pv[2];
            r[4] = F[0] * pv[0] + F[1] * pv[1] + F[2] * pv[2];
            r[5] = F[0]; r[6] = F[1]; r[7] = F[2];
        }
    }
}

/* IndentationTipUpdater::update, indentation.cu:414-511 */
static void tip_update(or_sim* s) {
    const or_params* p = &s->p;
    if (s->step == p->indRetractionStep) s->V = -s->V;
    if (s->step % p->nav != 0) return;
    float f[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int i = 0; i < s->n; i++) {
        f[0] += s->tipF[i].x; f[1] += s->tipF[i].y; f[2] += s->tipF[i].z;
        s->tipF[i].x = 0.0f; s->tipF[i].y = 0.0f; s->tipF[i].z = 0.0f;
    }
    f[0] /= (float)p->nav; f[1] /= (float)p->nav; f[2] /= (float)p->nav;
    f[3] = sqrtf(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    s->dx = (float)(s->step / p->nav) * s->V;
    if (p->indMoveSurface) {
        for (int d = 0; d < 3; d++) s->micaR[d] = s->micaR0[d] + s->dir[d] * s->dx;
    } else {
        for (int d = 0; d < 3; d++) s->chip[d] = s->chip0[d] + s->dir[d] * s->dx;
    }
    float dr[3] = {s->chip[0] - s->tip[0], s->chip[1] - s->tip[1], s->chip[2] - s->tip[2]};
    const float drw = sqrtf(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
    float df[3];
    for (int d = 0; d < 3; d++) df[d] = -f[d] + dr[d] * p->indKs;
    float mult = ((p->timestep * ((float)p->nav)) / (p->indTipZeta));
    if (p->indFixTrans) {
        float ct = mult * df[0] * s->dir[0] + mult * df[1] * s->dir[1] + mult * df[2] * s->dir[2];
        for (int d = 0; d < 3; d++) s->tip[d] += ct * s->dir[d];
    } else {
        for (int d = 0; d < 3; d++) s->tip[d] += mult * df[d];
    }
    /* averages and the output row (indentation.cu:497-536), every
     * indentationOutputFreq steps */
    for (int d = 0; d < 3; d++) {
        s->chipAv[d] += s->chip[d];
        s->tipAv[d] += s->tip[d];
        s->fav[d] += f[d];
    }
    s->fav[3] += f[3];
    s->kDeltaXAv += drw * p->indKs;
    if (p->indOutputFreq > 0 && s->step % p->indOutputFreq == 0) {
        const int factor = p->indOutputFreq / p->nav;
        for (int d = 0; d < 3; d++) {
            s->chipAv[d] /= factor;
            s->tipAv[d] /= factor;
            s->fav[d] /= factor;
        }
        s->fav[3] /= factor;
        s->kDeltaXAv /= factor;
        const float fmod = sqrtf(s->fav[0] * s->fav[0] + s->fav[1] * s->fav[1] + s->fav[2] * s->fav[2]);
        double* r = push_row(&s->irow, &s->nirow, &s->capirow, 15);
        r[0] = (double)s->step; r[1] = s->dx; r[2] = s->fav[3]; r[3] = fmod; r[4] = fabsf(drw * p->indKs);
        r[5] = s->kDeltaXAv; r[6] = s->fav[0]; r[7] = s->fav[1]; r[8] = s->fav[2];
        r[9] = s->tip[0]; r[10] = s->tip[1]; r[11] = s->tip[2];
        r[12] = s->chip[0]; r[13] = s->chip[1]; r[14] = s->chip[2];
        for (int d = 0; d < 3; d++) s->chipAv[d] = s->tipAv[d] = s->fav[d] = 0.0f;
        s->fav[3] = s->kDeltaXAv = 0.0f;
    }
}

/* TemperatureUpdater::update + setTemperature, langevin.cu:49-82 */
static void temperature_update(or_sim* s) {
    if (s->step % s->p.tempFreq == 0) {
        float temp = s->p.initialT + (float)(s->step / s->p.tempFreq) * s->p.deltaT;
        s->var = sqrtf(2.0f * temp * s->p.zeta / s->p.timestep);
    }
}

/* OutputManager::update (energies, T, native count), output_manager.cu:73-183 */
static void output_update(or_sim* s) {
    if (s->step % s->p.outputtiming == 0) {
        float e[3 * 64];
        float* ee = s->ntr <= 64 ? e : (float*)xcalloc((size_t)3 * s->ntr, sizeof(float));
        or_sum_energies(s, ee);
        if (s->ndat == s->capdat) {
            s->capdat = s->capdat ? 2 * s->capdat : 64;
            s->dat = (double*)realloc(s->dat, sizeof(double) * 6 * s->capdat);
        }
        double tempav = 0.0;
        for (int i = 0; i < s->N; i++) tempav += s->T[i];
        tempav /= ((double)s->N * s->p.nav);
        int q = 0;
        for (int k = 0; k < s->nn; k++) {
            f4 a = s->coord[s->ni[k]], b = s->coord[s->nj[k]];
            float rx = a.x - b.x, ry = a.y - b.y, rz = a.z - b.z;
            float dr = sqrtf(rx * rx + ry * ry + rz * rz);
            if (dr < s->p.R_limit_bond || fabsf(s->nr0[k] - dr) <= s->p.R_limit) q++;
        }
        double* row = &s->dat[6 * s->ndat++];
        row[0] = (double)s->step; row[1] = (float)tempav;
        row[2] = ee[0]; row[3] = ee[s->ntr]; row[4] = ee[2 * s->ntr]; row[5] = q;
        if (ee != e) free(ee);
    }
    if (s->step % s->p.nav == 0) memset(s->T, 0, sizeof(float) * s->n);
}

int or_get_pull_rows(const or_sim* s, double* out, int cap) {
    int m = s->nprow < cap ? s->nprow : cap;
    if (m) memcpy(out, s->prow, sizeof(double) * 8 * m);
    return s->nprow;
}

int or_get_ind_rows(const or_sim* s, double* out, int cap) {
    int m = s->nirow < cap ? s->nirow : cap;
    if (m) memcpy(out, s->irow, sizeof(double) * 15 * m);
    return s->nirow;
}

int or_get_dat(const or_sim* s, double* out, int cap) {
    int m = s->ndat < cap ? s->ndat : cap;
    memcpy(out, s->dat, sizeof(double) * 6 * m);
    return s->ndat;
}

/* runGPU, gsop.cu:111-163, in initFF's creation order (gsop.cu:80-101) */
long long or_run(or_sim* s, long long numsteps) {
    const or_params* p = &s->p;
    long long done = 0;
    int stride = s->stride;
    if (s->step % p->pairs_freq == 0) or_build_verlet(s); /* gsop.cu:126 */
    for (long long b = s->step / stride; b <= numsteps / stride; b++) {
        if (p->indentation) tip_update(s);
        if (p->pulling && s->pullingUpdater) pulling_update(s);
        if (s->step % p->possiblepairs_freq == 0) or_build_possible(s);
        output_update(s);
        if (s->step % p->pairs_freq == 0) or_build_verlet(s);
        if (p->heating) temperature_update(s);
        if (p->hi_on && !p->hi_exact && s->step % p->hi_epsilon_freq == 0)
            if (tea_update(s)) return -1;
        for (int j = 0; j < stride; j++) {
            forces_pass(s);
            if (p->hi_on) {
                hi_integrate(s);
                s->step += 1;
            } else {
                integrate_pass(s);
            }
        }
        if (p->hi_on) s->step -= stride;
        s->step += stride;
        done += stride;
    }
    return done;
}

/* FP64 twin for zero-temperature trajectories: same cadence, same list
 * logic, double coordinates and double force arithmetic. */
long long or_run_f64(or_sim* s, long long numsteps, double* cout) {
    const or_params* p = &s->p;
    const int n = s->n;
    if (!(p->minimization || p->temperature == 0.0f) || p->indentation || p->heating) {
        snprintf(g_err, sizeof g_err, "or_run_f64 models T=0 runs without indentation or heating");
        return -1;
    }
    double* x = (double*)xcalloc((size_t)3 * n, sizeof(double));
    double* fr = (double*)xcalloc((size_t)3 * n, sizeof(double));
    for (int i = 0; i < n; i++) { x[3 * i] = s->coord[i].x; x[3 * i + 1] = s->coord[i].y; x[3 * i + 2] = s->coord[i].z; }
    double hz = (double)p->timestep / p->zeta;
    int stride = s->stride;
    long long done = 0;
    /* the lists are built from float images of the double coordinates */
#define SYNC_F32() do { for (int i_ = 0; i_ < n; i_++) { s->coord[i_].x = (float)x[3*i_]; s->coord[i_].y = (float)x[3*i_+1]; s->coord[i_].z = (float)x[3*i_+2]; } } while (0)
    double cant[3] = {0.0, 0.0, 0.0};
    if (s->cant) for (int d = 0; d < 3; d++) cant[d] = s->cant[0][d];
    double ext[3] = {0, 0, 0};
    int have_ext = 0;
    if (s->step % p->pairs_freq == 0) or_build_verlet(s);
    for (long long b = s->step / stride; b <= numsteps / stride; b++) {
        if (p->pulling && s->pullingUpdater) {
            int e = p->pulledEnd;
            ext[0] = p->k_trans * (cant[0] - x[3 * e]);
            ext[1] = p->k_trans * (cant[1] - x[3 * e + 1]);
            ext[2] = p->k_trans * (cant[2] - x[3 * e + 2]);
            double xt = (double)p->deltax * (double)(s->step / p->pullFreq);
            for (int d = 0; d < 3; d++) cant[d] = s->cant0[0][d] + xt * s->pullVec[0][d];
            have_ext = 1;
        }
        SYNC_F32();
        if (s->step % p->possiblepairs_freq == 0) or_build_possible(s);
        if (s->step % p->pairs_freq == 0) or_build_verlet(s);
        for (int j = 0; j < stride; j++) {
            /* forces in double with the list membership of the float run */
            for (int i = 0; i < n; i++) { fr[3*i] = fr[3*i+1] = fr[3*i+2] = 0.0; }
            for (int i = 0; i < n; i++) {
                double fx = 0, fy = 0, fz = 0;
                for (int k = 0; k < s->cov_cnt[i]; k++) {
                    int o = s->cov_j[k * n + i];
                    double dx = x[3*o] - x[3*i], dy = x[3*o+1] - x[3*i+1], dz = x[3*o+2] - x[3*i+2];
                    double r = sqrt(dx*dx + dy*dy + dz*dz), dr = r - s->cov_r0[k * n + i];
                    double R2 = (double)p->R_limit * p->R_limit;
                    double w = p->kspring_cov * dr / (r * (1.0 - dr*dr / R2));
                    fx += w*dx; fy += w*dy; fz += w*dz;
                }
                for (int k = 0; k < s->nat_cnt[i]; k++) {
                    int o = s->nat_j[k * n + i];
                    double dx = x[3*o] - x[3*i], dy = x[3*o+1] - x[3*i+1], dz = x[3*o+2] - x[3*i+2];
                    double q = (double)s->nat_r02[k * n + i] / (dx*dx + dy*dy + dz*dz), p6 = q*q*q;
                    double w = q * (p6*p6 - p6) * (double)s->nat_m12[k * n + i];
                    fx += w*dx; fy += w*dy; fz += w*dz;
                }
                for (int k = 0; k < s->pair_cnt[i]; k++) {
                    int o = s->pair_j[k * n + i];
                    double dx = x[3*o] - x[3*i], dy = x[3*o+1] - x[3*i+1], dz = x[3*o+2] - x[3*i+2];
                    double r2 = dx*dx + dy*dy + dz*dz, rc = p->pairs_cutoff;
                    if (r2 < rc * rc) {
                        double a2 = (double)p->a * p->a, q = a2 / r2;
                        double w = -6.0 * p->el / a2 * q*q*q*q;
                        fx += w*dx; fy += w*dy; fz += w*dz;
                    }
                }
                fr[3*i] = fx; fr[3*i+1] = fy; fr[3*i+2] = fz;
            }
            for (int i = 0; i < n; i++) {
                if (p->pulling) {
                    f4 e = s->ext[i];
                    if (e.w == 1.0f) { fr[3*i] = fr[3*i+1] = fr[3*i+2] = 0.0; }
                    if (e.w == 2.0f) {
                        if (have_ext) { fr[3*i] += ext[0]; fr[3*i+1] += ext[1]; fr[3*i+2] += ext[2]; }
                        else { fr[3*i] += e.x; fr[3*i+1] += e.y; fr[3*i+2] += e.z; }
                    }
                }
                x[3*i] += hz * fr[3*i]; x[3*i+1] += hz * fr[3*i+1]; x[3*i+2] += hz * fr[3*i+2];
            }
        }
        s->step += stride;
        done += stride;
    }
    SYNC_F32();
#undef SYNC_F32
    memcpy(cout, x, sizeof(double) * 3 * n);
    free(x); free(fr);
    return done;
}

/* CPU baseline: the reference's per-kernel loops (forces then Langevin)
 * with list rebuilds at their cadences, OpenMP over beads. */
double or_bench_steps(or_sim* s, int steps, int nthreads) {
    s->nthreads = nthreads < 1 ? 1 : nthreads;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int k = 0; k < steps; k++) {
        if (s->step % s->p.possiblepairs_freq == 0) or_build_possible(s);
        if (s->step % s->p.pairs_freq == 0) or_build_verlet(s);
        forces_pass(s);
        integrate_pass(s);
        s->step++;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    s->nthreads = 1;
    return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}
