// hi.hip -- hydrodynamic interactions: the Rotne-Prager-Yamakawa tensor with
// the TEA (truncated expansion) integrator and the exact Cholesky path.
//
// Reference: src/Integrators/bdhitea_kernel.cu (integrateTea_prepare :33-65,
// integrateTea_RPY :67-87, integrateCholesky_D :89-152,
// integrateCholesky_decompose :154-190, integrateTea_epsilon :212-249,
// integrateTea_force :252-300, integrateCholesky_L :303-361,
// integrateTea_kernel[_unlisted] :364-515) and src/Integrators/bdhitea.cu
// (TeaUpdater::update :133-176).
//
// The per-step preparation (noise into rforce, forces into mforce) is done by
// k_step for HI jobs; these kernels do the rest.  Neighbour coordinates come
// from the step-start buffer, so no kernel reads what it writes.
//
// The exact path factorises the RPY matrix (3N x 3N, padded to a multiple
// of 64 with an identity block) by a right-looking blocked Cholesky:
//   diagonal 64x64 blocks in FP64 in LDS, with their inverse;
//   the panel below as A L^-T with FP64 accumulation, also emitted as
//   split-BF16 pairs (hi = bf16(x), lo = bf16(x - hi));
//   the trailing update C -= P_I P_J^T on MFMA, either split-BF16 with three
//   products (hi*hi + hi*lo + lo*hi, FP32 accumulate) or FP32 MFMA.
#include <hip/hip_runtime.h>

#include <cmath>

#include "devmath.h"
#include "hi.h"
#include "jobdesc.h"

namespace sop {

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct F6 {
    float xx, xy, xz, yy, yz, zz;
};

// D_ij / D_ii for a normalised separation with its length in w.
__device__ __forceinline__ F6 rpy(const float4 dr, float a) {
    F6 d;
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

// b - a, normalised, length in w
__device__ __forceinline__ float4 unit_sep(const float4 a, const float4 b) {
    float4 dr = b;
    dr.x -= a.x;
    dr.y -= a.y;
    dr.z -= a.z;
    dr.w = sqrtf(dr.x * dr.x + dr.y * dr.y + dr.z * dr.z);
    dr.x /= dr.w;
    dr.y /= dr.w;
    dr.z /= dr.w;
    return dr;
}

__device__ __forceinline__ int wave_of() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// tile loop shared with kernels.hip's layout (64-bead tiles, 4 per block)
#define FOR_TILES(ntiles)                                                 \
    const int lane = threadIdx.x & 63;                                    \
    for (int t = blockIdx.x * 4 + wave_of(); t < (ntiles); t += gridDim.x * 4)

__device__ __forceinline__ void hi_finish(const DevJob& J, int d, long long s, float fx, float fy, float fz,
                                          float4 coord) {
    if (J.flags & JF_PULLING) {
        const float4 e = J.ext[d];
        if (e.w == 1.0f) {
            fx = 0.0f; fy = 0.0f; fz = 0.0f;
        }
    }
    const float m = J.hOverZeta;
    const float dx = m * fx, dy = m * fy, dz = m * fz;
    coord.x += dx;
    coord.y += dy;
    coord.z += dz;
    J.coord[(s + 1) & 1][d] = coord;
    const bool reset = (s % J.stride == 0) && (s % J.nav == 0);
    const float T0 = reset ? 0.0f : J.T[d];
    J.T[d] = T0 + J.tempNorm * (dx * dx + dy * dy + dz * dz);
}

// ---------------------------------------------------------------------------
// TEA
// ---------------------------------------------------------------------------
__device__ __forceinline__ void eps_add(const float4 c1, const float4 c2, float a, float4& sum) {
    const float4 dr = unit_sep(c1, c2);
    const F6 d = rpy(dr, a);
    float4 p;
    p.w = d.xx + 2 * d.xy + 2 * d.xz + d.yy + 2 * d.yz + d.zz;
    p.x = d.xx * d.xx + d.xy * d.xy + d.xz * d.xz;
    p.y = d.xy * d.xy + d.yy * d.yy + d.yz * d.yz;
    p.z = d.xz * d.xz + d.yz * d.yz + d.zz * d.zz;
    sum.x += p.x;
    sum.y += p.y;
    sum.z += p.z;
    sum.w += p.w;
}

__global__ __launch_bounds__(256) void k_hi_eps(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                                const int2* __restrict__ tiles, int ntiles) {
    FOR_TILES(ntiles) {
        const int2 td = tiles[t];
        const DevJob& J = jobs[td.x];
        if (!(J.flags & JF_HI) || (J.flags & JF_HI_EXACT)) continue;
        const long long s = st[td.x].step;
        if (!(s < J.end_step && s % J.stride == 0 && s % J.hi_eps_freq == 0)) continue;
        const int d = td.y + lane;
        const int n = J.n;
        if (d >= n) continue;
        const float4* __restrict__ cur = J.coord[s & 1];
        const float4 c = cur[d];
        const float a = J.hi_a;
        float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
        if (!(J.flags & JF_HI_LISTED)) {
            for (int i = 0; i < n; ++i)
                if (i != d) eps_add(c, cur[i], a, sum);
        } else {
            for (int k = 0; k < J.cov_cnt[d]; ++k) eps_add(c, cur[J.cov_j[(size_t)k * n + d]], a, sum);
            for (int k = 0; k < J.nat_cnt[d]; ++k) eps_add(c, cur[J.nat_j[(size_t)k * n + d]], a, sum);
            for (int k = 0; k < J.pair_cnt[d]; ++k) eps_add(c, cur[list_at(J.pair_j, J.idx16, (size_t)k * n + d)], a, sum);
        }
        J.ci[d] = sum;
        J.eps[d] = sum.w;
    }
}

// TeaUpdater::update (bdhitea.cu:133-176): epsilon and beta' in FP64.
__global__ void k_hi_beta(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st, int njobs) {
    const int jb = blockIdx.x * blockDim.x + threadIdx.x;
    if (jb >= njobs) return;
    const DevJob& J = jobs[jb];
    if (!(J.flags & JF_HI) || (J.flags & JF_HI_EXACT)) return;
    DevJobState& S = st[jb];
    const long long s = S.step;
    if (!(s < J.end_step && s % J.stride == 0 && s % J.hi_eps_freq == 0)) return;
    const int N = J.n;
    double epsilon = 0.0;
    for (int i = 0; i < N; ++i) epsilon += J.eps[i];
    if (N > 1) epsilon /= 3. * N * (3. * N - 3.);
    if (epsilon > 1.0) {
        if (J.hi_capricious) S.overflow |= 4;
        epsilon = 1.0;
    }
    if (epsilon > J.hi_epsmax) S.overflow |= 4;
    const double a = (3. * N - 1.) * epsilon * epsilon - (3. * N - 2.) * epsilon;
    if (fabs(a) < 1e-7) {
        S.beta = .5f;
        if (J.hi_capricious && J.hi_a > 0.0f) S.overflow |= 8;
    } else {
        S.beta = (float)((1. - sqrt(1. - a)) / a);
    }
    S.eps_traj = epsilon;
}

__device__ __forceinline__ void tea_force(const float4 c1, const float4 c2, const float4 mf, const float4 rf,
                                          const float ci0, const float ci1, const float ci2, float a, float& fx,
                                          float& fy, float& fz) {
    const float4 dr = unit_sep(c1, c2);
    float gx = mf.x, gy = mf.y, gz = mf.z;
    gx += rf.x * ci0;
    gy += rf.y * ci1;
    gz += rf.z * ci2;
    const F6 D = rpy(dr, a);
    fx += D.xx * gx + D.xy * gy + D.xz * gz;
    fy += D.xy * gx + D.yy * gy + D.yz * gz;
    fz += D.xz * gx + D.yz * gy + D.zz * gz;
}

__global__ __launch_bounds__(256) void k_hi_tea(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                                const int2* __restrict__ tiles, int ntiles, int kstep) {
    FOR_TILES(ntiles) {
        const int2 td = tiles[t];
        const DevJob& J = jobs[td.x];
        if (!(J.flags & JF_HI) || (J.flags & JF_HI_EXACT)) continue;
        const DevJobState& S = st[td.x];
        const long long s = S.step + kstep;
        if (s >= J.end_step) continue;
        const int d = td.y + lane;
        const int n = J.n;
        if (d >= n) continue;
        const float4* __restrict__ cur = J.coord[s & 1];
        float4 cs = J.ci[d];
        const float4 f0 = J.mforce[d];
        const float4 df = J.rforce[d];
        const float beta = S.beta;
        cs.w = beta * beta;
        float ci0 = 1.f / sqrtf(1.f + cs.w * cs.x);
        float ci1 = 1.f / sqrtf(1.f + cs.w * cs.y);
        float ci2 = 1.f / sqrtf(1.f + cs.w * cs.z);
        const float4 coord = cur[d];
        float fx = f0.x, fy = f0.y, fz = f0.z;
        fx += df.x * ci0;
        fy += df.y * ci1;
        fz += df.z * ci2;
        ci0 *= beta;
        ci1 *= beta;
        ci2 *= beta;
        const float a = J.hi_a;
        float tx, ty, tz;
        if (!(J.flags & JF_HI_LISTED)) {
            for (int i = 0; i < n; ++i) {
                if (i == d) continue;
                tx = 0.f; ty = 0.f; tz = 0.f;
                tea_force(coord, cur[i], J.mforce[i], J.rforce[i], ci0, ci1, ci2, a, tx, ty, tz);
                fx += tx; fy += ty; fz += tz;
            }
        } else {
            for (int k = 0; k < J.cov_cnt[d]; ++k) {
                const int i = J.cov_j[(size_t)k * n + d];
                tx = 0.f; ty = 0.f; tz = 0.f;
                tea_force(coord, cur[i], J.mforce[i], J.rforce[i], ci0, ci1, ci2, a, tx, ty, tz);
                fx += tx; fy += ty; fz += tz;
            }
            for (int k = 0; k < J.nat_cnt[d]; ++k) {
                const int i = J.nat_j[(size_t)k * n + d];
                tx = 0.f; ty = 0.f; tz = 0.f;
                tea_force(coord, cur[i], J.mforce[i], J.rforce[i], ci0, ci1, ci2, a, tx, ty, tz);
                fx += tx; fy += ty; fz += tz;
            }
            for (int k = 0; k < J.pair_cnt[d]; ++k) {
                const int i = list_at(J.pair_j, J.idx16, (size_t)k * n + d);
                tx = 0.f; ty = 0.f; tz = 0.f;
                tea_force(coord, cur[i], J.mforce[i], J.rforce[i], ci0, ci1, ci2, a, tx, ty, tz);
                fx += tx; fy += ty; fz += tz;
            }
        }
        hi_finish(J, d, s, fx, fy, fz, coord);
    }
}

// ---------------------------------------------------------------------------
// exact path
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool refactor_due(const DevJob& J, long long s) {
    return s < J.end_step && (s % J.hi_chol_freq) == 0;
}

// D/D_ii lower block triangle into M (refactorisation steps) and
// F_i + sum_j D_ij F_j into fdf.
__global__ __launch_bounds__(256) void k_chol_D(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                                const int2* __restrict__ tiles, int ntiles, int kstep) {
    FOR_TILES(ntiles) {
        const int2 td = tiles[t];
        const DevJob& J = jobs[td.x];
        if (!(J.flags & JF_HI_EXACT)) continue;
        const long long s = st[td.x].step + kstep;
        if (s >= J.end_step) continue;
        const int d = td.y + lane;
        const int n = J.n;
        if (d >= n) continue;
        const bool build = refactor_due(J, s);
        const float4* __restrict__ cur = J.coord[s & 1];
        const size_t ld = J.chol_ld;
        float* __restrict__ M = J.chol_M;
        const float4 c = cur[d];
        const float4 f0 = J.mforce[d];
        float fx = f0.x, fy = f0.y, fz = f0.z;
        const float a = J.hi_a;
        for (int i = 0; i < n; ++i) {
            const size_t r = 3 * (size_t)i, col = 3 * (size_t)d;
            if (i == d) {
                if (build)
                    for (int p = 0; p < 3; ++p)
                        for (int q = 0; q < 3; ++q) M[(r + p) * ld + col + q] = p == q ? 1.f : 0.f;
                continue;
            }
            const float4 dr = unit_sep(c, cur[i]);
            const F6 D = rpy(dr, a);
            if (build && i > d) {
                float* m0 = M + r * ld + col;
                m0[0] = D.xx; m0[1] = D.xy; m0[2] = D.xz;
                m0 += ld;
                m0[0] = D.xy; m0[1] = D.yy; m0[2] = D.yz;
                m0 += ld;
                m0[0] = D.xz; m0[1] = D.yz; m0[2] = D.zz;
            }
            const float4 g = J.mforce[i];
            fx += D.xx * g.x + D.xy * g.y + D.xz * g.z;
            fy += D.xy * g.x + D.yy * g.y + D.yz * g.z;
            fz += D.xz * g.x + D.yz * g.y + D.zz * g.z;
        }
        J.fdf[d] = make_float4(fx, fy, fz, 0.0f);
    }
}

// Noise sum_j B_ij rforce_j for i's three rows, one wavefront per bead, then
// the Euler step.  faithful: each 3x3 block applied as the symmetric matrix
// of its lower triangle, as integrateCholesky_L reads it
// (bdhitea_kernel.cu:322-337); otherwise the factor itself.
__global__ __launch_bounds__(256) void k_chol_L(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                                const int2* __restrict__ tiles, int ntiles, int kstep) {
    const int lane = threadIdx.x & 63;
    // one wavefront per bead: bead index = tile*64 + r for r in 0..63
    for (int t = blockIdx.x * 4 + wave_of(); t < ntiles; t += gridDim.x * 4) {
        const int2 td = tiles[t];
        const DevJob& J = jobs[td.x];
        if (!(J.flags & JF_HI_EXACT)) continue;
        const long long s = st[td.x].step + kstep;
        if (s >= J.end_step) continue;
        const int n = J.n;
        const size_t ld = J.chol_ld;
        const float* __restrict__ M = J.chol_M;
        const float* __restrict__ g = reinterpret_cast<const float*>(J.rforce);
        const bool faithful = (J.flags & JF_HI_FAITHFUL) != 0;
        const int dend = min(td.y + 64, n);
        for (int d = td.y; d < dend; ++d) {
            const int r = 3 * d;
            float a0 = 0.f, a1 = 0.f, a2 = 0.f;
            for (int c = lane; c < r + 3; c += 64) {
                const int blk = c / 3, b = c - 3 * blk;
                const float gc = g[4 * blk + b];
                float b0, b1, b2;
                if (faithful) {
                    const int c0 = 3 * blk;
                    // B[a][c0+b] = L[r+max(a,b)][c0+min(a,b)]
                    b0 = M[(size_t)(r + b) * ld + c0];
                    b1 = M[(size_t)(r + (b > 1 ? b : 1)) * ld + c0 + (b < 1 ? b : 1)];
                    b2 = M[(size_t)(r + 2) * ld + c0 + b];
                    if (blk == d) {  // inside the diagonal block only the lower part exists
                        b0 = M[(size_t)(r + b) * ld + c0];
                    }
                } else {
                    b0 = c <= r ? M[(size_t)r * ld + c] : 0.f;
                    b1 = c <= r + 1 ? M[(size_t)(r + 1) * ld + c] : 0.f;
                    b2 = M[(size_t)(r + 2) * ld + c];
                }
                a0 += b0 * gc;
                a1 += b1 * gc;
                a2 += b2 * gc;
            }
            // fixed-order tree reduction across the wavefront
            for (int off = 32; off > 0; off >>= 1) {
                a0 += __shfl_xor(a0, off);
                a1 += __shfl_xor(a1, off);
                a2 += __shfl_xor(a2, off);
            }
            if (lane == 0) {
                const float4 f = J.fdf[d];
                hi_finish(J, d, s, f.x + a0, f.y + a1, f.z + a2, J.coord[s & 1][d]);
            }
        }
    }
}

// ---- blocked Cholesky -------------------------------------------------------
static constexpr int NB = 64;

__device__ __forceinline__ unsigned short f2bf(float x) {  // round to nearest even
    unsigned int u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

// Factor the k-th diagonal block in FP64 and store its inverse.
__global__ __launch_bounds__(256) void k_potrf_diag(float* __restrict__ M, size_t ld, int k, double* __restrict__ linv,
                                                    const int* __restrict__ flag) {
    if (flag && !*flag) return;
    __shared__ double s[NB][NB + 1];
    __shared__ double v[NB][NB + 1];
    const int tid = threadIdx.x;
    const size_t o = (size_t)k * NB;
    for (int e = tid; e < NB * NB; e += 256) {
        const int r = e / NB, c = e % NB;
        s[r][c] = c <= r ? (double)M[(o + r) * ld + o + c] : 0.0;
    }
    __syncthreads();
    for (int j = 0; j < NB; ++j) {
        if (tid == 0) s[j][j] = sqrt(s[j][j]);
        __syncthreads();
        const double djj = s[j][j];
        for (int r = j + 1 + tid; r < NB; r += 256) s[r][j] /= djj;
        __syncthreads();
        const int rem = NB - 1 - j;
        for (int e = tid; e < rem * rem; e += 256) {
            const int r = j + 1 + e / rem, c = j + 1 + e % rem;
            if (c <= r) s[r][c] -= s[r][j] * s[c][j];
        }
        __syncthreads();
    }
    for (int e = tid; e < NB * NB; e += 256) {
        const int r = e / NB, c = e % NB;
        if (c <= r) M[(o + r) * ld + o + c] = (float)s[r][c];
    }
    // inverse of the lower-triangular factor, one column per thread
    if (tid < NB) {
        const int c = tid;
        for (int r = 0; r < NB; ++r) v[r][c] = 0.0;
        for (int r = c; r < NB; ++r) {
            double x = r == c ? 1.0 : 0.0;
            for (int q = c; q < r; ++q) x -= s[r][q] * v[q][c];
            v[r][c] = x / s[r][r];
        }
    }
    __syncthreads();
    for (int e = tid; e < NB * NB; e += 256) linv[e] = v[e / NB][e % NB];
}

// Panel below block k: L_ik = A_ik Linv^T, FP64 accumulation, one row per
// thread; also writes the split-BF16 copy of the panel row.
__global__ __launch_bounds__(256) void k_trsm(float* __restrict__ M, size_t ld, int k, int rows_end,
                                              const double* __restrict__ linv, unsigned short* __restrict__ phi,
                                              unsigned short* __restrict__ plo, const int* __restrict__ flag) {
    if (flag && !*flag) return;
    __shared__ double li[NB][NB + 1];
    for (int e = threadIdx.x; e < NB * NB; e += 256) li[e / NB][e % NB] = linv[e];
    __syncthreads();
    const size_t o = (size_t)k * NB;
    const size_t row = o + NB + (size_t)blockIdx.x * 256 + threadIdx.x;
    if (row >= (size_t)rows_end) return;
    float a[NB];
    float* __restrict__ Ar = M + row * ld + o;
#pragma unroll
    for (int t = 0; t < NB; ++t) a[t] = Ar[t];
#pragma unroll 4
    for (int c = 0; c < NB; ++c) {
        double acc = 0.0;
#pragma unroll
        for (int t = 0; t < NB; ++t)
            if (t <= c) acc += (double)a[t] * li[c][t];
        const float x = (float)acc;
        Ar[c] = x;
        const unsigned short h = f2bf(x);
        phi[row * NB + c] = h;
        plo[row * NB + c] = f2bf(x - bf2f(h));
    }
}

__device__ __forceinline__ void tri_index(int idx, int& I, int& J) {
    int i = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= idx) ++i;
    while (i * (i + 1) / 2 > idx) --i;
    I = i;
    J = idx - i * (i + 1) / 2;
}

// Trailing update with split-BF16 MFMA: C_IJ -= P_I P_J^T over the 64-wide
// panel, three products per K-chunk, FP32 accumulation.  One 64x64 tile
// per workgroup; wavefront w owns rows 16w..16w+15.
// v_mfma_f32_16x16x32_bf16 lane layout: A[m=l%16][k=8(l/16)+0..7],
// B[k=8(l/16)+0..7][n=l%16], C[4(l/16)+0..3][l%16].
__global__ __launch_bounds__(256) void k_syrk_bf16x3(float* __restrict__ M, size_t ld, int k,
                                                     const unsigned short* __restrict__ phi,
                                                     const unsigned short* __restrict__ plo,
                                                     const int* __restrict__ flag) {
    if (flag && !*flag) return;
    int I, Jt;
    tri_index(blockIdx.x, I, Jt);
    const size_t base = (size_t)(k + 1) * NB;
    const size_t r0 = base + (size_t)I * NB, c0 = base + (size_t)Jt * NB;
    const int l = threadIdx.x & 63, w = wave_of();
    const int lr = l & 15, lk = (l >> 4) * 8;
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const size_t arow = r0 + 16 * w + lr;
#pragma unroll
    for (int kc = 0; kc < NB; kc += 32) {
        const bf16x8 ah = *reinterpret_cast<const bf16x8*>(phi + arow * NB + kc + lk);
        const bf16x8 al = *reinterpret_cast<const bf16x8*>(plo + arow * NB + kc + lk);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const size_t brow = c0 + 16 * t + lr;
            const bf16x8 bh = *reinterpret_cast<const bf16x8*>(phi + brow * NB + kc + lk);
            const bf16x8 bl = *reinterpret_cast<const bf16x8*>(plo + brow * NB + kc + lk);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const size_t row = r0 + 16 * w + 4 * (l >> 4) + q;
            const size_t col = c0 + 16 * t + lr;
            if (I == Jt && col > row) continue;
            M[row * ld + col] -= acc[t][q];
        }
    }
}

// FP32 path of the same update: v_mfma_f32_16x16x4_f32, lane layout
// A[m=l%16][k=l/16], B[k=l/16][n=l%16], C as above; operands read from the
// panel in M itself.
__global__ __launch_bounds__(256) void k_syrk_f32(float* __restrict__ M, size_t ld, int k,
                                                  const int* __restrict__ flag) {
    if (flag && !*flag) return;
    int I, Jt;
    tri_index(blockIdx.x, I, Jt);
    const size_t base = (size_t)(k + 1) * NB, pc = (size_t)k * NB;
    const size_t r0 = base + (size_t)I * NB, c0 = base + (size_t)Jt * NB;
    const int l = threadIdx.x & 63, w = wave_of();
    const int lr = l & 15, lk = l >> 4;
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const float* __restrict__ arow = M + (r0 + 16 * w + lr) * ld + pc;
#pragma unroll 4
    for (int kc = 0; kc < NB; kc += 4) {
        const float av = arow[kc + lk];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const float bv = M[(c0 + 16 * t + lr) * ld + pc + kc + lk];
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const size_t row = r0 + 16 * w + 4 * (l >> 4) + q;
            const size_t col = c0 + 16 * t + lr;
            if (I == Jt && col > row) continue;
            M[row * ld + col] -= acc[t][q];
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int grid_tiles(int ntiles) { return ntiles > 0 ? (ntiles + 3) / 4 : 1; }

hipError_t launch_hi_tea_update(const DevJob* jobs, DevJobState* st, const int2* tiles, int ntiles, int njobs,
                                hipStream_t stream) {
    hipLaunchKernelGGL(k_hi_eps, dim3(grid_tiles(ntiles)), dim3(256), 0, stream, jobs, st, tiles, ntiles);
    hipLaunchKernelGGL(k_hi_beta, dim3((njobs + 63) / 64), dim3(64), 0, stream, jobs, st, njobs);
    return hipGetLastError();
}

hipError_t launch_hi_tea_step(const DevJob* jobs, const DevJobState* st, const int2* tiles, int ntiles, int kstep,
                              hipStream_t stream) {
    hipLaunchKernelGGL(k_hi_tea, dim3(grid_tiles(ntiles)), dim3(256), 0, stream, jobs, st, tiles, ntiles, kstep);
    return hipGetLastError();
}

hipError_t launch_chol_factor(float* M, size_t ld, int mp, int mode, double* linv, unsigned short* phi,
                              unsigned short* plo, const int* flag, hipStream_t stream) {
    const int nblk = mp / NB;
    for (int k = 0; k < nblk; ++k) {
        hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(256), 0, stream, M, ld, k, linv, flag);
        const int below = mp - (k + 1) * NB;
        if (below <= 0) break;
        hipLaunchKernelGGL(k_trsm, dim3((below + 255) / 256), dim3(256), 0, stream, M, ld, k, mp, linv, phi, plo,
                           flag);
        const int T = below / NB;
        const int tiles = T * (T + 1) / 2;
        if (mode == SOP_CHOL_F32)
            hipLaunchKernelGGL(k_syrk_f32, dim3(tiles), dim3(256), 0, stream, M, ld, k, flag);
        else
            hipLaunchKernelGGL(k_syrk_bf16x3, dim3(tiles), dim3(256), 0, stream, M, ld, k, phi, plo, flag);
    }
    return hipGetLastError();
}

hipError_t launch_chol_step(const DevJob* jobs, const DevJobState* st, const int2* tiles, int ntiles, int kstep,
                            const HiExactWork* w, hipStream_t stream) {
    hipLaunchKernelGGL(k_chol_D, dim3(grid_tiles(ntiles)), dim3(256), 0, stream, jobs, st, tiles, ntiles, kstep);
    if (w) (void)launch_chol_factor(w->M, w->ld, w->mp, w->mode, w->linv, w->phi, w->plo, nullptr, stream);
    hipLaunchKernelGGL(k_chol_L, dim3(grid_tiles(ntiles)), dim3(256), 0, stream, jobs, st, tiles, ntiles, kstep);
    return hipGetLastError();
}

}  // namespace sop
