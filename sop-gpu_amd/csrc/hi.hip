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
// of 64 with an identity block) by a blocked Cholesky over outer panels of
// kCholPanel columns:
//   the panel's diagonal region, left-looking over its 64x64 blocks in one
//   workgroup: each diagonal block factored in FP64 in LDS, with its
//   inverse (FP64), and the blocks below it in the region solved against
//   it in FP64;
//   the panel's rows below the region, 64 per workgroup, solved the same
//   way with the inverses (X = (A - X' L'^T) Li^T, FP64 accumulation),
//   also emitted as split-BF16 pairs (hi = bf16(x), lo = bf16(x - hi));
//   the trailing update C -= P_I P_J^T on MFMA, either split-BF16 with three
//   products (hi*hi + hi*lo + lo*hi, FP32 accumulate) or FP32 MFMA, the
//   next panel's columns first and the rest on a second stream, overlapped
//   with the next panel's factorisation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "../../include/sopgpu.h"
#include "devmath.h"
#include "hi.h"
#include "hostutil.h"
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
    if ((J.flags & JF_PULLING) && J.bflag[d] == 1) {
        fx = 0.0f; fy = 0.0f; fz = 0.0f;
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

// The all-pairs HI sums (epsilon, TEA step, exact D.F) run kHiSplit
// 1024-thread workgroups per 64-bead tile, 16 beads each, so that one job
// of a few thousand beads fills the chip.
constexpr int kHiSplit = 4;

// Per-bead sums of the TEA epsilon (bdhitea_kernel.cu:212-249): without
// lists, for each chunk of 64 partners j every thread of the workgroup
// forms one (bead, partner) term into LDS and lanes 0-15 add them to their
// bead's sums in j order (the same bits as the bead-by-bead loop); with
// lists, wavefront 0 of the tile's first workgroup walks bead i's lists.
__global__ __launch_bounds__(1024) void k_hi_eps(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                                 const int2* __restrict__ tiles, int ntiles) {
    __shared__ float term[4][16][65];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int2 td = tiles[blockIdx.x / kHiSplit];
    const int q = blockIdx.x % kHiSplit;
    const DevJob& J = jobs[td.x];
    if (!(J.flags & JF_HI) || (J.flags & JF_HI_EXACT)) return;
    const long long s = st[td.x].step;
    if (!(s < J.end_step && s % J.stride == 0 && s % J.hi_eps_freq == 0)) return;
    const bool listed = (J.flags & JF_HI_LISTED) != 0;
    if (listed && (q != 0 || w != 0)) return;
    // listed: lane = bead of the 64-bead tile; unlisted: t & 15 = bead of
    // this workgroup's 16, t >> 4 = partner of the chunk
    const int d = listed ? td.y + lane : td.y + 16 * q + (t & 15);
    const int n = J.n;
    const float4* __restrict__ cur = J.coord[s & 1];
    const float4 c = d < n ? cur[d] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float a = J.hi_a;
    float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
    if (listed) {
        if (d >= n) return;
        for (int k = 0; k < J.cov_cnt[d]; ++k) eps_add(c, cur[J.cov_j[(size_t)k * n + d]], a, sum);
        for (int k = 0; k < J.nat_cnt[d]; ++k) eps_add(c, cur[J.nat_j[(size_t)k * n + d]], a, sum);
        for (int k = 0; k < J.pair_cnt[d]; ++k)
            eps_add(c, cur[list_at(J.pair_j, J.idx16, (size_t)k * J.pair_ld + d)], a, sum);
        J.ci[d] = sum;
        J.eps[d] = sum.w;
        return;
    }
    const int b16 = t & 15, jl = t >> 4;
    for (int base = 0; base < n; base += 64) {
        const int i = base + jl;
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i < n && i != d && d < n) {
            float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
            eps_add(c, cur[i], a, z);
            p = z;
        }
        term[0][b16][jl] = p.x;
        term[1][b16][jl] = p.y;
        term[2][b16][jl] = p.z;
        term[3][b16][jl] = p.w;
        __syncthreads();
        if (t < 16 && d < n) {
            const int m = min(64, n - base);
            for (int k = 0; k < m; ++k) {
                if (base + k == d) continue;
                sum.x += term[0][b16][k];
                sum.y += term[1][b16][k];
                sum.z += term[2][b16][k];
                sum.w += term[3][b16][k];
            }
        }
        __syncthreads();
    }
    if (t < 16 && d < n) {
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

// The TEA step (bdhitea_kernel.cu:364-515).  Unlisted: kHiSplit 1024-thread
// workgroups per 64-bead tile of a TEA job, each serving 16 beads, so a
// single job of a few thousand beads still fills the chip; for each chunk of
// 64 partners every thread forms one (bead, partner) term into LDS, then 16
// lanes of wavefront 0 add them to their beads' sums in partner order, the
// order of the reference's sequential loop.  Listed: wavefront 0 of the
// tile's first workgroup walks the lists.
__global__ __launch_bounds__(1024) void k_hi_tea(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                                 const int2* __restrict__ tiles, int ntiles, int kstep) {
    __shared__ float term[3][16][65];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int2 td = tiles[blockIdx.x / kHiSplit];
    const int q = blockIdx.x % kHiSplit;
    const DevJob& J = jobs[td.x];
    if (!(J.flags & JF_HI) || (J.flags & JF_HI_EXACT)) return;
    const DevJobState& S = st[td.x];
    const long long s = S.step + kstep;
    if (s >= J.end_step) return;
    const bool listed = (J.flags & JF_HI_LISTED) != 0;
    if (listed && (q != 0 || w != 0)) return;
    // listed: lane = bead of the 64-bead tile; unlisted: t & 15 = bead of
    // this workgroup's 16, t >> 4 = partner of the chunk
    const int d = listed ? td.y + lane : td.y + 16 * q + (t & 15);
    const int n = J.n;
    const bool own = d < n;
    const float4* __restrict__ cur = J.coord[s & 1];
    float ci0 = 0.f, ci1 = 0.f, ci2 = 0.f, fx = 0.f, fy = 0.f, fz = 0.f;
    float4 coord = make_float4(0.f, 0.f, 0.f, 0.f);
    if (own) {
        float4 cs = J.ci[d];
        const float4 f0 = J.mforce[d];
        const float4 df = J.rforce[d];
        const float beta = S.beta;
        cs.w = beta * beta;
        ci0 = 1.f / sqrtf(1.f + cs.w * cs.x);
        ci1 = 1.f / sqrtf(1.f + cs.w * cs.y);
        ci2 = 1.f / sqrtf(1.f + cs.w * cs.z);
        coord = cur[d];
        fx = f0.x, fy = f0.y, fz = f0.z;
        fx += df.x * ci0;
        fy += df.y * ci1;
        fz += df.z * ci2;
        ci0 *= beta;
        ci1 *= beta;
        ci2 *= beta;
    }
    const float a = J.hi_a;
    float tx, ty, tz;
    if (listed) {
        if (!own) return;
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
            const int i = list_at(J.pair_j, J.idx16, (size_t)k * J.pair_ld + d);
            tx = 0.f; ty = 0.f; tz = 0.f;
            tea_force(coord, cur[i], J.mforce[i], J.rforce[i], ci0, ci1, ci2, a, tx, ty, tz);
            fx += tx; fy += ty; fz += tz;
        }
        hi_finish(J, d, s, fx, fy, fz, coord);
        return;
    }
    const int b16 = t & 15, jl = t >> 4;
    for (int base = 0; base < n; base += 64) {
        const int i = base + jl;
        tx = 0.f; ty = 0.f; tz = 0.f;
        if (i < n && i != d && own) tea_force(coord, cur[i], J.mforce[i], J.rforce[i], ci0, ci1, ci2, a, tx, ty, tz);
        term[0][b16][jl] = tx;
        term[1][b16][jl] = ty;
        term[2][b16][jl] = tz;
        __syncthreads();
        if (t < 16 && own) {
            const int m = min(64, n - base);
            for (int k = 0; k < m; ++k) {
                if (base + k == d) continue;
                fx += term[0][b16][k];
                fy += term[1][b16][k];
                fz += term[2][b16][k];
            }
        }
        __syncthreads();
    }
    if (t < 16 && own) hi_finish(J, d, s, fx, fy, fz, coord);
}

// ---------------------------------------------------------------------------
// exact path
// ---------------------------------------------------------------------------

// F_i + sum_j D_ij F_j into fdf, summed in bead order.  kHiSplit 1024-thread
// workgroups per 64-bead tile of the (single) exact-HI job of the launch,
// 16 beads each: for each chunk of 64 beads j every thread forms one term
// D_ij F_j (the three row sums, written exactly as the sequential loop
// writes them) into LDS, then lanes 0-15 add them to their bead's sums in j
// order, so every sum is the same bits as the bead-by-bead loop.
__global__ __launch_bounds__(1024) void k_chol_D(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                                 const int2* __restrict__ tiles, int ntiles, int kstep) {
    __shared__ float term[3][16][65];
    const int t = threadIdx.x, b16 = t & 15, jl = t >> 4;
    const int2 td = tiles[blockIdx.x / kHiSplit];
    const DevJob& J = jobs[td.x];
    if (!(J.flags & JF_HI_EXACT)) return;  // one job per launch: uniform
    const long long s = st[td.x].step + kstep;
    if (s >= J.end_step) return;
    const int n = J.n;
    // this thread's bead i: one of the workgroup's 16 (the sums' owners are
    // lanes 0-15 of wavefront 0); jl: its partner in each chunk of 64
    const int d = td.y + 16 * (blockIdx.x % kHiSplit) + b16;
    const float4* __restrict__ cur = J.coord[s & 1];
    const float4 c = d < n ? cur[d] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float a = J.hi_a;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    if (t < 16 && d < n) {
        const float4 f0 = J.mforce[d];
        fx = f0.x, fy = f0.y, fz = f0.z;
    }
    for (int base = 0; base < n; base += 64) {
        const int j = base + jl;
        float tx = 0.f, ty = 0.f, tz = 0.f;
        if (j < n && j != d && d < n) {
            const float4 dr = unit_sep(c, cur[j]);
            const F6 D = rpy(dr, a);
            const float4 g = J.mforce[j];
            tx = D.xx * g.x + D.xy * g.y + D.xz * g.z;
            ty = D.xy * g.x + D.yy * g.y + D.yz * g.z;
            tz = D.xz * g.x + D.yz * g.y + D.zz * g.z;
        }
        term[0][b16][jl] = tx;
        term[1][b16][jl] = ty;
        term[2][b16][jl] = tz;
        __syncthreads();
        if (t < 16 && d < n) {
            const int m = min(64, n - base);
            for (int k = 0; k < m; ++k) {
                if (base + k == d) continue;
                fx += term[0][b16][k];
                fy += term[1][b16][k];
                fz += term[2][b16][k];
            }
        }
        __syncthreads();
    }
    if (t < 16 && d < n) J.fdf[d] = make_float4(fx, fy, fz, 0.0f);
}

// D/D_ii lower block triangle into M on refactorisation steps: block
// (blockIdx.x, blockIdx.y) of 64x64 bead pairs (i in tile x, d in tile y,
// d <= i), identity 3x3 blocks on the diagonal; the same RPY blocks k_chol_D
// sums, written row by row (consecutive threads: consecutive d).
__global__ __launch_bounds__(256) void k_chol_build(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                                    const int2* __restrict__ tiles, int kstep) {
    if (blockIdx.y > blockIdx.x) return;
    const int2 ti = tiles[blockIdx.x], tdd = tiles[blockIdx.y];
    const DevJob& J = jobs[ti.x];
    const long long s = st[ti.x].step + kstep;
    const int n = J.n;
    const float4* __restrict__ cur = J.coord[s & 1];
    const size_t ld = J.chol_ld;
    float* __restrict__ M = J.chol_M;
    const float a = J.hi_a;
    const int dl = threadIdx.x & 63;
    const int d = tdd.y + dl;
    if (d >= n) return;
    const float4 c = cur[d];
    for (int il = threadIdx.x >> 6; il < 64; il += 4) {
        const int i = ti.y + il;
        if (i >= n || i < d) continue;
        const size_t r = 3 * (size_t)i, col = 3 * (size_t)d;
        float* m0 = M + r * ld + col;
        if (i == d) {
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) m0[p * ld + q] = p == q ? 1.f : 0.f;
            continue;
        }
        const F6 D = rpy(unit_sep(c, cur[i]), a);
        m0[0] = D.xx; m0[1] = D.xy; m0[2] = D.xz;
        m0 += ld;
        m0[0] = D.xy; m0[1] = D.yy; m0[2] = D.yz;
        m0 += ld;
        m0[0] = D.xz; m0[1] = D.yz; m0[2] = D.zz;
    }
}

// Noise sum_j B_ij rforce_j for i's three rows, one wavefront per bead, then
// the Euler step.  faithful: each 3x3 block applied as the symmetric matrix
// of its lower triangle, as integrateCholesky_L reads it
// (bdhitea_kernel.cu:322-337); otherwise the factor itself.
__global__ __launch_bounds__(256) void k_chol_L(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                                const int2* __restrict__ tiles, int ntiles, int kstep) {
    const int lane = threadIdx.x & 63;
    // wavefront gw: bead (gw % 64) of tile gw / 64
    for (int gw = blockIdx.x * 4 + wave_of(); gw < ntiles * 64; gw += gridDim.x * 4) {
        const int2 td = tiles[gw >> 6];
        const DevJob& J = jobs[td.x];
        if (!(J.flags & JF_HI_EXACT)) continue;
        const long long s = st[td.x].step + kstep;
        if (s >= J.end_step) continue;
        const int n = J.n;
        const int d = td.y + (gw & 63);
        if (d >= n) continue;
        const size_t ld = J.chol_ld;
        const float* __restrict__ M = J.chol_M;
        const float* __restrict__ g = reinterpret_cast<const float*>(J.rforce);
        const bool faithful = (J.flags & JF_HI_FAITHFUL) != 0;
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

// ---- blocked Cholesky -------------------------------------------------------
static constexpr int NB = 64;    // inner block: diagonal factor, panel solve
static constexpr int NB2 = kCholPanel;  // outer panel: depth of the trailing update

__device__ __forceinline__ unsigned short f2bf(float x) {  // round to nearest even
    unsigned int u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

// A lane's double from another lane of the wavefront (uniform lane index).
__device__ __forceinline__ double readlane_d(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 1/sqrt(p) in FP64: the hardware estimate refined by two Newton steps
// (y <- y (3 - p y^2) / 2), which leaves it within a few ulp of FP64.
__device__ __forceinline__ double rsqrt_d(double p) {
    double y = __builtin_amdgcn_rsq(p);
    y = y * fma(-0.5 * p * y, y, 1.5);
    y = y * fma(-0.5 * p * y, y, 1.5);
    return y;
}

// 16-column steps of the diagonal factorisation, run by one wavefront on
// registers: lanes 0..15 hold the rows of the 16x16 diagonal sub-block,
// lanes 16..63 the rows below it.  Step J (a template parameter, so every
// register index is a constant) scales column J by 1/L[J][J] and subtracts
// L[r][J] L[c][J] from the columns c > J, with L[c][J] broadcast from the
// wavefront's LDS row `col`: for lanes below the sub-block this is the
// forward substitution of the panel solve, so both run as one pass.  Lanes
// 0..15 accumulate scratch above their diagonal, which is never written
// back.  rinv[J] receives 1/L[J][J].
template <int J>
__device__ __forceinline__ void chol16_step(double (&v)[16], int r, double* __restrict__ col,
                                            double* __restrict__ rinv) {
    const double y = rsqrt_d(readlane_d(v[J], J));
    v[J] *= y;  // lane J: L[J][J] = p / sqrt(p); below: a[r][J] / L[J][J]
    if (r == J) rinv[J] = y;
    if constexpr (J + 1 < 16) {
        col[r] = v[J];
#pragma unroll
        for (int c = J + 1; c < 16; ++c) v[c] = fma(-v[J], col[c], v[c]);
        chol16_step<J + 1>(v, r, col, rinv);
    }
}

static constexpr int LDS_ROW = NB + 1;  // row stride of the FP64 block being factored
static constexpr int LDT = NB + 2;      // row stride of the transposed FP64 blocks (16-byte rows)

// Factor the 64x64 FP64 block s (lower triangle, row-major) in place with
// 256 threads: four 16-column sub-blocks, each factored together with the
// rows below it by wavefront 0 in registers (no barrier inside), then a
// rank-16 update of the rest of the block by all threads.  dinv receives
// the reciprocals of the diagonal; col is a 64-double LDS scratch row.
__device__ __forceinline__ void potrf64(double (*s)[LDS_ROW], double* __restrict__ dinv, double* __restrict__ col) {
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
#pragma unroll 1
    for (int sb = 0; sb < NB / 16; ++sb) {
        const int q0 = 16 * sb;
        if (tid < 64) {
            const int r = tid, row = q0 + r;  // lanes 0..15: sub-block rows; 16..63: rows below
            const bool live = row < NB;
            double v[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) v[c] = live && (r >= 16 || c <= r) ? s[row][q0 + c] : 0.0;
            chol16_step<0>(v, r, col, dinv + q0);
            if (live) {
#pragma unroll
                for (int c = 0; c < 16; ++c)
                    if (r >= 16 || c <= r) s[row][q0 + c] = v[c];
            }
        }
        __syncthreads();
        const int b0 = q0 + 16;  // rank-16 update of rows/columns b0 .. 63
        if (b0 < NB) {
            double acc[3][3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) acc[a][b] = 0.0;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                double lr[3], lc[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const int r = b0 + ty + 16 * a, c = b0 + tx + 16 * a;
                    lr[a] = r < NB ? s[r][q0 + q] : 0.0;
                    lc[a] = c < NB ? s[c][q0 + q] : 0.0;
                }
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b) acc[a][b] = fma(lr[a], lc[b], acc[a][b]);
            }
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const int r = b0 + ty + 16 * a, c = b0 + tx + 16 * b;
                    if (r < NB && c <= r) s[r][c] -= acc[a][b];
                }
            __syncthreads();
        }
    }
}

// Off-diagonal 16x16 blocks of the inverse at distance D = I - J:
// Li_IJ = -Li_II (sum_{J<=K<I} L_IK Li_KJ), thread (ty, tx) one element of
// each block, sums split over two accumulators.
template <int D>
__device__ __forceinline__ void inverse_level(const double (*s)[LDS_ROW], double* __restrict__ LiT,
                                              double (*T)[16][17], int ty, int tx) {
#pragma unroll
    for (int J = 0; J + D < 4; ++J) {
        const int I = J + D;
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int t = 16 * J; t < 16 * I; t += 2) {
            a0 = fma(s[16 * I + ty][t], LiT[(16 * J + tx) * LDT + t], a0);
            a1 = fma(s[16 * I + ty][t + 1], LiT[(16 * J + tx) * LDT + t + 1], a1);
        }
        T[J][ty][tx] = a0 + a1;
    }
    __syncthreads();
#pragma unroll
    for (int J = 0; J + D < 4; ++J) {
        const int I = J + D;
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int t = 0; t < 16; t += 2) {
            a0 = fma(LiT[(16 * I + t) * LDT + 16 * I + ty], T[J][t][tx], a0);
            a1 = fma(LiT[(16 * I + t + 1) * LDT + 16 * I + ty], T[J][t + 1][tx], a1);
        }
        LiT[(16 * J + tx) * LDT + 16 * I + ty] = -(a0 + a1);
    }
    __syncthreads();
}

// Inverse of the lower-triangular 64x64 factor L (FP64, in s; dinv holds
// the reciprocals of its diagonal) by 16x16 blocks, 256 threads: the four
// diagonal blocks by forward substitution (wavefront I, lane c: column c of
// inv(L_II)), then the blocks at distance d = I - J = 1, 2, 3.  Output
// transposed: LiT[a*LDT + b] = Li[b][a], zero for a > b.
__device__ __forceinline__ void tri_inverse64(const double (*s)[LDS_ROW], double* __restrict__ LiT,
                                              double (*T)[16][17], const double* __restrict__ dinv) {
    const int tid = threadIdx.x;
    for (int e = tid; e < NB * LDT; e += 256) LiT[e] = 0.0;
    __syncthreads();
    {
        const int I = tid >> 6, c = tid & 63, q0 = 16 * I;
        if (c < 16) {
            double y[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                double acc = 0.0;
#pragma unroll
                for (int t = 0; t < r; ++t) acc = fma(s[q0 + r][q0 + t], y[t], acc);
                y[r] = r < c ? 0.0 : (r == c ? dinv[q0 + r] : -acc * dinv[q0 + r]);
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) LiT[(q0 + c) * LDT + q0 + r] = y[r];
        }
    }
    __syncthreads();
    const int ty = tid >> 4, tx = tid & 15;
    inverse_level<1>(s, LiT, T, ty, tx);
    inverse_level<2>(s, LiT, T, ty, tx);
    inverse_level<3>(s, LiT, T, ty, tx);
}

// The 64x64 block of M at rows r0.., columns kr.. (column block k of the
// panel starting at column c0) minus sum_{j<k} X_j L_j^T over the panel's
// first 64k columns, where X are rows r0.. and L rows kr.. of the panel.
// Wave w computes rows 16w..16w+15 on MFMA, from the split-BF16 copy of the
// panel (three products) or from M in FP32.  The result goes to LDS in
// FP64, row-major (dst[r*dld + c]) or transposed (dst[c*dld + r]).
// v_mfma_f32_16x16x32_bf16 lane layout: A[m=l%16][k=8(l/16)+0..7],
// B[k=8(l/16)+0..7][n=l%16], C[4(l/16)+0..3][l%16]; v_mfma_f32_16x16x4_f32:
// A[m=l%16][k=l/16], B[k=l/16][n=l%16], C as above.
template <int MODE>
__device__ __forceinline__ void block_update(const float* __restrict__ M, size_t ld, size_t r0, size_t kr,
                                             size_t c0, int k, const unsigned short* __restrict__ phi,
                                             const unsigned short* __restrict__ plo, double* __restrict__ dst,
                                             int dld, bool trans) {
    const int l = threadIdx.x & 63, w = wave_of();
    const int lr = l & 15;
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int K = NB * k;
    if constexpr (MODE == SOP_CHOL_BF16X3) {
        const int lk = (l >> 4) * 8;
        const size_t arow = (r0 + 16 * w + lr) * NB2 + lk;
#pragma unroll 1
        for (int jc = 0; jc < K; jc += NB)
#pragma unroll
        for (int kc = jc; kc < jc + NB; kc += 32) {
            const bf16x8 ah = *reinterpret_cast<const bf16x8*>(phi + arow + kc);
            const bf16x8 al = *reinterpret_cast<const bf16x8*>(plo + arow + kc);
            bf16x8 bh[4], bl[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const size_t brow = (kr + 16 * t + lr) * NB2 + lk + kc;
                bh[t] = *reinterpret_cast<const bf16x8*>(phi + brow);
                bl[t] = *reinterpret_cast<const bf16x8*>(plo + brow);
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[t], acc[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[t], acc[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[t], acc[t], 0, 0, 0);
        }
    } else {
        const int lk = l >> 4;
        const float* __restrict__ arow = M + (r0 + 16 * w + lr) * ld + c0 + lk;
#pragma unroll 1
        for (int jc = 0; jc < K; jc += NB)
#pragma unroll 4
        for (int kc = jc; kc < jc + NB; kc += 4) {
            const float av = arow[kc];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float bv = M[(kr + 16 * t + lr) * ld + c0 + lk + kc];
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[t], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = 16 * w + 4 * (l >> 4) + q, col = 16 * t + lr;
            const double v = (double)(M[(r0 + row) * ld + kr + col] - acc[t][q]);
            dst[trans ? col * dld + row : row * dld + col] = v;
        }
}

// X = A Li^T for one 64x64 block in FP64: At[t*LDT + r] = A[r][t] and
// LiT[t*LDT + c] = Li[c][t] (zero for t > c).  Thread (rb, cb) computes
// rows 4rb..4rb+3 and columns 4cb..4cb+3.  X goes to M (rows r0..,
// columns col0..) in FP32 and, in split-BF16 mode, to the panel copy
// (columns pcol0.. of the panel).
template <int MODE>
__device__ __forceinline__ void trsm_block(const double* __restrict__ At, const double* __restrict__ LiT,
                                           float* __restrict__ M, size_t ld, size_t r0, size_t col0,
                                           unsigned short* __restrict__ phi, unsigned short* __restrict__ plo,
                                           int pcol0) {
    const int tid = threadIdx.x, rb = tid >> 4, cb = tid & 15;
    double x[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) x[i][j] = 0.0;
    const int tend = 4 * cb + 4;
    for (int t = 0; t < tend; ++t) {
        const double2 a01 = *reinterpret_cast<const double2*>(At + t * LDT + 4 * rb);
        const double2 a23 = *reinterpret_cast<const double2*>(At + t * LDT + 4 * rb + 2);
        const double2 b01 = *reinterpret_cast<const double2*>(LiT + t * LDT + 4 * cb);
        const double2 b23 = *reinterpret_cast<const double2*>(LiT + t * LDT + 4 * cb + 2);
        const double a[4] = {a01.x, a01.y, a23.x, a23.y}, b[4] = {b01.x, b01.y, b23.x, b23.y};
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) x[i][j] = fma(a[i], b[j], x[i][j]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const size_t row = r0 + 4 * rb + i;
        const float4 v = make_float4((float)x[i][0], (float)x[i][1], (float)x[i][2], (float)x[i][3]);
        *reinterpret_cast<float4*>(M + row * ld + col0 + 4 * cb) = v;
        if constexpr (MODE == SOP_CHOL_BF16X3) {
            const float vv[4] = {v.x, v.y, v.z, v.w};
            ushort4 h, lo;
            unsigned short hh[4], ll[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                hh[j] = f2bf(vv[j]);
                ll[j] = f2bf(vv[j] - bf2f(hh[j]));
            }
            h = make_ushort4(hh[0], hh[1], hh[2], hh[3]);
            lo = make_ushort4(ll[0], ll[1], ll[2], ll[3]);
            *reinterpret_cast<ushort4*>(phi + row * NB2 + pcol0 + 4 * cb) = h;
            *reinterpret_cast<ushort4*>(plo + row * NB2 + pcol0 + 4 * cb) = lo;
        }
    }
}

// Diagonal region of the outer panel at column c0 (kw columns, nb = kw/64
// blocks), one workgroup, left-looking over its blocks k: A_kk minus the
// panel's earlier columns, its FP64 factor and inverse (to linv, block k,
// transposed), then the blocks I > k below it in the region as
// (A_Ik - sum_{j<k} X_Ij L_kj^T) Li_kk^T.
template <int MODE>
__global__ __launch_bounds__(256) void k_panel_diag(float* __restrict__ M, size_t ld, int c0, int kw,
                                                    double* __restrict__ linv, unsigned short* __restrict__ phi,
                                                    unsigned short* __restrict__ plo, const int* __restrict__ flag) {
    if (flag && !*flag) return;
    // the diagonal block being factored (row stride LDS_ROW) and, once it
    // is inverted, the transposed block being solved (row stride LDT)
    __shared__ __attribute__((aligned(16))) double SA[NB * LDT];
    __shared__ __attribute__((aligned(16))) double LiT[NB * LDT];
    __shared__ double T[3][16][17];
    __shared__ double dinv[NB];
    __shared__ double col[NB];
    double(*S)[LDS_ROW] = reinterpret_cast<double(*)[LDS_ROW]>(SA);
    double* const At = SA;
    const int nb = kw / NB, tid = threadIdx.x;
#pragma unroll 1
    for (int k = 0; k < nb; ++k) {
        const size_t kr = (size_t)c0 + (size_t)NB * k;
        block_update<MODE>(M, ld, kr, kr, c0, k, phi, plo, &S[0][0], LDS_ROW, false);
        __syncthreads();
        potrf64(S, dinv, col);
        for (int e = tid; e < NB * NB; e += 256) {
            const int r = e / NB, c = e % NB;
            if (c <= r) M[(kr + r) * ld + kr + c] = (float)S[r][c];
        }
        tri_inverse64(S, LiT, T, dinv);
        for (int e = tid; e < NB * NB; e += 256) linv[(size_t)k * NB * NB + e] = LiT[(e / NB) * LDT + e % NB];
#pragma unroll 1
        for (int I = k + 1; I < nb; ++I) {
            const size_t r0 = (size_t)c0 + (size_t)NB * I;
            block_update<MODE>(M, ld, r0, kr, c0, k, phi, plo, At, LDT, true);
            __syncthreads();
            trsm_block<MODE>(At, LiT, M, ld, r0, kr, phi, plo, NB * k);
            __syncthreads();
        }
    }
}

// Rows of the outer panel below its diagonal region, 64 per workgroup, with
// the diagonal blocks' inverses from linv: for k = 0..nb-1,
// X_Rk = (A_Rk - sum_{j<k} X_Rj L_kj^T) Li_kk^T.
template <int MODE>
__global__ __launch_bounds__(256) void k_panel_rows(float* __restrict__ M, size_t ld, int c0, int kw, int mp,
                                                    const double* __restrict__ linv, unsigned short* __restrict__ phi,
                                                    unsigned short* __restrict__ plo, const int* __restrict__ flag) {
    if (flag && !*flag) return;
    __shared__ __attribute__((aligned(16))) double At[NB * LDT];
    __shared__ __attribute__((aligned(16))) double LiT[NB * LDT];
    const size_t r0 = (size_t)c0 + kw + (size_t)blockIdx.x * NB;
    if (r0 >= (size_t)mp) return;
    const int nb = kw / NB;
#pragma unroll 1
    for (int k = 0; k < nb; ++k) {
        const size_t kr = (size_t)c0 + (size_t)NB * k;
        for (int e = threadIdx.x; e < NB * NB; e += 256) LiT[(e / NB) * LDT + e % NB] = linv[(size_t)k * NB * NB + e];
        block_update<MODE>(M, ld, r0, kr, c0, k, phi, plo, At, LDT, true);
        __syncthreads();
        trsm_block<MODE>(At, LiT, M, ld, r0, kr, phi, plo, NB * k);
        __syncthreads();
    }
}

// Trailing update after an outer panel: C_IJ -= P_I P_J^T over the panel's
// K = kw columns, on 128x128 tiles of the lower triangle from row/column
// s0 on.  Wave w computes the 64x64 quarter (w>>1, w&1) as 4x4 MFMA blocks;
// operands are staged through LDS in K-chunks.  Rows at or beyond mp are
// read as zero and not written.
static constexpr int GT = 128;   // trailing tile edge
static constexpr int KCB = 32;   // K-chunk of the split-BF16 update (bf16 elements)
static constexpr int KCF = 16;   // K-chunk of the FP32 update (floats)
static constexpr int LPB = KCB + 8;  // padded LDS row (bf16 elements)
static constexpr int LPF = KCF + 4;  // padded LDS row (floats)

// Tile (blockIdx.x, blockIdx.y) of the update region: rows [rs, mp),
// columns [cs, ce).  Tiles wholly above the diagonal have nothing to do.
__device__ __forceinline__ bool big_tile(int rs, int cs, size_t& r0, size_t& c0) {
    r0 = (size_t)rs + (size_t)blockIdx.x * GT;
    c0 = (size_t)cs + (size_t)blockIdx.y * GT;
    return c0 < r0 + GT;
}

__device__ __forceinline__ void big_store(float* __restrict__ M, size_t ld, int mp, int ce, size_t r0, size_t c0,
                                          const f32x4 (&acc)[4][4]) {
    const int l = threadIdx.x & 63, w = wave_of();
    const size_t rb = r0 + 64 * (w >> 1), cb = c0 + 64 * (w & 1);
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 4; ++bj)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const size_t row = rb + 16 * bi + 4 * (l >> 4) + q;
                const size_t col = cb + 16 * bj + (l & 15);
                if (row >= (size_t)mp || col >= (size_t)ce || col > row) continue;
                M[row * ld + col] -= acc[bi][bj][q];
            }
}

// Operands of the next K-chunk are loaded into registers while the MFMAs of
// the current chunk run from LDS; they are stored to LDS after the chunk.
__global__ __launch_bounds__(256) void k_gemm_bf16x3(float* __restrict__ M, size_t ld, int rs, int cs, int ce, int mp,
                                                     int kw, const unsigned short* __restrict__ phi,
                                                     const unsigned short* __restrict__ plo) {
    __shared__ __attribute__((aligned(16))) unsigned short sa[2][GT][LPB];  // hi, lo of the row tile
    __shared__ __attribute__((aligned(16))) unsigned short sb[2][GT][LPB];  // hi, lo of the column tile
    size_t r0, c0;
    if (!big_tile(rs, cs, r0, c0)) return;
    const int tid = threadIdx.x, l = tid & 63, w = wave_of();
    const int lr = l & 15, lk = (l >> 4) * 8;
    f32x4 acc[4][4];
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = f32x4{0.f, 0.f, 0.f, 0.f};
    // loader: thread t moves 16 bf16 (two 16-byte words) of row t/2 per array
    const int ldr = tid >> 1, ldk = (tid & 1) * 16;
    const size_t ga = r0 + ldr, gb = c0 + ldr;
    const bool va = ga < (size_t)mp, vb = gb < (size_t)ce;
    uint4 v[8];
    auto fetch = [&](int kc) {
        const uint4 z = make_uint4(0u, 0u, 0u, 0u);
        if (va) {
            const uint4* ph = reinterpret_cast<const uint4*>(phi + ga * NB2 + kc + ldk);
            const uint4* pl = reinterpret_cast<const uint4*>(plo + ga * NB2 + kc + ldk);
            v[0] = ph[0], v[1] = ph[1], v[2] = pl[0], v[3] = pl[1];
        } else {
            v[0] = v[1] = v[2] = v[3] = z;
        }
        if (vb) {
            const uint4* ph = reinterpret_cast<const uint4*>(phi + gb * NB2 + kc + ldk);
            const uint4* pl = reinterpret_cast<const uint4*>(plo + gb * NB2 + kc + ldk);
            v[4] = ph[0], v[5] = ph[1], v[6] = pl[0], v[7] = pl[1];
        } else {
            v[4] = v[5] = v[6] = v[7] = z;
        }
    };
    auto stash = [&]() {
        *reinterpret_cast<uint4*>(&sa[0][ldr][ldk]) = v[0];
        *reinterpret_cast<uint4*>(&sa[0][ldr][ldk + 8]) = v[1];
        *reinterpret_cast<uint4*>(&sa[1][ldr][ldk]) = v[2];
        *reinterpret_cast<uint4*>(&sa[1][ldr][ldk + 8]) = v[3];
        *reinterpret_cast<uint4*>(&sb[0][ldr][ldk]) = v[4];
        *reinterpret_cast<uint4*>(&sb[0][ldr][ldk + 8]) = v[5];
        *reinterpret_cast<uint4*>(&sb[1][ldr][ldk]) = v[6];
        *reinterpret_cast<uint4*>(&sb[1][ldr][ldk + 8]) = v[7];
    };
    fetch(0);
    stash();
    __syncthreads();
    const int ra = 64 * (w >> 1) + lr, rbb = 64 * (w & 1) + lr;
    for (int kc = 0; kc < kw; kc += KCB) {
        const bool more = kc + KCB < kw;
        if (more) fetch(kc + KCB);
        bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            ah[b] = *reinterpret_cast<const bf16x8*>(&sa[0][ra + 16 * b][lk]);
            al[b] = *reinterpret_cast<const bf16x8*>(&sa[1][ra + 16 * b][lk]);
            bh[b] = *reinterpret_cast<const bf16x8*>(&sb[0][rbb + 16 * b][lk]);
            bl[b] = *reinterpret_cast<const bf16x8*>(&sb[1][rbb + 16 * b][lk]);
        }
        // the three products in separate sweeps over the sixteen accumulators,
        // so consecutive MFMAs never wait on each other's results
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
            for (int bj = 0; bj < 4; ++bj)
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bi], bh[bj], acc[bi][bj], 0, 0, 0);
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
            for (int bj = 0; bj < 4; ++bj)
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bi], bl[bj], acc[bi][bj], 0, 0, 0);
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
            for (int bj = 0; bj < 4; ++bj)
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[bi], bh[bj], acc[bi][bj], 0, 0, 0);
        if (more) {
            __syncthreads();  // every wave has read this chunk's operands
            stash();
            __syncthreads();
        }
    }
    big_store(M, ld, mp, ce, r0, c0, acc);
}

// FP32 trailing update: operands are the panel columns [pc, pc+kw) of M.
// LDS rows hold a K-chunk of 16 permuted so that lane group g = l/16 reads
// k = g, g+4, g+8, g+12 (its four MFMA steps) as one float4.  The next
// chunk is loaded into registers during the current chunk's MFMAs.
__global__ __launch_bounds__(256) void k_gemm_f32(float* __restrict__ M, size_t ld, int rs, int cs, int ce, int mp,
                                                  int pc, int kw) {
    __shared__ __attribute__((aligned(16))) float sa[GT][LPF];
    __shared__ __attribute__((aligned(16))) float sb[GT][LPF];
    size_t r0, c0;
    if (!big_tile(rs, cs, r0, c0)) return;
    const int tid = threadIdx.x, l = tid & 63, w = wave_of();
    const int lr = l & 15, lg = l >> 4;
    f32x4 acc[4][4];
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = f32x4{0.f, 0.f, 0.f, 0.f};
    // loader: thread t moves 8 floats (two float4, k = 8h..8h+7) of row t/2
    const int ldr = tid >> 1, h = tid & 1;
    const size_t ga = r0 + ldr, gb = c0 + ldr;
    const bool va = ga < (size_t)mp, vb = gb < (size_t)ce;
    float4 a0, a1, b0, b1;
    auto fetch = [&](int kc) {
        a0 = a1 = b0 = b1 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (va) {
            const float4* pa = reinterpret_cast<const float4*>(M + ga * ld + pc + kc + 8 * h);
            a0 = pa[0], a1 = pa[1];
        }
        if (vb) {
            const float4* pb = reinterpret_cast<const float4*>(M + gb * ld + pc + kc + 8 * h);
            b0 = pb[0], b1 = pb[1];
        }
    };
    // element k (0..15) goes to position (k%4)*4 + k/4
    auto stash = [&]() {
        const int q0 = 2 * h, q1 = 2 * h + 1;  // k/4 of a0 and a1
        sa[ldr][0 + q0] = a0.x, sa[ldr][4 + q0] = a0.y, sa[ldr][8 + q0] = a0.z, sa[ldr][12 + q0] = a0.w;
        sa[ldr][0 + q1] = a1.x, sa[ldr][4 + q1] = a1.y, sa[ldr][8 + q1] = a1.z, sa[ldr][12 + q1] = a1.w;
        sb[ldr][0 + q0] = b0.x, sb[ldr][4 + q0] = b0.y, sb[ldr][8 + q0] = b0.z, sb[ldr][12 + q0] = b0.w;
        sb[ldr][0 + q1] = b1.x, sb[ldr][4 + q1] = b1.y, sb[ldr][8 + q1] = b1.z, sb[ldr][12 + q1] = b1.w;
    };
    fetch(0);
    stash();
    __syncthreads();
    const int ra = 64 * (w >> 1) + lr, rbb = 64 * (w & 1) + lr;
    for (int kc = 0; kc < kw; kc += KCF) {
        const bool more = kc + KCF < kw;
        if (more) fetch(kc + KCF);
        float4 av[4], bv[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            av[b] = *reinterpret_cast<const float4*>(&sa[ra + 16 * b][4 * lg]);
            bv[b] = *reinterpret_cast<const float4*>(&sb[rbb + 16 * b][4 * lg]);
        }
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
            for (int bj = 0; bj < 4; ++bj)
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[bi].x, bv[bj].x, acc[bi][bj], 0, 0, 0);
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
            for (int bj = 0; bj < 4; ++bj)
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[bi].y, bv[bj].y, acc[bi][bj], 0, 0, 0);
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
            for (int bj = 0; bj < 4; ++bj)
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[bi].z, bv[bj].z, acc[bi][bj], 0, 0, 0);
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
            for (int bj = 0; bj < 4; ++bj)
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[bi].w, bv[bj].w, acc[bi][bj], 0, 0, 0);
        if (more) {
            __syncthreads();
            stash();
            __syncthreads();
        }
    }
    big_store(M, ld, mp, ce, r0, c0, acc);
}

// Update of the next panel's columns (on the critical path, so built for
// latency): C -= P_I P_J^T on 64x64 tiles, the four waves of a workgroup
// splitting K (wave w takes the panel's columns 64w..64w+63, nothing when
// beyond kw) and loading their MFMA operands straight from memory, all at
// once; the four partial tiles are summed through LDS in wave order.
template <int MODE>
__global__ __launch_bounds__(256) void k_gemm_ks(float* __restrict__ M, size_t ld, int rs, int cs, int ce, int mp,
                                                 int pc, int kw, const unsigned short* __restrict__ phi,
                                                 const unsigned short* __restrict__ plo) {
    __shared__ float red[4][NB][NB + 1];
    const size_t r0 = (size_t)rs + (size_t)blockIdx.x * NB, c0 = (size_t)cs + (size_t)blockIdx.y * NB;
    if (c0 > r0 + NB - 1) return;  // wholly above the diagonal
    const int tid = threadIdx.x, l = tid & 63, w = wave_of(), lr = l & 15;
    f32x4 acc[4][4];
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int k0 = NB * w;
    if (k0 < kw) {
        if constexpr (MODE == SOP_CHOL_BF16X3) {
            const int lk = (l >> 4) * 8;
            bf16x8 ah[2][4], al[2][4], bh[2][4], bl[2][4];
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const size_t ao = (r0 + 16 * b + lr) * NB2 + k0 + 32 * h + lk;
                    const size_t bo = (c0 + 16 * b + lr) * NB2 + k0 + 32 * h + lk;
                    ah[h][b] = *reinterpret_cast<const bf16x8*>(phi + ao);
                    al[h][b] = *reinterpret_cast<const bf16x8*>(plo + ao);
                    bh[h][b] = *reinterpret_cast<const bf16x8*>(phi + bo);
                    bl[h][b] = *reinterpret_cast<const bf16x8*>(plo + bo);
                }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
#pragma unroll
                for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 4; ++bj)
                        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[h][bi], bh[h][bj], acc[bi][bj], 0, 0, 0);
#pragma unroll
                for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 4; ++bj)
                        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[h][bi], bl[h][bj], acc[bi][bj], 0, 0, 0);
#pragma unroll
                for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 4; ++bj)
                        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[h][bi], bh[h][bj], acc[bi][bj], 0, 0, 0);
            }
        } else {
            // lane group g = l/16 holds k = kc+4g..kc+4g+3; MFMA step i uses
            // component i, so step i covers k = kc+4g+i over the groups g
            const int lg = l >> 4;
            float4 av[4][4], bv[4][4];
#pragma unroll
            for (int h = 0; h < 4; ++h)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    av[h][b] = *reinterpret_cast<const float4*>(M + (r0 + 16 * b + lr) * ld + pc + k0 + 16 * h + 4 * lg);
                    bv[h][b] = *reinterpret_cast<const float4*>(M + (c0 + 16 * b + lr) * ld + pc + k0 + 16 * h + 4 * lg);
                }
#pragma unroll
            for (int h = 0; h < 4; ++h) {
#pragma unroll
                for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 4; ++bj)
                        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[h][bi].x, bv[h][bj].x, acc[bi][bj], 0, 0, 0);
#pragma unroll
                for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 4; ++bj)
                        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[h][bi].y, bv[h][bj].y, acc[bi][bj], 0, 0, 0);
#pragma unroll
                for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 4; ++bj)
                        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[h][bi].z, bv[h][bj].z, acc[bi][bj], 0, 0, 0);
#pragma unroll
                for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 4; ++bj)
                        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[h][bi].w, bv[h][bj].w, acc[bi][bj], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 4; ++bj)
#pragma unroll
            for (int q = 0; q < 4; ++q) red[w][16 * bi + 4 * (l >> 4) + q][16 * bj + lr] = acc[bi][bj][q];
    __syncthreads();
#pragma unroll 4
    for (int e = tid; e < NB * NB; e += 256) {
        const int r = e / NB, c = e % NB;
        const size_t row = r0 + r, col = c0 + c;
        if (row >= (size_t)mp || col >= (size_t)ce || col > row) continue;
        const float s = ((red[0][r][c] + red[1][r][c]) + red[2][r][c]) + red[3][r][c];
        M[row * ld + col] -= s;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

hipError_t launch_hi_tea_update(const DevJob* jobs, DevJobState* st, const int2* tiles, int ntiles, int njobs,
                                hipStream_t stream) {
    if (ntiles > 0)
        hipLaunchKernelGGL(k_hi_eps, dim3(ntiles * kHiSplit), dim3(1024), 0, stream, jobs, st, tiles, ntiles);
    hipLaunchKernelGGL(k_hi_beta, dim3((njobs + 63) / 64), dim3(64), 0, stream, jobs, st, njobs);
    return hipGetLastError();
}

hipError_t launch_hi_tea_step(const DevJob* jobs, const DevJobState* st, const int2* tiles, int ntiles, int kstep,
                              hipStream_t stream) {
    if (ntiles > 0)
        hipLaunchKernelGGL(k_hi_tea, dim3(ntiles * kHiSplit), dim3(1024), 0, stream, jobs, st, tiles, ntiles, kstep);
    return hipGetLastError();
}

// Rows [rs, mp) x columns [cs, ce) of the trailing matrix minus the outer
// panel (columns [pc, pc+kw)) times its transpose.
static void trailing_update(float* M, size_t ld, int rs, int cs, int ce, int mp, int pc, int kw, int mode,
                            const unsigned short* phi, const unsigned short* plo, hipStream_t stream,
                            bool latency = false) {
    if (rs >= mp || cs >= ce) return;
    if (latency) {
        const dim3 g((mp - rs) / NB, (ce - cs) / NB);
        if (mode == SOP_CHOL_F32)
            hipLaunchKernelGGL(k_gemm_ks<SOP_CHOL_F32>, g, dim3(256), 0, stream, M, ld, rs, cs, ce, mp, pc, kw, phi,
                               plo);
        else
            hipLaunchKernelGGL(k_gemm_ks<SOP_CHOL_BF16X3>, g, dim3(256), 0, stream, M, ld, rs, cs, ce, mp, pc, kw,
                               phi, plo);
        return;
    }
    const dim3 g((mp - rs + GT - 1) / GT, (ce - cs + GT - 1) / GT);
    if (mode == SOP_CHOL_F32)
        hipLaunchKernelGGL(k_gemm_f32, g, dim3(256), 0, stream, M, ld, rs, cs, ce, mp, pc, kw);
    else
        hipLaunchKernelGGL(k_gemm_bf16x3, g, dim3(256), 0, stream, M, ld, rs, cs, ce, mp, kw, phi, plo);
}

// Right-looking blocked factorisation with a lookahead of one outer panel.
// After panel p is factored, the columns of panel p+1 are updated on the
// main stream (T1) and the rest of the trailing matrix on `aux` (T2), so
// panel p+1 is factored while T2(p) runs.  T1(p+1) and T2(p) both write
// panel p+2's columns, so the main stream waits for T2(p) before T1(p+1).
// With aux == nullptr the same updates run in order on `stream`.
hipError_t launch_chol_factor(float* M, size_t ld, int mp, int mode, double* linv, unsigned short* phi,
                              unsigned short* plo, const int* flag, hipStream_t stream, hipStream_t aux,
                              hipEvent_t ev_f, hipEvent_t ev_t2) {
    const bool two = aux != nullptr && ev_f != nullptr && ev_t2 != nullptr;
    bool t2_pending = false;
    for (int c0 = 0; c0 < mp; c0 += NB2) {
        const int kw = mp - c0 < NB2 ? mp - c0 : NB2;
        // split-BF16 copies of consecutive panels alternate between two
        // buffers: T2 of panel p still reads its copy while panel p+1 is factored
        const size_t pbuf = (size_t)((c0 / NB2) & 1) * mp * NB2;
        unsigned short* const ph = phi + pbuf;
        unsigned short* const pl = plo + pbuf;
        const int below = mp - c0 - kw;
        if (mode == SOP_CHOL_F32) {
            hipLaunchKernelGGL(k_panel_diag<SOP_CHOL_F32>, dim3(1), dim3(256), 0, stream, M, ld, c0, kw, linv, ph, pl,
                               flag);
            if (below > 0)
                hipLaunchKernelGGL(k_panel_rows<SOP_CHOL_F32>, dim3(below / NB), dim3(256), 0, stream, M, ld, c0, kw,
                                   mp, linv, ph, pl, flag);
        } else {
            hipLaunchKernelGGL(k_panel_diag<SOP_CHOL_BF16X3>, dim3(1), dim3(256), 0, stream, M, ld, c0, kw, linv, ph,
                               pl, flag);
            if (below > 0)
                hipLaunchKernelGGL(k_panel_rows<SOP_CHOL_BF16X3>, dim3(below / NB), dim3(256), 0, stream, M, ld, c0,
                                   kw, mp, linv, ph, pl, flag);
        }
        const int s0 = c0 + kw;
        if (s0 >= mp) break;
        const int s1 = s0 + NB2 < mp ? s0 + NB2 : mp;
        if (two) {
            hipError_t e = hipEventRecord(ev_f, stream);  // panel p factored
            if (e != hipSuccess) return e;
            if (t2_pending && (e = hipStreamWaitEvent(stream, ev_t2, 0)) != hipSuccess) return e;
            trailing_update(M, ld, s0, s0, s1, mp, c0, kw, mode, ph, pl, stream, true);
            if ((e = hipStreamWaitEvent(aux, ev_f, 0)) != hipSuccess) return e;
            trailing_update(M, ld, s1, s1, mp, mp, c0, kw, mode, ph, pl, aux);
            if ((e = hipEventRecord(ev_t2, aux)) != hipSuccess) return e;
            t2_pending = true;
        } else {  // the same two updates in stream order
            trailing_update(M, ld, s0, s0, s1, mp, c0, kw, mode, ph, pl, stream, true);
            trailing_update(M, ld, s1, s1, mp, mp, c0, kw, mode, ph, pl, stream);
        }
    }
    if (t2_pending) {
        hipError_t e = hipStreamWaitEvent(stream, ev_t2, 0);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t launch_chol_step(const DevJob* jobs, const DevJobState* st, const int2* tiles, int ntiles, int kstep,
                            const HiExactWork* w, hipStream_t stream) {
    hipLaunchKernelGGL(k_chol_D, dim3(ntiles * kHiSplit), dim3(1024), 0, stream, jobs, st, tiles, ntiles, kstep);
    if (w) {
        hipLaunchKernelGGL(k_chol_build, dim3(ntiles, ntiles), dim3(256), 0, stream, jobs, st, tiles, kstep);
        // the factorisation's error covers the event record/wait pairs that
        // order its two streams: report it before the noise reads the factor
        const hipError_t e = launch_chol_factor(w->M, w->ld, w->mp, w->mode, w->linv, w->phi, w->plo, nullptr,
                                                stream, w->aux, w->ev_f, w->ev_t2);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_chol_L, dim3(ntiles * 16), dim3(256), 0, stream, jobs, st, tiles, ntiles, kstep);
    return hipGetLastError();
}

}  // namespace sop

// ---------------------------------------------------------------------------
// Factorisation on caller-owned device memory with nothing allocated
// (sop_chol_factor_noalloc): the workspace of launch_chol_factor -- the FP64
// inverses of a panel's diagonal blocks and the two split-BF16 panel
// buffers -- sized for matrices up to kNoAllocMaxOrder, lives in one static
// device array that every GPU the library runs on holds until the process
// ends.  Its device address is taken once per device with
// hipGetSymbolAddress; the array's name is never used as an address on the
// host.
// ---------------------------------------------------------------------------
namespace sop {

constexpr int kNoAllocMaxOrder = 12288;  // 3N for N = 4,096 beads
constexpr size_t kNoAllocPanelBytes = (size_t)2 * kNoAllocMaxOrder * kCholPanel * sizeof(unsigned short);
constexpr size_t kNoAllocScratchBytes = kCholLinvBytes + 2 * kNoAllocPanelBytes;  // linv | phi | plo
static_assert(kCholLinvBytes % 256 == 0 && kNoAllocPanelBytes % 256 == 0, "scratch sections stay 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (the current device).
hipError_t scratch_on(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

// Rows n..mp-1 of M become the identity padding: zero left of the
// diagonal, one on it (the allocating entry point pads the same way).
__global__ __launch_bounds__(256) void k_chol_pad(float* __restrict__ M, size_t ld, int n, int mp) {
    const int r = n + (int)blockIdx.y;
    const int c = (int)(blockIdx.x * 256 + threadIdx.x);
    if (r >= mp || c > r) return;
    M[(size_t)r * ld + c] = c == r ? 1.0f : 0.0f;
}

}  // namespace

}  // namespace sop

extern "C" int sop_chol_factor_noalloc(int device, int n, float* M, unsigned long long ld, int mode, void* stream) {
    using namespace sop;
    if (n <= 0 || !M) return fail(SOP_E_ARG, "cholesky (no allocation): bad arguments");
    if (n > kNoAllocMaxOrder)
        return fail(SOP_E_ARG, "cholesky (no allocation): order " + std::to_string(n) + " exceeds the maximum " +
                                   std::to_string(kNoAllocMaxOrder));
    if (mode != SOP_CHOL_BF16X3 && mode != SOP_CHOL_F32) return fail(SOP_E_ARG, "cholesky (no allocation): unknown mode");
    const int mp = (n + 63) / 64 * 64;
    if (ld < (unsigned long long)mp)
        return fail(SOP_E_ARG, "cholesky (no allocation): leading dimension " + std::to_string(ld) +
                                   " is below the padded order " + std::to_string(mp));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(SOP_E_HIP, "no HIP device available");
    if (device < 0 || device >= ndev || device >= kMaxDevices) return fail(SOP_E_ARG, "device index out of range");
    if (hipSetDevice(device) != hipSuccess) return fail(SOP_E_HIP, "hipSetDevice failed");
    unsigned char* base = nullptr;
    hipError_t e = scratch_on(device, &base);
    if (e != hipSuccess)
        return fail(SOP_E_HIP, std::string("cholesky (no allocation): scratch address: ") + hipGetErrorString(e));
    double* linv = reinterpret_cast<double*>(base);
    unsigned short* phi = reinterpret_cast<unsigned short*>(base + kCholLinvBytes);
    unsigned short* plo = reinterpret_cast<unsigned short*>(base + kCholLinvBytes + kNoAllocPanelBytes);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (mp > n) hipLaunchKernelGGL(k_chol_pad, dim3((mp + 255) / 256, mp - n), dim3(256), 0, s, M, (size_t)ld, n, mp);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_chol_factor(M, (size_t)ld, mp, mode, linv, phi, plo, nullptr, s);
    if (e != hipSuccess) return fail(SOP_E_HIP, std::string("cholesky (no allocation): ") + hipGetErrorString(e));
    return 0;
}
