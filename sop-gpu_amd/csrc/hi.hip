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
// kCholPanel columns, each factored by 64-column steps (k_chol_step), one
// launch per step over the 64-row blocks from the step's diagonal down:
//   the previous step's column block is solved against the inverse of its
//   diagonal block (X = B Li^T on FP64 MFMA) and emitted in FP32 and as
//   split-BF16 pairs (hi = bf16(x), lo = bf16(x - hi));
//   the step's column block is updated by the columns to its left that no
//   trailing update has applied yet (the previous panel's and this
//   panel's), on MFMA -- split-BF16 with three products (hi*hi + hi*lo +
//   lo*hi, FP32 accumulate) or FP32 -- with the K range split over the
//   workgroup's four wavefronts;
//   the diagonal workgroup factors its block in FP64 (16-column passes in
//   registers, rank-16 updates and the inverse's off-diagonal blocks on FP64
//   MFMA) and forms the inverse the next step solves with.
// The rest of the trailing matrix is updated once per outer panel by wide
// MFMA kernels on a second stream, overlapped with the next panel's steps.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>
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
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

struct F6 {
    float xx, xy, xz, yy, yz, zz;
};

// D_ij / D_ii for a normalised separation with its length in w.
__device__ __forceinline__ F6 rpy(const float4 dr, float a) {
    F6 d;
    // the correctly rounded quotients of the reference's expressions; where
    // the operands lie in div_rn_fast's range (ra in (2, 2^59]: every
    // numerator below is then in [2^-60, 1) and every divisor in (2, 2^59]),
    // by the short division, else by the IEEE one -- the same bits
    float ra = div_rn_fast(dr.w, a);
    if (!(div_fast_numerator(dr.w) && a >= kDivFastMinB && a <= kDivFastMaxB)) ra = div_ieee(dr.w, a);
    float crr, cii;
    if (ra > 2.f) {
        constexpr float two_thirds = 2.f / 3.f;
        float q0 = div_rn_fast(0.75f, ra), q1 = div_rn_fast(2.f, ra), q3 = div_rn_fast(two_thirds, ra);
        float q2 = div_rn_fast(q1, ra), q4 = div_rn_fast(q3, ra);
        if (!(ra <= 0x1p59f)) {
            q0 = div_ieee(0.75f, ra);
            q1 = div_ieee(2.f, ra);
            q2 = div_ieee(q1, ra);
            q3 = div_ieee(two_thirds, ra);
            q4 = div_ieee(q3, ra);
        }
        crr = q0 * (1.f - q2);
        cii = q0 * (1.f + q4);
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
    // correctly rounded quotients: the short division where w and the
    // components lie in its range (w in [2^-60, 2^60], each component 0 or
    // of magnitude at least 2^-60; at most w), else the IEEE one
    const float x = dr.x, y = dr.y, z = dr.z, w = dr.w;
    dr.x = div_rn_fast(x, w);
    dr.y = div_rn_fast(y, w);
    dr.z = div_rn_fast(z, w);
    const bool fast = w >= kDivFastMinA && w <= kDivFastMaxA && div_fast_numerator(fabsf(x)) &&
                      div_fast_numerator(fabsf(y)) && div_fast_numerator(fabsf(z));
    if (!fast) {
        dr.x = div_ieee(x, w);
        dr.y = div_ieee(y, w);
        dr.z = div_ieee(z, w);
    }
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

// Ordered all-pairs sums of a workgroup's 16 beads.  Wavefront w owns bead
// dbase + w: for each chunk of kHiChunk partners its lanes form the NC
// components of the (bead, partner) terms -- partner j = base + lane + 64u
// -- into an LDS buffer, and after the chunk's barrier lanes 16c + b of
// wavefront 0 add component c of bead dbase + b in partner order, four
// terms per LDS read, while the other wavefronts already form the next
// chunk into the second buffer.  Terms that the sequential loop skips (the
// bead itself, partners past n) are -0.0f, which leaves every sum unchanged
// (x + -0 = x for every x), so each sum is the same sequence of additions
// as the reference's bead-by-bead loop.  `term(j, v)` fills the wavefront's
// term for partner j; `acc` enters as the summing lane's starting value and
// leaves as its sum.
constexpr int kHiChunk = 256, kHiRow = kHiChunk + 4;
template <int NC, typename Term>
__device__ __forceinline__ void hi_ordered_sums(float (*buf)[NC][16][kHiRow], int n, Term term, float& acc) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int sc = lane >> 4, sb = lane & 15;
    int c = 0;
    for (int base = 0; base < n; base += kHiChunk, ++c) {
        float(*tb)[16][kHiRow] = buf[c & 1];
#pragma unroll
        for (int u = 0; u < kHiChunk / 64; ++u) {
            float v[NC];
            term(base + lane + 64 * u, v);
#pragma unroll
            for (int k = 0; k < NC; ++k) tb[k][w][lane + 64 * u] = v[k];
        }
        __syncthreads();
        if (w == 0 && sc < NC) {
            const float4* __restrict__ row = reinterpret_cast<const float4*>(&tb[sc][sb][0]);
#pragma unroll 8
            for (int k = 0; k < kHiChunk / 4; ++k) {
                const float4 q = row[k];
                acc += q.x;
                acc += q.y;
                acc += q.z;
                acc += q.w;
            }
        }
    }
}

// Per-bead sums of the TEA epsilon (bdhitea_kernel.cu:212-249): without
// lists, the ordered chunk sums of hi_ordered_sums (the same bits as the
// bead-by-bead loop); with lists, wavefront 0 of the tile's first
// workgroup walks bead i's lists.
__global__ __launch_bounds__(1024) void k_hi_eps(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                                 const int2* __restrict__ tiles, int ntiles) {
    __shared__ __attribute__((aligned(16))) float eterm[2][4][16][kHiRow];
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
    // unlisted: wavefront w serves bead dbase + w (c and d above were the
    // listed path's); lanes 16c + b of wavefront 0 sum component c of bead dbase + b
    const int dbase = td.y + 16 * q, dw = dbase + w;
    const float4 cw = dw < n ? cur[dw] : make_float4(0.f, 0.f, 0.f, 0.f);
    float acc = 0.f;
    hi_ordered_sums<4>(eterm, n,
                       [&](int i, float(&v)[4]) {
                           float4 z = make_float4(-0.f, -0.f, -0.f, -0.f);
                           if (i < n && i != dw && dw < n) {
                               z = make_float4(0.f, 0.f, 0.f, 0.f);
                               eps_add(cw, cur[i], a, z);
                           }
                           v[0] = z.x, v[1] = z.y, v[2] = z.z, v[3] = z.w;
                       },
                       acc);
    // component c of bead dbase + b is in lane 16c + b of wavefront 0
    const float sy = __shfl(acc, lane + 16), sz = __shfl(acc, lane + 32), sw = __shfl(acc, lane + 48);
    if (w == 0 && lane < 16 && dbase + lane < n) {
        J.ci[dbase + lane] = make_float4(acc, sy, sz, sw);
        J.eps[dbase + lane] = sw;
    }
}

// TeaUpdater::update (bdhitea.cu:133-176): epsilon and beta' in FP64.  One
// wavefront per job: the per-bead values are staged in LDS 256 at a time and
// added in bead order by every lane (the reference's sequential double sum,
// the same bits; the lanes' sums are equal).
__global__ __launch_bounds__(64) void k_hi_beta(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st,
                                                int njobs) {
    const int jb = blockIdx.x, lane = threadIdx.x;
    if (jb >= njobs) return;
    const DevJob& J = jobs[jb];
    if (!(J.flags & JF_HI) || (J.flags & JF_HI_EXACT)) return;
    DevJobState& S = st[jb];
    const long long s = S.step;
    if (!(s < J.end_step && s % J.stride == 0 && s % J.hi_eps_freq == 0)) return;
    const int N = J.n;
    // the values staged kBetaChunk at a time, all of a chunk's loads in flight
    // together, then added in bead order
    constexpr int kBetaChunk = 4096;
    __shared__ __attribute__((aligned(16))) float chunk[kBetaChunk];
    double epsilon = 0.0;
    for (int base = 0; base < N; base += kBetaChunk) {
        float v[kBetaChunk / 64];
#pragma unroll
        for (int u = 0; u < kBetaChunk / 64; ++u) {
            const int i = base + 64 * u + lane;
            v[u] = i < N ? J.eps[i] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < kBetaChunk / 64; ++u) chunk[64 * u + lane] = v[u];
        __syncthreads();
        const int m = min(kBetaChunk, N - base);
        const float4* __restrict__ c4 = reinterpret_cast<const float4*>(chunk);
        int k = 0;
        for (; k + 4 <= m; k += 4) {
            const float4 q = c4[k >> 2];
            epsilon += (double)q.x;
            epsilon += (double)q.y;
            epsilon += (double)q.z;
            epsilon += (double)q.w;
        }
        for (; k < m; ++k) epsilon += (double)chunk[k];
        __syncthreads();
    }
    if (lane != 0) return;  // the sum is uniform; one lane updates the state
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
// single job of a few thousand beads still fills the chip; the terms are
// added in partner order, the order of the reference's sequential loop, by
// hi_ordered_sums.  Listed: wavefront 0 of the tile's first workgroup walks
// the lists.
__global__ __launch_bounds__(1024) void k_hi_tea(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                                 const int2* __restrict__ tiles, int ntiles, int kstep) {
    __shared__ __attribute__((aligned(16))) float tterm[2][3][16][kHiRow];
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
    // unlisted: wavefront w serves bead dbase + w with that bead's ci
    // (uniform over the wavefront); lanes 16c + b of wavefront 0 sum
    // component c of bead dbase + b from its starting value
    const int dbase = td.y + 16 * q, dw = dbase + w, ds = dbase + (lane & 15);
    const bool ownw = dw < n;
    float wi0 = 0.f, wi1 = 0.f, wi2 = 0.f;
    float4 cw = make_float4(0.f, 0.f, 0.f, 0.f);
    if (ownw) {
        const float4 cs = J.ci[dw];
        const float bb = S.beta * S.beta;
        wi0 = 1.f / sqrtf(1.f + bb * cs.x) * S.beta;
        wi1 = 1.f / sqrtf(1.f + bb * cs.y) * S.beta;
        wi2 = 1.f / sqrtf(1.f + bb * cs.z) * S.beta;
        cw = cur[dw];
    }
    float acc = 0.f;
    if (w == 0 && lane < 48 && ds < n) {  // the bead's starting value, component lane / 16
        const float4 cs = J.ci[ds], f0 = J.mforce[ds], df = J.rforce[ds];
        const float bb = S.beta * S.beta;
        const int k = lane >> 4;
        const float cc = k == 0 ? cs.x : (k == 1 ? cs.y : cs.z);
        const float f = k == 0 ? f0.x : (k == 1 ? f0.y : f0.z), g = k == 0 ? df.x : (k == 1 ? df.y : df.z);
        acc = f;
        acc += g * (1.f / sqrtf(1.f + bb * cc));
    }
    hi_ordered_sums<3>(tterm, n,
                       [&](int i, float(&v)[3]) {
                           float x = -0.f, y = -0.f, z = -0.f;
                           if (i < n && i != dw && ownw) {
                               x = 0.f, y = 0.f, z = 0.f;
                               tea_force(cw, cur[i], J.mforce[i], J.rforce[i], wi0, wi1, wi2, a, x, y, z);
                           }
                           v[0] = x, v[1] = y, v[2] = z;
                       },
                       acc);
    const float sy = __shfl(acc, lane + 16), sz = __shfl(acc, lane + 32);
    if (w == 0 && lane < 16 && ds < n) hi_finish(J, ds, s, acc, sy, sz, cur[ds]);
}

// ---------------------------------------------------------------------------
// exact path
// ---------------------------------------------------------------------------

// F_i + sum_j D_ij F_j into fdf, summed in bead order.  kHiSplit 1024-thread
// workgroups per 64-bead tile of the (single) exact-HI job of the launch,
// 16 beads each, wavefront w forming bead dbase + w's terms D_ij F_j (the
// three row sums, written exactly as the sequential loop writes them) and
// wavefront 0 adding them in j order (hi_ordered_sums), so every sum is the
// same bits as the bead-by-bead loop.
__global__ __launch_bounds__(1024) void k_chol_D(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                                 const int2* __restrict__ tiles, int ntiles, int kstep) {
    __shared__ __attribute__((aligned(16))) float dterm[2][3][16][kHiRow];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int2 td = tiles[blockIdx.x / kHiSplit];
    const DevJob& J = jobs[td.x];
    if (!(J.flags & JF_HI_EXACT)) return;  // one job per launch: uniform
    const long long s = st[td.x].step + kstep;
    if (s >= J.end_step) return;
    const int n = J.n;
    const int dbase = td.y + 16 * (blockIdx.x % kHiSplit), dw = dbase + w, ds = dbase + (lane & 15);
    const float4* __restrict__ cur = J.coord[s & 1];
    const float4 c = dw < n ? cur[dw] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float a = J.hi_a;
    float acc = 0.f;
    if (w == 0 && lane < 48 && ds < n) {
        const float4 f0 = J.mforce[ds];
        acc = lane < 16 ? f0.x : (lane < 32 ? f0.y : f0.z);
    }
    hi_ordered_sums<3>(dterm, n,
                       [&](int j, float(&v)[3]) {
                           float tx = -0.f, ty = -0.f, tz = -0.f;
                           if (j < n && j != dw && dw < n) {
                               const float4 dr = unit_sep(c, cur[j]);
                               const F6 D = rpy(dr, a);
                               const float4 g = J.mforce[j];
                               tx = D.xx * g.x + D.xy * g.y + D.xz * g.z;
                               ty = D.xy * g.x + D.yy * g.y + D.yz * g.z;
                               tz = D.xz * g.x + D.yz * g.y + D.zz * g.z;
                           }
                           v[0] = tx, v[1] = ty, v[2] = tz;
                       },
                       acc);
    const float sy = __shfl(acc, lane + 16), sz = __shfl(acc, lane + 32);
    if (w == 0 && lane < 16 && ds < n) J.fdf[ds] = make_float4(acc, sy, sz, 0.0f);
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
        // unrolled so that several columns' loads are in flight per lane; each
        // lane still adds its columns in order
#pragma unroll 4
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

// The diagonal machinery below (factor and inverse of a 64x64 block, the
// panel solve with the inverse) runs in the scalar type R: FP64 for the FP32
// factorisation, FP32 for the split-BF16 one (whose trailing updates carry
// errors near 1e-6 relative; an FP32 diagonal adds ~1e-7, and halves the
// cost of the serial pivot chain and of the inverse's MFMA work).
template <int MODE>
using CholReal = typename std::conditional<MODE == SOP_CHOL_BF16X3, float, double>::type;
template <typename R>
using R4 = typename std::conditional<std::is_same<R, double>::value, f64x4, f32x4>::type;

// A lane's value from another lane of the wavefront (uniform lane index).
__device__ __forceinline__ double readlane_r(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float readlane_r(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// 1/sqrt(p): the hardware estimate refined by one Newton step (y <- y (3 -
// p y^2) / 2).  FP64: estimate error below 2^-22, after the step near 2^-44
// -- far below the FP32 rounding of the factor it feeds; FP32: within an
// ulp or two of the FP32 value.
__device__ __forceinline__ double rsqrt_r(double p) {
    const double y = __builtin_amdgcn_rsq(p);
    return y * fma(-0.5 * p * y, y, 1.5);
}
__device__ __forceinline__ float rsqrt_r(float p) {
    const float y = __builtin_amdgcn_rsqf(p);
    return y * fmaf(-0.5f * p * y, y, 1.5f);
}

// acc += A B on the 16x16x4 MFMA of R.  Lane layout of both: A[i=l%16][k=l/16],
// B[k=l/16][j=l%16]; D row drow(l/16, v), column l%16 in register v: FP64
// interleaves the rows (l/16 + 4v), FP32 blocks them (4(l/16) + v).
__device__ __forceinline__ f64x4 mfma16(double a, double b, f64x4 acc) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma16(float a, float b, f32x4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
}
template <typename R>
__device__ __forceinline__ int drow(int lk, int v) {
    return std::is_same<R, double>::value ? lk + 4 * v : 4 * lk + v;
}

// 16-column steps of the diagonal factorisation, run by one wavefront on
// registers: lanes 0..15 hold the rows of the 16x16 diagonal sub-block,
// lanes 16..63 the rows below it.  Step J (a template parameter, so every
// register index is a constant) scales column J by 1/L[J][J] and subtracts
// L[r][J] L[c][J] from the columns c > J, with L[c][J] read from lane c
// (readlane, no LDS round trip): for lanes below the sub-block this is the
// forward substitution of the panel solve, so both run as one pass.  Lanes
// 0..15 accumulate scratch above their diagonal, which is never written
// back.  y[J] receives 1/L[J][J] (uniform).
template <int J, typename R>
__device__ __forceinline__ void chol16_step(R (&v)[16], R (&y)[16]) {
    y[J] = rsqrt_r(readlane_r(v[J], J));
    v[J] *= y[J];  // lane J: L[J][J] = p / sqrt(p); below: a[r][J] / L[J][J]
    if constexpr (J + 1 < 16) {
#pragma unroll
        for (int c = J + 1; c < 16; ++c) v[c] = fma(-v[J], readlane_r(v[J], c), v[c]);
        chol16_step<J + 1>(v, y);
    }
}

static constexpr int LDS_ROW = NB + 1;  // row stride of the block being factored
// row stride of the transposed blocks (Li^T): 16-byte rows, and for FP32 the
// four row groups of an MFMA operand read in distinct bank ranges
template <typename R>
constexpr int ldt() { return std::is_same<R, double>::value ? NB + 2 : NB + 16; }

// Column block K of the 64x64 block s (rows 16K..63) factored by
// wavefront 0 in registers: lane r holds row 16K + r (lanes 0..15 the
// diagonal sub-block, the others the rows below, forward-substituted in
// the same pass).  Lanes past row 63 read a clamped row and keep zeros;
// lanes 0..15 write their scratch above the diagonal back, where nothing
// reads it.  dinv receives the reciprocals of the sub-block's diagonal.
template <typename R>
__device__ __forceinline__ void chol_pass(R (*s)[LDS_ROW], R* __restrict__ dinv, int K) {
    const int r = threadIdx.x & 63, q0 = 16 * K, row = q0 + r;
    const bool live = row < NB;
    const int rr = live ? row : NB - 1;
    R v[16], y[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const R e = s[rr][q0 + c];
        v[c] = live && (r >= 16 || c <= r) ? e : R(0);
    }
    chol16_step<0>(v, y);
    if (live) {
#pragma unroll
        for (int c = 0; c < 16; ++c) s[row][q0 + c] = v[c];
    }
    if (r == 0) {
#pragma unroll
        for (int c = 0; c < 16; ++c) dinv[q0 + c] = y[c];
    }
}

// C_IJ -= L_IK L_JK^T for the 16x16 tile (I, J) of s on MFMA, one wavefront.
template <typename R>
__device__ __forceinline__ void rank16_tile(R (*s)[LDS_ROW], int I, int J, int K) {
    const int l = threadIdx.x & 63, lr = l & 15, lk = l >> 4;
    R4<R> acc;
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] = s[16 * I + drow<R>(lk, v)][16 * J + lr];
#pragma unroll
    for (int k0 = 0; k0 < 16; k0 += 4)
        acc = mfma16(-s[16 * I + lr][16 * K + k0 + lk], s[16 * J + lr][16 * K + k0 + lk], acc);
#pragma unroll
    for (int v = 0; v < 4; ++v) s[16 * I + drow<R>(lk, v)][16 * J + lr] = acc[v];
}

// Column c (lane c < 16) of inv(L_KK) by forward substitution, y_t =
// (e_c - sum_{u<t} L_tu y_u) / L_tt with the sums accumulated right-looking
// so each y_t waits for one product; written transposed to LiT (LiT[a*LDT
// + b] = Li[b][a]).
template <typename R>
__device__ __forceinline__ void inv_diag16(const R (*s)[LDS_ROW], R* __restrict__ LiT, const R* __restrict__ dinv,
                                           int K) {
    constexpr int LDT = ldt<R>();
    const int c = threadIdx.x & 63, q0 = 16 * K;
    if (c < 16) {
        R acc[16], y[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = R(0);
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            y[t] = t < c ? R(0) : (t == c ? dinv[q0 + t] : -acc[t] * dinv[q0 + t]);
#pragma unroll
            for (int r = t + 1; r < 16; ++r) acc[r] = fma(s[q0 + r][q0 + t], y[t], acc[r]);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) LiT[(q0 + c) * LDT + q0 + r] = y[r];
    }
}

// Off-diagonal 16x16 block (I, J) of the inverse, J < I, on MFMA by one
// wavefront, in two halves: T = sum_{J<=K<I} L_IK Li_KJ into the
// wavefront's LDS tile (needs the rows J..I-1 of the inverse), then
// Li_IJ = -Li_II T (needs Li_II).
template <typename R>
__device__ __forceinline__ void inv_sum16(const R (*s)[LDS_ROW], const R* __restrict__ LiT, R (*T)[17], int I,
                                          int J) {
    constexpr int LDT = ldt<R>();
    const int l = threadIdx.x & 63, lr = l & 15, lk = l >> 4;
    R4<R> acc = {R(0), R(0), R(0), R(0)};
    for (int k0 = 0; k0 < 16 * (I - J); k0 += 4)
        acc = mfma16(s[16 * I + lr][16 * J + k0 + lk], LiT[(16 * J + lr) * LDT + 16 * J + k0 + lk], acc);
#pragma unroll
    for (int v = 0; v < 4; ++v) T[drow<R>(lk, v)][lr] = acc[v];
}
template <typename R>
__device__ __forceinline__ void inv_apply16(R* __restrict__ LiT, const R (*T)[17], int I, int J) {
    constexpr int LDT = ldt<R>();
    const int l = threadIdx.x & 63, lr = l & 15, lk = l >> 4;
    R4<R> li = {R(0), R(0), R(0), R(0)};
#pragma unroll
    for (int k0 = 0; k0 < 16; k0 += 4) li = mfma16(-LiT[(16 * I + k0 + lk) * LDT + 16 * I + lr], T[k0 + lk][lr], li);
#pragma unroll
    for (int v = 0; v < 4; ++v) LiT[(16 * J + lr) * LDT + 16 * I + drow<R>(lk, v)] = li[v];
}

// Factor the 64x64 block s (lower triangle, row-major) in place and form
// the inverse of the factor, transposed, in LiT (zero above its diagonal),
// with 256 threads.  The factor runs in four rounds, one per 16-column
// block K: wavefront 0 applies the previous block's rank-16 update to block
// K's tiles, then factors block K in registers; meanwhile wavefronts 2 and
// 3 apply that update to the tiles right of block K, and wavefront 1 forms
// row K-1 of the inverse (its diagonal block by substitution, the blocks
// left of it on MFMA) -- so the inverse trails the factor by one round
// instead of following it.  The inverse's last row is formed after the
// last round.  dinv receives the reciprocals of the factor's diagonal.
template <typename R>
__device__ __forceinline__ void potrf_inv64(R (*s)[LDS_ROW], R* __restrict__ LiT, R (*T)[16][17],
                                            R* __restrict__ dinv) {
    constexpr int LDT = ldt<R>();
    const int tid = threadIdx.x, w = wave_of();
    if (w == 0)
        chol_pass(s, dinv, 0);
    else
        for (int e = tid - 64; e < NB * LDT; e += 192) LiT[e] = R(0);
    __syncthreads();
#pragma unroll 1
    for (int K = 1; K < NB / 16; ++K) {
        const int k = K - 1;  // the block factored in the previous round
        if (w == 0) {
            for (int I = K; I < NB / 16; ++I) rank16_tile(s, I, K, k);
            chol_pass(s, dinv, K);
        } else if (w == 1) {
            inv_diag16(s, LiT, dinv, k);
            for (int J = 0; J < k; ++J) {
                inv_sum16(s, LiT, T[0], k, J);
                inv_apply16(LiT, T[0], k, J);
            }
        } else {
            // tiles (I, J), K < J <= I, dealt to wavefronts 2 and 3 in turn
            int t = 0;
            for (int J = K + 1; J < NB / 16; ++J)
                for (int I = J; I < NB / 16; ++I, ++t)
                    if ((t & 1) == w - 2) rank16_tile(s, I, J, k);
        }
        __syncthreads();
    }
    // row 3 of the inverse: its diagonal block (wavefront 0) beside the sums
    // of the blocks left of it (wavefronts 1..3), then the products
    constexpr int RL = NB / 16 - 1;
    if (w == 0)
        inv_diag16(s, LiT, dinv, RL);
    else
        inv_sum16(s, LiT, T[w - 1], RL, w - 1);
    __syncthreads();
    if (w != 0) inv_apply16(LiT, T[w - 1], RL, w - 1);
    __syncthreads();
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

// Tile (blockIdx.x, blockIdx.y) of the update region: rows and columns
// [s, mp).  Tiles wholly above the diagonal have nothing to do.
__device__ __forceinline__ bool big_tile(int s, size_t& r0, size_t& c0) {
    r0 = (size_t)s + (size_t)blockIdx.x * GT;
    c0 = (size_t)s + (size_t)blockIdx.y * GT;
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
__global__ __launch_bounds__(256) void k_gemm_bf16x3(float* __restrict__ M, size_t ld, int s, int mp, int kw,
                                                     const unsigned short* __restrict__ phi,
                                                     const unsigned short* __restrict__ plo) {
    __shared__ __attribute__((aligned(16))) unsigned short sa[2][GT][LPB];  // hi, lo of the row tile
    __shared__ __attribute__((aligned(16))) unsigned short sb[2][GT][LPB];  // hi, lo of the column tile
    const int ce = mp;
    size_t r0, c0;
    if (!big_tile(s, r0, c0)) return;
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

// Split-BF16 trailing update on 256x256 tiles for the large trailing
// regions: four times the MFMA work of a 128x128 tile per operand byte
// staged, half the operand traffic per flop.  One workgroup of eight
// wavefronts per CU (two per SIMD); wavefront w forms the 64x128 block
// (w>>1, w&1) as 2x4 tiles of 32x32 on v_mfma_f32_32x32x16_bf16 (lane l:
// A[i=l%32][k=8(l/32)..+7], B[k=8(l/32)..+7][j=l%32], D row
// 8(q/4)+4(l/32)+q%4, column l%32 in accumulator register q).  Per
// 32-column chunk the hi and lo halves of the tile's 256 panel rows and 256
// panel columns (64 KB) are staged in LDS in 16-byte units, unit u of row r
// stored at u ^ ((r >> 1) & 3) so that the eight lanes of each read phase
// hit distinct banks; the next chunk is loaded into registers while the
// current one is multiplied.  Tiles are numbered over the lower triangle
// row by row and dealt to the XCDs in contiguous ranges (workgroup i runs
// on XCD i % 8), so the tiles an XCD holds at once share their row panels
// in its L2.
static constexpr int BT = 256;  // tile edge
static constexpr int BK = 32;   // K chunk (bf16 elements): four 16-byte units per row
constexpr int kXcds = 8;        // MI355X: workgroup i is dispatched to XCD i % 8

__device__ __forceinline__ int swz(int r, int u) { return u ^ ((r >> 1) & 3); }

// The loader of k_trail_bf16x3: thread t moves unit t%4 of rows t/4 and
// t/4 + 128, hi and lo, of the row panel (rows from r0) and of the column
// panel (rows from c0) -- eight 16-byte loads, zeros past row mp -- into v,
// and v into LDS.  Written as macros: the same code as helper functions or
// lambdas makes this compiler (ROCm 7.2 clang) crash in its machine copy
// propagation pass.
#define TRAIL_FETCH(KC)                                                                       \
    for (int pnl = 0; pnl < 2; ++pnl)                                                         \
        for (int q = 0; q < 2; ++q) {                                                         \
            const size_t g = (pnl ? c0 : r0) + (size_t)(lrow + 128 * q);                      \
            if (g < (size_t)mp) {                                                             \
                const size_t off = g * NB2 + (size_t)(KC) + 8 * lu;                           \
                v[pnl][0][q] = *reinterpret_cast<const uint4*>(phi + off);                    \
                v[pnl][1][q] = *reinterpret_cast<const uint4*>(plo + off);                    \
            } else {                                                                          \
                v[pnl][0][q] = v[pnl][1][q] = make_uint4(0u, 0u, 0u, 0u);                     \
            }                                                                                 \
        }
#define TRAIL_STASH()                                                                         \
    for (int pnl = 0; pnl < 2; ++pnl)                                                         \
        for (int h = 0; h < 2; ++h)                                                           \
            for (int q = 0; q < 2; ++q) {                                                     \
                const int r = lrow + 128 * q;                                                 \
                sm[pnl][h][r][swz(r, lu)] = v[pnl][h][q];                                     \
            }

__global__ __launch_bounds__(512) void k_trail_bf16x3(float* __restrict__ M, size_t ld, int s, int mp, int kw,
                                                      int ntr, const unsigned short* __restrict__ phi,
                                                      const unsigned short* __restrict__ plo) {
    __shared__ uint4 sm[2][2][BT][BK / 8];  // [row panel, column panel][hi, lo][row][unit]: 64 KB
    const int ntile = ntr * (ntr + 1) / 2, per = (ntile + kXcds - 1) / kXcds;
    const int t = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
    if (t >= ntile) return;
    int bi = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while (bi * (bi + 1) / 2 > t) --bi;
    while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
    const int bj = t - bi * (bi + 1) / 2;
    const size_t r0 = (size_t)s + (size_t)BT * bi, c0 = (size_t)s + (size_t)BT * bj;
    const int tid = threadIdx.x, l = tid & 63, w = wave_of();
    // loader: thread t moves unit t%4 of rows t/4 and t/4 + 128, hi and lo,
    // of the row panel and of the column panel: eight 16-byte loads
    const int lu = tid & 3, lrow = tid >> 2;
    uint4 v[2][2][2];
    f32x16 acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x16{};
    TRAIL_FETCH(0);
    TRAIL_STASH();
    __syncthreads();
    const int wr = 64 * (w >> 1), wc = 128 * (w & 1), li = l & 31, lk = l >> 5;
    for (int kc = 0; kc < kw; kc += BK) {
        const bool more = kc + BK < kw;
        if (more) { TRAIL_FETCH(kc + BK); }
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {  // two K = 16 substeps per chunk
            const int u = 2 * ks + lk;
            bf16x8 ah[2], al[2], bh[4], bl[4];
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int ra = wr + 32 * b + li;
                const uint4 xa = sm[0][0][ra][swz(ra, u)], ya = sm[0][1][ra][swz(ra, u)];
                ah[b] = *reinterpret_cast<const bf16x8*>(&xa);
                al[b] = *reinterpret_cast<const bf16x8*>(&ya);
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int rb = wc + 32 * b + li;
                const uint4 xb = sm[1][0][rb][swz(rb, u)], yb = sm[1][1][rb][swz(rb, u)];
                bh[b] = *reinterpret_cast<const bf16x8*>(&xb);
                bl[b] = *reinterpret_cast<const bf16x8*>(&yb);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            __syncthreads();  // every wave has read this chunk
            TRAIL_STASH();
            __syncthreads();
        }
    }
    // C -= acc over the lower triangle of the tile, rows below mp
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const size_t row = r0 + wr + 32 * i + 8 * (q >> 2) + 4 * lk + (q & 3);
                const size_t col = c0 + wc + 32 * j + li;
                if (row >= (size_t)mp || col > row) continue;
                M[row * ld + col] -= acc[i][j][q];
            }
}
#undef TRAIL_FETCH
#undef TRAIL_STASH

// FP32 trailing update: operands are the panel columns [pc, pc+kw) of M.
// LDS rows hold a K-chunk of 16 permuted so that lane group g = l/16 reads
// k = g, g+4, g+8, g+12 (its four MFMA steps) as one float4.  The next
// chunk is loaded into registers during the current chunk's MFMAs.
__global__ __launch_bounds__(256) void k_gemm_f32(float* __restrict__ M, size_t ld, int s, int mp, int pc, int kw) {
    __shared__ __attribute__((aligned(16))) float sa[GT][LPF];
    __shared__ __attribute__((aligned(16))) float sb[GT][LPF];
    const int ce = mp;
    size_t r0, c0;
    if (!big_tile(s, r0, c0)) return;
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

// ---- the factorisation by fused 64-column steps -----------------------------
// Within an outer panel (columns c0 .. c0+kw-1, nb = kw/64 column blocks),
// step j (0 <= j <= nb) is one launch over the 64-row blocks g >= c0/64 + j
// (workgroup x serves g = c0/64 + j + x; d = c0/64 + j is the step's
// diagonal block):
//  (1) for j >= 1, each row block's solve against column block j-1:
//      X_g = B_g Li^T with B_g the block as the previous step left it and
//      Li^T the inverse of L_{j-1,j-1} (FP64 accumulation), written to M and,
//      split-BF16 mode, to the panel copy;
//  (2) for j < nb, column block j of each row block minus the products of
//      every column to its left that no trailing update has applied: the
//      previous outer panel's columns (its lookahead update, folded in here)
//      and this panel's columns 0 .. 64j-1, on MFMA:
//      B_g = A_gj - sum_k X_gk X_dk^T.  Workgroup 0 (g = d) factors B in
//      FP64 into L_jj and its inverse; the others store B for step j+1.
// Every workgroup forms X_d for (2) itself (the diagonal row block's solve,
// the same arithmetic as workgroup 0's), so no workgroup waits for another.
// Lookahead: the diagonal workgroup's chain (solve, update, factor, inverse)
// is the step's critical path, so the long part of its update -- every
// column left of j-1 -- is formed one step early: while step j-1 runs, one
// extra workgroup solves row block d for column block j-2 itself and forms
// P = A_dd - sum_{k < j-1} X_dk X_dk^T into a step buffer; step j's
// diagonal workgroup then only subtracts X_d,j-1 X_d,j-1^T (64 columns)
// before it factors.  The panel's first step has no such block and forms
// the whole sum.
// The trailing matrix beyond the next panel is updated by the wide MFMA
// kernels on the second stream, one pass per outer panel.
static constexpr int XS = NB + 4;  // row stride of the FP32 X blocks in LDS (16-byte rows)

struct CholStep {
    float* M;
    size_t ld;
    int mp;
    int c0, kw;                      // this outer panel
    int pc, pkw;                     // the previous outer panel (pkw = 0: none)
    int j;                           // step 0 .. nb
    const void* li_in;               // Li^T of column block j-1 (64 x 64 CholReal, row a = column a of Li)
    void* li_out;                    // Li^T of column block j
    const float* b_in;               // B of column block j-1, mp x 64 by row
    float* b_out;                    // B of column block j
    const float* p_in;               // the diagonal block's P formed by the previous step (j >= 1)
    float* p_out;                    // the next diagonal block's P (when look)
    int look;                        // 1: an extra workgroup forms P for step j+1
    unsigned short *phi, *plo;       // this panel's split-BF16 copy, mp x NB2
    const unsigned short *pphi, *pplo;  // the previous panel's copy
};

// The 64 values of B a lane feeds the solve's MFMAs: B[16w + l%16][4s + l/16]
// for s = 0..15 (B is 64 x 64 FP32, rows r0.. of a row-major mp x 64 buffer).
__device__ __forceinline__ void load_solve_rows(const float* __restrict__ b, size_t r0, float (&bv)[16]) {
    const int l = threadIdx.x & 63, w = wave_of();
    const float* __restrict__ row = b + (r0 + 16 * w + (l & 15)) * NB + (l >> 4);
#pragma unroll
    for (int q = 0; q < 16; ++q) bv[q] = row[4 * q];
}

// X = B Li^T for one 64x64 block on MFMA in CholReal (B from
// load_solve_rows, LiT[t*LDT + c] = Li[c][t], zero for t > c), to Xs (LDS,
// row stride XS) and, when `out`, to M (rows r0.., columns col0..) and the
// panel copy (columns pcol0..).  Wave w forms rows 16w..16w+15; output
// block jb (columns 16jb..) sums t < 16(jb+1) only, where Li^T is not zero.
template <int MODE>
__device__ __forceinline__ void solve_block(const float (&bv)[16], const CholReal<MODE>* __restrict__ LiT,
                                            float* __restrict__ Xs, bool out, float* __restrict__ M, size_t ld,
                                            size_t r0, size_t col0, unsigned short* __restrict__ phi,
                                            unsigned short* __restrict__ plo, int pcol0) {
    using R = CholReal<MODE>;
    constexpr int LDT = ldt<R>();
    const int l = threadIdx.x & 63, w = wave_of(), lr = l & 15, lk = l >> 4;
    R4<R> acc[4];
#pragma unroll
    for (int jb = 0; jb < 4; ++jb) acc[jb] = R4<R>{R(0), R(0), R(0), R(0)};
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int t0 = 4 * q;
        const R a = (R)bv[q];
#pragma unroll
        for (int jb = 0; jb < 4; ++jb)
            if (t0 < 16 * (jb + 1)) acc[jb] = mfma16(a, LiT[(t0 + lk) * LDT + 16 * jb + lr], acc[jb]);
    }
#pragma unroll
    for (int jb = 0; jb < 4; ++jb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = 16 * w + drow<R>(lk, v), c = 16 * jb + lr;
            const float x = (float)acc[jb][v];
            Xs[row * XS + c] = x;
            if (!out) continue;
            M[(r0 + row) * ld + col0 + c] = x;
            if constexpr (MODE == SOP_CHOL_BF16X3) {
                const unsigned short h = f2bf(x);
                phi[(r0 + row) * NB2 + pcol0 + c] = h;
                plo[(r0 + row) * NB2 + pcol0 + c] = f2bf(x - bf2f(h));
            }
        }
}

// acc += X_g X_d^T (the full 64x64 block, 4x4 MFMA tiles) over this
// wavefront's share of the columns left of column block j-1 that no
// trailing update has applied: the previous panel's pkw columns, then this
// panel's first 64(j-1).  Those columns are cut into chunks (32 columns
// split-BF16, 16 FP32) dealt to the four wavefronts in turn; each keeps one
// chunk's operands in flight while the MFMAs of the previous one run.
// SYM (the diagonal block, g = d): the rows are loaded once and only the
// tiles on and below the diagonal are formed (the factorisation reads the
// lower triangle); each of them sums the same products in the same order.
template <int MODE, bool SYM>
__device__ __forceinline__ void mfma_split(f32x4 (&acc)[4][4], const CholStep& a, size_t rg, size_t rd) {
    const int l = threadIdx.x & 63, w = wave_of(), lr = l & 15;
    constexpr int CW = MODE == SOP_CHOL_BF16X3 ? 32 : 16;
    const int n1 = a.pkw / CW, n = n1 + (a.j >= 2 ? NB * (a.j - 1) : 0) / CW;
    if constexpr (MODE == SOP_CHOL_BF16X3) {
        const int lk = (l >> 4) * 8;
        struct Ops {
            bf16x8 ah[4], al[4], bh[4], bl[4];
        };
        auto load = [&](Ops& o, int c) {
            const unsigned short* __restrict__ ph = c < n1 ? a.pphi : a.phi;
            const unsigned short* __restrict__ pl = c < n1 ? a.pplo : a.plo;
            const int col = (c < n1 ? c : c - n1) * CW + lk;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const size_t ao = (rg + 16 * b + lr) * NB2 + col, bo = (rd + 16 * b + lr) * NB2 + col;
                o.ah[b] = *reinterpret_cast<const bf16x8*>(ph + ao);
                o.al[b] = *reinterpret_cast<const bf16x8*>(pl + ao);
                if constexpr (!SYM) {
                    o.bh[b] = *reinterpret_cast<const bf16x8*>(ph + bo);
                    o.bl[b] = *reinterpret_cast<const bf16x8*>(pl + bo);
                }
            }
        };
        auto mma = [&](const Ops& o) {
            const bf16x8(&bh)[4] = SYM ? o.ah : o.bh;
            const bf16x8(&bl)[4] = SYM ? o.al : o.bl;
#pragma unroll
            for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                for (int bj = 0; bj < 4; ++bj)
                    if (!SYM || bj <= bi)
                        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(o.ah[bi], bh[bj], acc[bi][bj], 0, 0, 0);
#pragma unroll
            for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                for (int bj = 0; bj < 4; ++bj)
                    if (!SYM || bj <= bi)
                        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(o.ah[bi], bl[bj], acc[bi][bj], 0, 0, 0);
#pragma unroll
            for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                for (int bj = 0; bj < 4; ++bj)
                    if (!SYM || bj <= bi)
                        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(o.al[bi], bh[bj], acc[bi][bj], 0, 0, 0);
        };
        Ops o0, o1;
        if (w < n) load(o0, w);
        for (int c = w; c < n; c += 8) {
            if (c + 4 < n) load(o1, c + 4);
            mma(o0);
            if (c + 4 >= n) break;
            if (c + 8 < n) load(o0, c + 8);
            mma(o1);
        }
    } else {
        // lane group g = l/16 holds k = 4g .. 4g+3 of a 16-column chunk (one
        // 16-byte load); MFMA step i takes component i
        const int lg = l >> 4;
        struct Ops {
            float4 a[4], b[4];
        };
        auto load = [&](Ops& o, int c) {
            const size_t col = (c < n1 ? (size_t)a.pc + (size_t)c * CW : (size_t)a.c0 + (size_t)(c - n1) * CW) + 4 * lg;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                o.a[b] = *reinterpret_cast<const float4*>(a.M + (rg + 16 * b + lr) * a.ld + col);
                if constexpr (!SYM) o.b[b] = *reinterpret_cast<const float4*>(a.M + (rd + 16 * b + lr) * a.ld + col);
            }
        };
        auto mma = [&](const Ops& o) {
            const float4(&ob)[4] = SYM ? o.a : o.b;
#pragma unroll
            for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                for (int bj = 0; bj < 4; ++bj)
                    if (!SYM || bj <= bi) {
                        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(o.a[bi].x, ob[bj].x, acc[bi][bj], 0, 0, 0);
                        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(o.a[bi].y, ob[bj].y, acc[bi][bj], 0, 0, 0);
                        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(o.a[bi].z, ob[bj].z, acc[bi][bj], 0, 0, 0);
                        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(o.a[bi].w, ob[bj].w, acc[bi][bj], 0, 0, 0);
                    }
        };
        Ops o0, o1;
        if (w < n) load(o0, w);
        for (int c = w; c < n; c += 8) {
            if (c + 4 < n) load(o1, c + 4);
            mma(o0);
            if (c + 4 >= n) break;
            if (c + 8 < n) load(o0, c + 8);
            mma(o1);
        }
    }
}

// The same over the 64 columns of the X blocks just solved, from LDS; in
// split-BF16 mode each value is split as the panel copy holds it.
template <int MODE>
__device__ __forceinline__ void mfma_lds(f32x4 (&acc)[4], const float* __restrict__ Xg, const float* __restrict__ Xd) {
    const int l = threadIdx.x & 63, w = wave_of(), lr = l & 15;
    if constexpr (MODE == SOP_CHOL_BF16X3) {
        const int lk = (l >> 4) * 8;
#pragma unroll
        for (int kc = 0; kc < NB; kc += 32) {
            bf16x8 ah, al, bh[4], bl[4];
            const float* a = Xg + (16 * w + lr) * XS + kc + lk;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const unsigned short h = f2bf(a[e]);
                ah[e] = (short)h;
                al[e] = (short)f2bf(a[e] - bf2f(h));
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float* b = Xd + (16 * t + lr) * XS + kc + lk;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const unsigned short h = f2bf(b[e]);
                    bh[t][e] = (short)h;
                    bl[t][e] = (short)f2bf(b[e] - bf2f(h));
                }
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[t], acc[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[t], acc[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[t], acc[t], 0, 0, 0);
        }
    } else {
        const int lg = l >> 4;  // k = kc+4g .. kc+4g+3, as in mfma_rows
#pragma unroll
        for (int kc = 0; kc < NB; kc += 16) {
            const float4 a = *reinterpret_cast<const float4*>(Xg + (16 * w + lr) * XS + kc + 4 * lg);
            float4 b[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) b[t] = *reinterpret_cast<const float4*>(Xd + (16 * t + lr) * XS + kc + 4 * lg);
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b[t].x, acc[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b[t].y, acc[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, b[t].z, acc[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, b[t].w, acc[t], 0, 0, 0);
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_chol_step(CholStep a) {
    using R = CholReal<MODE>;
    constexpr int LDT = ldt<R>();
    // SA: the block factored (R, row stride LDS_ROW), and the partial sums of
    // the update (FP32); LiT: Li^T in, then the new inverse (R, stride LDT),
    // and partial sums; both sized for the larger use
    constexpr int kSA = NB * (NB + 2);  // doubles
    __shared__ __attribute__((aligned(16))) double SA[kSA];
    __shared__ __attribute__((aligned(16))) double LiTraw[kSA];
    __shared__ __attribute__((aligned(16))) float Xg[NB * XS];
    __shared__ __attribute__((aligned(16))) float Xd[NB * XS];
    __shared__ R T[3][16][17];
    __shared__ R dinv[NB];
    R* const LiT = reinterpret_cast<R*>(LiTraw);
    static_assert(NB * ldt<R>() * sizeof(R) <= sizeof(LiTraw), "Li^T fits its buffer");
    static_assert(NB * LDS_ROW * sizeof(R) <= sizeof(SA), "the factored block fits its buffer");
    const int tid = threadIdx.x;
    const int nb = a.kw / NB, j = a.j;
    // workgroup 0: the diagonal block; 1 (when a.look): the lookahead, which
    // serves row block d+1; then the row blocks d+1, d+2, ...
    const bool look = a.look && blockIdx.x == 1;
    const int x = look ? 1 : (int)blockIdx.x - (a.look && blockIdx.x > 1 ? 1 : 0);
    const size_t rd = (size_t)a.c0 + (size_t)NB * j;  // first row of the diagonal block d
    const size_t rg = rd + (size_t)NB * x;             // first row of this workgroup's block g
    const bool update = j < nb;
    const bool diag_p = x == 0 && j >= 1;              // the diagonal block arrives as P
    const int l = tid & 63, w = wave_of();
    // the block this workgroup updates: column block j (rows rg), or for the
    // lookahead the next diagonal block (rows and columns rg)
    const size_t cj = look ? rg : rd;
    // every global read first, so their latencies overlap: this row block of
    // column block j (or P), the B rows of the solves, the inverse Li^T
    float mblk[4][4], bg[16], bd[16];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = 16 * w + 4 * (l >> 4) + q, c = 16 * t + (l & 15);
            mblk[t][q] = !update ? 0.f : diag_p ? a.p_in[row * NB + c] : a.M[(rg + row) * a.ld + cj + c];
        }
    if (j >= 1) {
        load_solve_rows(a.b_in, rg, bg);
        if (update && x != 0 && !look) load_solve_rows(a.b_in, rd, bd);
        {  // Li^T: 16-byte loads, all in flight before the LDS stores
            using V = typename std::conditional<std::is_same<R, double>::value, double2, float4>::type;
            constexpr int EPV = 16 / sizeof(R), NV = NB * NB / (256 * EPV);
            const V* __restrict__ src = reinterpret_cast<const V*>(a.li_in);
            V lv[NV];
#pragma unroll
            for (int k = 0; k < NV; ++k) lv[k] = src[tid + 256 * k];
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                const int e = EPV * (tid + 256 * k);  // e .. e+EPV-1 in the same row
                *reinterpret_cast<V*>(LiT + (e >> 6) * LDT + (e & 63)) = lv[k];
            }
        }
        __syncthreads();
        const int pcol = NB * (j - 1);  // column block j-1 within the panel
        if (update && x != 0 && !look)  // X_d, for the update only
            solve_block<MODE>(bd, LiT, Xd, false, a.M, a.ld, rd, a.c0 + pcol, a.phi, a.plo, pcol);
        // the lookahead's rows are written by the workgroup that serves them
        solve_block<MODE>(bg, LiT, Xg, !look, a.M, a.ld, rg, a.c0 + pcol, a.phi, a.plo, pcol);
    }
    if (!update) return;
    __syncthreads();  // X blocks complete; SA and LiT are free
    f32x4 acc[4];
    if (diag_p) {
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    } else {
        // the columns left of j-1, K split over the wavefronts; the four
        // partial blocks are summed through LDS in wavefront order
        f32x4 part[4][4];
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
            for (int bj = 0; bj < 4; ++bj) part[bi][bj] = f32x4{0.f, 0.f, 0.f, 0.f};
        if (x == 0 || look)
            mfma_split<MODE, true>(part, a, rg, rg);
        else
            mfma_split<MODE, false>(part, a, rg, rd);
        constexpr int PS = NB + 1;
        float* const P = (w < 2 ? reinterpret_cast<float*>(SA) : reinterpret_cast<float*>(LiTraw)) + (w & 1) * NB * PS;
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
            for (int bj = 0; bj < 4; ++bj)
#pragma unroll
                for (int q = 0; q < 4; ++q) P[(16 * bi + 4 * (l >> 4) + q) * PS + 16 * bj + (l & 15)] = part[bi][bj][q];
        __syncthreads();
        const float* const P0 = reinterpret_cast<const float*>(SA);
        const float* const P2 = reinterpret_cast<const float*>(LiTraw);
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = (16 * w + 4 * (l >> 4) + q) * PS + 16 * t + (l & 15);
                acc[t][q] = ((P0[e] + P0[NB * PS + e]) + P2[e]) + P2[NB * PS + e];
            }
        __syncthreads();  // SA and LiT are reused below
    }
    if (j >= 1) mfma_lds<MODE>(acc, Xg, x == 0 || look ? Xg : Xd);
    R(*S)[LDS_ROW] = reinterpret_cast<R(*)[LDS_ROW]>(SA);
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = 16 * w + 4 * (l >> 4) + q, c = 16 * t + (l & 15);
            const float b = mblk[t][q] - acc[t][q];
            if (look)
                a.p_out[row * NB + c] = b;
            else if (x == 0)
                S[row][c] = (R)b;
            else
                a.b_out[(rg + row) * NB + c] = b;
        }
    if (x != 0 || look) return;
    __syncthreads();
    potrf_inv64(S, LiT, T, dinv);
#pragma unroll
    for (int k = 0; k < NB * NB / 256; ++k) {
        const int e = tid + 256 * k, r = e / NB, c = e % NB;
        if (c <= r) a.M[(rd + r) * a.ld + rd + c] = (float)S[r][c];
    }
    {  // 16-byte LDS reads, then the stores
        using V = typename std::conditional<std::is_same<R, double>::value, double2, float4>::type;
        constexpr int EPV = 16 / sizeof(R), NV = NB * NB / (256 * EPV);
        V lv[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const int e = EPV * (tid + 256 * k);
            lv[k] = *reinterpret_cast<const V*>(LiT + (e >> 6) * LDT + (e & 63));
        }
        V* __restrict__ dst = reinterpret_cast<V*>(a.li_out);
#pragma unroll
        for (int k = 0; k < NV; ++k) dst[tid + 256 * k] = lv[k];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

hipError_t launch_hi_tea_update(const DevJob* jobs, DevJobState* st, const int2* tiles, int ntiles, int njobs,
                                hipStream_t stream) {
    if (ntiles > 0)
        hipLaunchKernelGGL(k_hi_eps, dim3(ntiles * kHiSplit), dim3(1024), 0, stream, jobs, st, tiles, ntiles);
    hipLaunchKernelGGL(k_hi_beta, dim3(njobs), dim3(64), 0, stream, jobs, st, njobs);
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
// Regions whose 256x256 tiles would leave most CUs idle take 128x128 tiles.
static constexpr int kTrailBigMinTiles = 256;

static void trailing_update(float* M, size_t ld, int s, int mp, int pc, int kw, int mode, const unsigned short* phi,
                            const unsigned short* plo, hipStream_t stream) {
    if (s >= mp) return;
    const int nr = (mp - s + GT - 1) / GT;
    const dim3 g(nr, nr);
    const int ntr = (mp - s + BT - 1) / BT, ntile = ntr * (ntr + 1) / 2;
    if (mode == SOP_CHOL_F32)
        hipLaunchKernelGGL(k_gemm_f32, g, dim3(256), 0, stream, M, ld, s, mp, pc, kw);
    else if (ntile >= kTrailBigMinTiles)
        hipLaunchKernelGGL(k_trail_bf16x3, dim3((ntile + kXcds - 1) / kXcds * kXcds), dim3(512), 0, stream, M, ld, s,
                           mp, kw, ntr, phi, plo);
    else
        hipLaunchKernelGGL(k_gemm_bf16x3, g, dim3(256), 0, stream, M, ld, s, mp, kw, phi, plo);
}

// Blocked factorisation over outer panels of NB2 columns, each factored by
// fused 64-column steps (k_chol_step).  A panel's columns receive the
// previous panel's update inside its own steps; the rest of the trailing
// matrix (from the panel after the next one on) is updated by the wide MFMA
// kernels on `aux`, overlapped with the next panel.  Panel p waits for the
// trailing update of panel p-2, which wrote its columns and read the copy
// buffer panel p writes.  With aux == nullptr everything runs on `stream`.
hipError_t launch_chol_factor(float* M, size_t ld, int mp, int mode, const CholWork& w, hipStream_t stream,
                              hipStream_t aux, hipEvent_t ev_f, hipEvent_t ev_t2a, hipEvent_t ev_t2b) {
    const bool two = aux != nullptr && ev_f != nullptr && ev_t2a != nullptr && ev_t2b != nullptr;
    hipEvent_t ev_t2[2] = {ev_t2a, ev_t2b};
    bool pending[2] = {false, false};
    hipError_t e = hipSuccess;
    const int nblk = mp / NB;
    float* const pbuf = reinterpret_cast<float*>(w.linv + (size_t)2 * NB * NB);  // after the two inverses
    int pc = 0, pkw = 0;
    for (int c0 = 0, p = 0; c0 < mp; c0 += NB2, ++p) {
        const int kw = mp - c0 < NB2 ? mp - c0 : NB2, nb = kw / NB;
        unsigned short* const ph = w.phi + (size_t)(p & 1) * mp * NB2;
        unsigned short* const pl = w.plo + (size_t)(p & 1) * mp * NB2;
        const unsigned short* const qh = w.phi + (size_t)((p + 1) & 1) * mp * NB2;
        const unsigned short* const ql = w.plo + (size_t)((p + 1) & 1) * mp * NB2;
        if (two && pending[p & 1]) {
            if ((e = hipStreamWaitEvent(stream, ev_t2[p & 1], 0)) != hipSuccess) return e;
            pending[p & 1] = false;
        }
        for (int j = 0; j <= nb; ++j) {
            const int nwg = nblk - (c0 / NB + j);
            if (nwg <= 0) continue;
            const int look = j + 1 < nb ? 1 : 0;  // a next diagonal block in this panel
            const CholStep a{M, ld, mp, c0, kw, pc, pkw, j,
                             w.linv + (size_t)((j + 1) & 1) * NB * NB, w.linv + (size_t)(j & 1) * NB * NB,
                             w.bbuf + (size_t)((j + 1) & 1) * mp * NB, w.bbuf + (size_t)(j & 1) * mp * NB,
                             pbuf + (size_t)(j & 1) * NB * NB, pbuf + (size_t)((j + 1) & 1) * NB * NB, look,
                             ph, pl, qh, ql};
            if (mode == SOP_CHOL_F32)
                hipLaunchKernelGGL(k_chol_step<SOP_CHOL_F32>, dim3(nwg + look), dim3(256), 0, stream, a);
            else
                hipLaunchKernelGGL(k_chol_step<SOP_CHOL_BF16X3>, dim3(nwg + look), dim3(256), 0, stream, a);
        }
        const int s1 = c0 + kw + NB2;  // the first column the next panel does not fold in
        if (s1 < mp) {
            if (two) {
                if ((e = hipEventRecord(ev_f, stream)) != hipSuccess) return e;
                if ((e = hipStreamWaitEvent(aux, ev_f, 0)) != hipSuccess) return e;
                trailing_update(M, ld, s1, mp, c0, kw, mode, ph, pl, aux);
                if ((e = hipEventRecord(ev_t2[p & 1], aux)) != hipSuccess) return e;
                pending[p & 1] = true;
            } else {
                trailing_update(M, ld, s1, mp, c0, kw, mode, ph, pl, stream);
            }
        }
        pc = c0;
        pkw = kw;
    }
    for (int q = 0; q < 2; ++q)
        if (pending[q] && (e = hipStreamWaitEvent(stream, ev_t2[q], 0)) != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t launch_chol_step(const DevJob* jobs, const DevJobState* st, const int2* tiles, int ntiles, int kstep,
                            const HiExactWork* w, hipStream_t stream) {
    hipLaunchKernelGGL(k_chol_D, dim3(ntiles * kHiSplit), dim3(1024), 0, stream, jobs, st, tiles, ntiles, kstep);
    if (w) {
        hipLaunchKernelGGL(k_chol_build, dim3(ntiles, ntiles), dim3(256), 0, stream, jobs, st, tiles, kstep);
        // the factorisation's error covers the event record/wait pairs that
        // order its two streams: report it before the noise reads the factor
        const hipError_t e =
            launch_chol_factor(w->M, w->ld, w->mp, w->mode, w->work, stream, w->aux, w->ev_f, w->ev_t2[0], w->ev_t2[1]);
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
constexpr size_t kNoAllocBBytes = (size_t)2 * kNoAllocMaxOrder * 64 * sizeof(float);  // chol_bbuf_bytes
constexpr size_t kNoAllocPanelBytes = (size_t)2 * kNoAllocMaxOrder * kCholPanel * sizeof(unsigned short);
constexpr size_t kNoAllocScratchBytes = kCholLinvBytes + kNoAllocBBytes + 2 * kNoAllocPanelBytes;  // linv|bbuf|phi|plo
static_assert(kCholLinvBytes % 256 == 0 && kNoAllocBBytes % 256 == 0 && kNoAllocPanelBytes % 256 == 0,
              "scratch sections stay 256-byte aligned");

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
    if (ld % 4 != 0 || reinterpret_cast<uintptr_t>(M) % 16 != 0)
        return fail(SOP_E_ARG, "cholesky (no allocation): rows must start on 16-byte boundaries "
                               "(M 16-byte aligned, leading dimension a multiple of 4)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(SOP_E_HIP, "no HIP device available");
    if (device < 0 || device >= ndev || device >= kMaxDevices) return fail(SOP_E_ARG, "device index out of range");
    if (hipSetDevice(device) != hipSuccess) return fail(SOP_E_HIP, "hipSetDevice failed");
    unsigned char* base = nullptr;
    hipError_t e = scratch_on(device, &base);
    if (e != hipSuccess)
        return fail(SOP_E_HIP, std::string("cholesky (no allocation): scratch address: ") + hipGetErrorString(e));
    // the workspace of an order-mp factorisation, carved from the front of each section
    CholWork w{};
    w.linv = reinterpret_cast<double*>(base);
    w.bbuf = reinterpret_cast<float*>(base + kCholLinvBytes);
    w.phi = reinterpret_cast<unsigned short*>(base + kCholLinvBytes + kNoAllocBBytes);
    w.plo = reinterpret_cast<unsigned short*>(base + kCholLinvBytes + kNoAllocBBytes + kNoAllocPanelBytes);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (mp > n) hipLaunchKernelGGL(k_chol_pad, dim3((mp + 255) / 256, mp - n), dim3(256), 0, s, M, (size_t)ld, n, mp);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_chol_factor(M, (size_t)ld, mp, mode, w, s);
    if (e != hipSuccess) return fail(SOP_E_HIP, std::string("cholesky (no allocation): ") + hipGetErrorString(e));
    return 0;
}
