// kernels.hip -- CDNA4 kernels of the SOP Langevin hot path.
//
// k_step, k_step_lds
//              one launch per step for every admitted job: FENE + native
//              LJ 12-6 + repulsive LJ (+ indentation walls) gathered per bead
//              over slot-major lists, then the Hybrid-Taus noise, the
//              fixed/pulled-bead rules and the overdamped Euler update; the
//              LDS variant stages the coordinates of jobs of at most 5,120
//              beads in LDS.  Reference bodies: src/Potentials/covalent_kernel.cu:28-51,
//              native_kernel.cu:27-52, pairs_kernel.cu:33-59,
//              indentation_kernel.cu:71-128, src/Integrators/langevin_kernel.cu:17-58.
// k_possible, k_possible_lds, k_possible_cells
//              possible-pairs list: same-job beads j != i, ascending j, with
//              sqrtf(d^2) < threshold, excluding covalent and native partners
//              (src/Updaters/possiblepairlist_kernel.cu:18-71).
// k_sort_rows  row order of the Verlet lists (by list length) and the
//              bonded lists moved into it.
// k_verlet, k_verlet_lds
//              Verlet list: the possible list filtered by sqrtf(d^2) <
//              pairlist_cutoff in stored order (src/Updaters/pairlist_kernel.cu:33-76).
// k_job_update per-job updaters at stride boundaries: indentation tip,
//              force-ramp cantilever, heating schedule
//              (src/Potentials/indentation.cu:414-511, pulling.cu:152-214,
//              src/Integrators/langevin.cu:76-82).
// k_energies   per-bead energies for the output cadence
//              (covalent_kernel.cu:57-77, native_kernel.cu:57-82, pairs_kernel.cu:69-91).
//
// Work is distributed as 64-bead tiles (one wavefront each) that never cross
// a job boundary, so every wavefront reads one job's descriptor through
// scalar loads.  Tiles are assigned to XCDs in contiguous ranges: workgroups
// are dispatched round-robin over the 8 XCDs, so block b runs on XCD b%8, and
// giving XCD x the x-th slice of the tile list keeps each job's coordinates in
// one XCD's L2.
#include "kernels.h"

#include <climits>

#include "devmath.h"

namespace sop {

static constexpr int kTilesPerBlock = 4;  // 256-thread workgroups


__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

struct TileRange {
    int begin, end, step;
};

// Contiguous slice of the tile list for this workgroup's XCD.
__device__ __forceinline__ TileRange xcd_tiles(int ntiles) {
    const int nb = gridDim.x;
    const int nxcd = nb >= 8 ? 8 : 1;
    const int xcd = blockIdx.x % nxcd;
    const int r = blockIdx.x / nxcd;
    const int per_xcd_blocks = nb / nxcd;
    const int slice = (ntiles + nxcd - 1) / nxcd;
    TileRange t;
    t.begin = xcd * slice + r * kTilesPerBlock + wave_id();
    t.end = min((xcd + 1) * slice, ntiles);
    t.step = per_xcd_blocks * kTilesPerBlock;
    return t;
}

// The descriptor's arrays as global-memory pointers.  Through a generic
// pointer the compiler issues flat loads, which are counted with the LDS
// operations as well, so a wait for an LDS read also waits for every
// prefetch in flight; global loads are counted apart from the LDS reads.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gmem(const T* p) {
    return (const __attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}

// Coordinates gathered from global memory with global-address-space loads
// (cin[j]), for the kernels that gather from L2: through a generic pointer
// the gathers are flat loads, which count with the scalar loads too, so a
// wait for a descriptor field also drained every gather in flight.
typedef float f32x4_g __attribute__((ext_vector_type(4)));
typedef float f32x2_g __attribute__((ext_vector_type(2)));
// p[k] of a float2 array in global memory, by a global-address-space load.
__device__ __forceinline__ float2 ldg2(const float2* p, size_t k) {
    const f32x2_g v = ((const __attribute__((address_space(1))) f32x2_g*)p)[k];
    return make_float2(v.x, v.y);
}
struct CoordGlobal {
    const float4* p;
    __device__ __forceinline__ float4 operator[](int j) const {
        const f32x4_g v = ((const __attribute__((address_space(1))) f32x4_g*)p)[j];
        return make_float4(v.x, v.y, v.z, v.w);
    }
};
// The global-gather kernels also read the per-bead counts and the native
// parameters by global-address-space loads; the LDS kernels keep the
// generic loads (the global form costs them registers).
template <typename CP>
constexpr bool kGlobalCoords = std::is_same<CP, CoordGlobal>::value;

// ---------------------------------------------------------------------------
// fused step
// ---------------------------------------------------------------------------

// The scalar factors of the three force terms, shared by every step
// kernel so that each evaluates the same float expressions.
// Repulsive LJ, for d2 < rc2 (pairs_kernel.cu:41-56): the force on the bead
// is (x, y, z) * q; `fast` (the job's a2 and rc2 admit the short division)
// takes div_rn_fast for d2 >= kDivFastMinB.
__device__ __forceinline__ float pair_q(float d2, float a2, float m6, bool fast) {
    float q = div_rn_fast(a2, d2);
    if (!(fast && d2 >= kDivFastMinB)) q = div_ieee(a2, d2);
    q = q * q;
    return m6 * q * q;
}
// FENE (covalent_kernel.cu:36-48): the force is w * (x, y, z).
__device__ __forceinline__ float fene_w(float ks, float R2, float x, float y, float z, float r0) {
    float w = sqrtf(x * x + y * y + z * z);
    w = w - r0;
    return ks * w / ((w + r0) * (1.0f - w * w / R2));
}
// Native LJ 12-6 (native_kernel.cu:36-49), par = (r0^2, -12 eh / r0^2): the
// force is (x, y, z) * q.
__device__ __forceinline__ float native_q(float x, float y, float z, float2 par) {
    float q = par.x / (x * x + y * y + z * z);
    const float p6 = q * q * q;
    return q * (p6 * p6 - p6) * par.y;
}

// Repulsive LJ over a bead's slot-major Verlet entries, in list order.  The
// list words are streamed with non-temporal loads (each is used once per
// step) and fetched one batch of PB ahead: while batch b is computed the
// indices of batch b+1 and the coordinates of batch b are in flight.  PB is
// 8 for gathers from L2 and 6 for gathers from LDS (which keeps the LDS
// kernel within the 64 VGPRs of two workgroups per CU).
// TAILB: the last np % PB entries are loaded as one predicated batch
// instead of one after another (gathers from L2, few wavefronts to hide
// the latency of each).
template <int PB, typename IdxT, bool TAILB = false, typename CP = const float4*>
__device__ __forceinline__ void pair_loop(const __attribute__((address_space(1))) IdxT* __restrict__ pj, int np, int n, CP cin,
                                          const float4 c, float rc2, float a2, float m6, float& fx, float& fy,
                                          float& fz) {
    // every d2 below rc2 and above kDivFastMinB takes the short division
    const bool fast = div_fast_numerator(a2) && rc2 <= kDivFastMaxB;
#define SOP_PAIR(o)                                                  \
    {                                                                \
        const float x = (o).x - c.x, y = (o).y - c.y, z = (o).z - c.z; \
        const float d2 = x * x + y * y + z * z;                      \
        if (d2 < rc2) {                                              \
            const float q = pair_q(d2, a2, m6, fast);                \
            fx += x * q;                                             \
            fy += y * q;                                             \
            fz += z * q;                                             \
        }                                                            \
    }
    int k = 0;
    if (np >= PB) {
        int jn[PB];
#pragma unroll
        for (int u = 0; u < PB; ++u) jn[u] = (int)__builtin_nontemporal_load(pj + (size_t)u * n);
        for (; k + PB <= np; k += PB) {
            float4 o[PB];
#pragma unroll
            for (int u = 0; u < PB; ++u) o[u] = cin[jn[u]];
            if (k + 2 * PB <= np) {
#pragma unroll
                for (int u = 0; u < PB; ++u) jn[u] = (int)__builtin_nontemporal_load(pj + (size_t)(k + PB + u) * n);
            }
#pragma unroll
            for (int u = 0; u < PB; ++u) SOP_PAIR(o[u]);
        }
    }
    if constexpr (TAILB) {
        if (k < np) {
            int jt[PB];
#pragma unroll
            for (int u = 0; u < PB; ++u)
                jt[u] = k + u < np ? (int)__builtin_nontemporal_load(pj + (size_t)(k + u) * n) : 0;
            float4 ot[PB];
#pragma unroll
            for (int u = 0; u < PB; ++u) ot[u] = cin[jt[u]];
#pragma unroll
            for (int u = 0; u < PB; ++u)
                if (k + u < np) SOP_PAIR(ot[u]);
        }
    } else {
        for (; k < np; ++k) {
            const float4 o = cin[(int)__builtin_nontemporal_load(pj + (size_t)k * n)];
            SOP_PAIR(o);
        }
    }
#undef SOP_PAIR
}
// The fixed/pulled rules of the reference's extForces (pulling.cu:108-125,
// langevin_kernel.cu:43-50): a fixed bead's force is zeroed, a pulled bead's
// gets the cantilever force.  The one-byte flag is read for every bead, the
// 16-byte force record only for pulled beads.
__device__ __forceinline__ void ext_rules_flag(const DevJob& J, int i, int f, float& fx, float& fy, float& fz) {
    if (f == 1) {
        fx = 0.0f; fy = 0.0f; fz = 0.0f;
    } else if (f == 2) {
        const float4 e = J.ext[i];
        fx += e.x; fy += e.y; fz += e.z;
    }
}
__device__ __forceinline__ void ext_rules(const DevJob& J, int i, float& fx, float& fy, float& fz) {
    ext_rules_flag(J, i, J.bflag[i], fx, fy, fz);
}

// FENE (covalent_kernel.cu:36-48) and native LJ 12-6 (native_kernel.cu:36-49)
// of one bead: nc covalent and nn native partners from slot-major lists of
// slot stride n whose columns for this bead start at cj, cr, nj, npar
// (32- or 16-bit indices; bead order, or row order for sorted rows).
template <typename CJ, typename CR, typename NJ, typename CP>
__device__ __forceinline__ void bonded_terms(const DevJob& J, int nc, int nn, int n, CJ cj, CR cr, NJ nj,
                                             const float2* __restrict__ npar, CP cin,
                                             const float4 c, float& fx, float& fy, float& fz) {
    // FENE (covalent_kernel.cu:36-48)
    {
        for (int k = 0; k < nc; ++k) {
            const int j = cj[(size_t)k * n];
            const float r0 = cr[(size_t)k * n];
            const float4 o = cin[j];
            const float x = o.x - c.x, y = o.y - c.y, z = o.z - c.z;
            const float w = fene_w(J.ks, J.R2, x, y, z, r0);
            fx += w * x;
            fy += w * y;
            fz += w * z;
        }
    }
    // native LJ 12-6 (native_kernel.cu:36-49), gathers four ahead
    {
#define SOP_NAT(o, par)                                                 \
{                                                                   \
    const float x = (o).x - c.x, y = (o).y - c.y, z = (o).z - c.z;     \
    const float q = native_q(x, y, z, par);                         \
    fx += x * q;                                                    \
    fy += y * q;                                                    \
    fz += z * q;                                                    \
}
        int k = 0;
        for (; k + 4 <= nn; k += 4) {
            int jj[4];
            float2 pp[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                jj[u] = nj[(size_t)(k + u) * n];
                pp[u] = kGlobalCoords<CP> ? ldg2(npar, (size_t)(k + u) * n) : npar[(size_t)(k + u) * n];
            }
            float4 o[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) o[u] = cin[jj[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u) SOP_NAT(o[u], pp[u]);
        }
        for (; k < nn; ++k) {
            const float2 par = kGlobalCoords<CP> ? ldg2(npar, (size_t)k * n) : npar[(size_t)k * n];
            const float4 o = cin[nj[(size_t)k * n]];
            SOP_NAT(o, par);
        }
#undef SOP_NAT
    }
}

// The first part of a bead's step, before the repulsive pairs: the extra
// force record, then FENE and native LJ 12-6 (for sorted rows from the
// row-ordered copies -- a wavefront's 64 rows read 64 consecutive words per
// slot -- else in bead order, partner indices 16-bit where the job has a
// 16-bit copy).  i: bead, r: its Verlet row; c = cin[i].
template <bool FO, typename CP>
__device__ __forceinline__ void step_head(const DevJob& J, int i, int r, int n, CP cin, const float4 c, float& fx,
                                          float& fy, float& fz) {
    fx = 0.0f, fy = 0.0f, fz = 0.0f;
    if (J.flags & JF_EXTRA_FORCE) {
        const float4 e = J.fextra[i];
        fx = e.x; fy = e.y; fz = e.z;
        if constexpr (!FO) J.fextra[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (J.sort_rows) {
        const int cnt = gmem(J.rb_cnt)[r];
        bonded_terms(J, cnt & 0xffff, cnt >> 16, J.pair_ld, gmem(J.rb_cov_j) + r, gmem(J.rb_cov_r0) + r,
                     gmem(J.rb_nat_j) + r, J.rb_nat_par + r, cin, c, fx, fy, fz);
    } else if (J.nat_j16) {
        bonded_terms(J, kGlobalCoords<CP> ? gmem(J.cov_cnt)[i] : J.cov_cnt[i],
                     kGlobalCoords<CP> ? gmem(J.nat_cnt)[i] : J.nat_cnt[i], n, gmem(J.cov_j16) + i, gmem(J.cov_r0) + i,
                     gmem(J.nat_j16) + i, J.nat_par + i, cin, c, fx, fy, fz);
    } else {
        bonded_terms(J, kGlobalCoords<CP> ? gmem(J.cov_cnt)[i] : J.cov_cnt[i],
                     kGlobalCoords<CP> ? gmem(J.nat_cnt)[i] : J.nat_cnt[i], n, gmem(J.cov_j) + i, gmem(J.cov_r0) + i,
                     gmem(J.nat_j) + i, J.nat_par + i, cin, c, fx, fy, fz);
    }
}

// The per-bead state the tail reads (RNG state, temperature accumulator,
// fixed/pulled flag), loaded ahead by a caller with latency to hide.
struct TailPre {
    uint4 rng;
    float T;
    int flag;
};
__device__ __forceinline__ TailPre tail_prefetch(const DevJob& J, int i) {
    TailPre t;
    t.rng = J.rng[i];
    t.T = J.T[i];
    t.flag = (J.flags & JF_PULLING) ? (int)J.bflag[i] : 0;
    return t;
}

// The rest of a bead's step, after the repulsive pairs: indentation walls,
// then (FO) the forces written to cout, or the noise, the fixed/pulled rules
// and the Euler update written to `cout` in global memory.  Hydrodynamic
// jobs stop after the forces.  pre: the state loaded ahead (null: loaded
// here).
template <bool FO>
__device__ __forceinline__ void step_tail(const DevJob& J, const DevJobState& S, long long s, int i, const float4 c,
                                          float fx, float fy, float fz, float4* __restrict__ cout,
                                          const TailPre* pre = nullptr) {
    // indentation tip and substrate plane (indentation_kernel.cu:71-128)
    if (J.flags & JF_INDENTATION) {
        float4 fin = J.tipF[i];
        float x = c.x - S.tip[0], y = c.y - S.tip[1], z = c.z - S.tip[2];
        const float rr = sqrtf(x * x + y * y + z * z);
        float r2 = rr - J.tipR;
        r2 = 1.0f / r2;
        r2 = r2 * r2;
        float r6 = r2 * r2 * r2;
        float dw = J.tipAp * r6 + J.tipBp;
        dw = dw * r6 / (rr * (rr - J.tipR));
        const float dfx = x * dw, dfy = y * dw, dfz = z * dw;
        fin.x += dfx;
        fin.y += dfy;
        fin.z += dfz;
        if constexpr (!FO) J.tipF[i] = fin;
        x = c.x - S.micaR[0];
        y = c.y - S.micaR[1];
        z = c.z - S.micaR[2];
        float h = x * J.micaN[0] + y * J.micaN[1] + z * J.micaN[2];
        h = 1.0f / h;
        h = h * h;
        r6 = h * h * h;
        float sw = J.surfAp * r6 + J.surfBp;
        sw = sw * r6 * h;
        fx += J.micaN[0] * sw + dfx;
        fy += J.micaN[1] * sw + dfy;
        fz += J.micaN[2] * sw + dfz;
    }
    if constexpr (FO) {
        cout[i] = make_float4(fx, fy, fz, 0.0f);
        return;
    }
    // hydrodynamic jobs: noise and forces for the HI integrator
    // (integrateTea_prepare, bdhitea_kernel.cu:33-65); the update itself
    // runs in hi.hip once every bead's forces are known
    if (J.flags & JF_HI) {
        uint4 rs = pre ? pre->rng : J.rng[i];
        const float var = S.var;
        const float3 g = gauss3(rs);
        J.rng[i] = rs;
        J.rforce[i] = make_float4(g.x * var, g.y * var, g.z * var, 0.0f);
        if (J.flags & JF_PULLING) ext_rules_flag(J, i, pre ? pre->flag : (int)J.bflag[i], fx, fy, fz);
        J.mforce[i] = make_float4(fx, fy, fz, 0.0f);
        return;
    }
    // Langevin step (langevin_kernel.cu:24-56)
    if (!(J.flags & JF_MINIMIZATION)) {
        uint4 rs = pre ? pre->rng : J.rng[i];
        const float var = S.var;
        const float3 g = gauss3(rs);
        J.rng[i] = rs;
        fx += g.x * var;
        fy += g.y * var;
        fz += g.z * var;
    }
    if (J.flags & JF_PULLING) ext_rules_flag(J, i, pre ? pre->flag : (int)J.bflag[i], fx, fy, fz);
    const float m = J.hOverZeta;
    const float dx = m * fx, dy = m * fy, dz = m * fz;
    float4 cn = c;
    cn.x += dx;
    cn.y += dy;
    cn.z += dz;
    // the temperature accumulator is cleared at updater passes that fall
    // on a multiple of nav (output_manager.cu:124-126)
    const bool reset = (s % J.stride == 0) && (s % J.nav == 0);
    const float T0 = reset ? 0.0f : (pre ? pre->T : J.T[i]);
    J.T[i] = T0 + J.tempNorm * (dx * dx + dy * dy + dz * dz);
    cout[i] = cn;
}

// One bead's step: FENE, native and repulsive terms gathered from `cin`
// (global memory, or the job's coordinates staged in LDS), indentation
// walls, then noise, the fixed/pulled rules and the Euler update, written to
// `cout` in global memory.  Hydrodynamic jobs stop after the forces.
// FO (forces only): the same force terms, written to `cout` (indexed by
// bead) before the noise, with no state changed -- the parity hook of the
// force gather runs the kernel the step runs.
template <int PB, bool FO = false, bool TAILB = false, typename CP = const float4*>
__device__ __forceinline__ void step_bead(const DevJob& J, const DevJobState& S, long long s, int i, int r, int n,
                                          CP cin, float4* __restrict__ cout) {
    const float4 c = cin[i];
    float fx, fy, fz;
    step_head<FO>(J, i, r, n, cin, c, fx, fy, fz);
    // repulsive LJ over the Verlet list (pairs_kernel.cu:41-56)
    const int np = gmem(J.pair_cnt)[r];
    if (J.idx16)
        pair_loop<PB, unsigned short, TAILB>(gmem(static_cast<const unsigned short*>(J.pair_j)) + r, np, J.pair_ld,
                                             cin, c, J.rc2, J.a2, J.m6, fx, fy, fz);
    else
        pair_loop<PB, int, TAILB>(gmem(static_cast<const int*>(J.pair_j)) + r, np, J.pair_ld, cin, c, J.rc2, J.a2,
                                  J.m6, fx, fy, fz);
    step_tail<FO>(J, S, s, i, c, fx, fy, fz, cout);
}

// FO: forces only, into fout (bead-indexed), for every job of the tiles
// whatever its step (nothing is advanced).
template <bool FO>
__global__ __launch_bounds__(256) void k_step(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                              const int2* __restrict__ tiles, int ntiles, int kstep,
                                              float4* __restrict__ fout) {
    const int lane = threadIdx.x & 63;
    const TileRange tr = xcd_tiles(ntiles);
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int2 td = tiles[t];
        const DevJob& J = jobs[td.x];
        if (J.lds_step) continue;  // stepped by k_step_lds
        const long long s = st[td.x].step + kstep;
        if (!FO && s >= J.end_step) continue;
        const int r = td.y + lane;  // Verlet row; the lane steps bead perm[r]
        const int n = J.n;
        if (r >= n) continue;
        step_bead<8, FO, true>(J, st[td.x], s, gmem(J.perm)[r], r, n, CoordGlobal{J.coord[s & 1]},
                               FO ? fout : J.coord[(s + 1) & 1]);
    }
}

// The same step for jobs small enough to keep their coordinates in LDS.
// One 1024-thread workgroup per unit; a unit is up to 1024 Verlet rows in
// segments (job, first row, rows, LDS offset): a job's whole 1024-row
// blocks are units of their own, and the rows left over share units with
// other jobs' left-over rows.  The workgroup stages each segment's job
// coordinates in LDS (coalesced), so every neighbour gather of the step
// reads LDS.  Segments start at multiples of 64 threads: no wavefront spans
// two jobs.
// PB: Verlet entries fetched ahead per lane (pair_loop).  With PB <= 6 the
// kernel is built for two 1024-thread workgroups per CU (64 VGPRs), with
// PB = 8 for one (128 VGPRs); two 80 KB LDS stagings fit a CU either way.
template <bool FO, int PB>
__global__ __launch_bounds__(1024, PB <= 6 ? 8 : 4) void k_step_lds(const DevJob* __restrict__ jobs,
                                                                 const DevJobState* __restrict__ st,
                                                                 const int2* __restrict__ units,
                                                                 const int4* __restrict__ segs, int nunits, int kstep,
                                                                 float4* __restrict__ fout) {
    extern __shared__ float4 lds_coord[];
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int2 ud = units[u];  // first segment, segments
        __syncthreads();           // the previous unit's gathers are done
        for (int q = 0; q < ud.y; ++q) {
            const int4 sg = segs[ud.x + q];
            const DevJob& J = jobs[sg.x];
            const long long s = st[sg.x].step + kstep;
            if (!FO && s >= J.end_step) continue;
            const float4* __restrict__ g = J.coord[s & 1];
            for (int k = threadIdx.x; k < J.n; k += blockDim.x) lds_coord[sg.w + k] = g[k];
        }
        __syncthreads();
        int t = threadIdx.x;
        for (int q = 0; q < ud.y; ++q) {
            const int4 sg = segs[ud.x + q];
            const int span = (sg.z + 63) & ~63;
            if (t < span) {
                const DevJob& J = jobs[sg.x];
                const long long s = st[sg.x].step + kstep;
                const int r = sg.y + t;  // Verlet row; the thread steps bead perm[r]
                if (t < sg.z && (FO || s < J.end_step))
                    step_bead<PB, FO>(J, st[sg.x], s, J.perm[r], r, J.n, lds_coord + sg.w,
                                      FO ? fout : J.coord[(s + 1) & 1]);
                break;
            }
            t -= span;
        }
    }
}

// The step with each bead's force terms spread over a workgroup, for jobs
// whose rows are too few to fill the GPU one row per lane (a single chain of
// a thousand beads fills 16 wavefronts).  One 1024-thread workgroup steps R
// Verlet rows of an LDS-step job, with the job's coordinates staged in LDS.
// A row's terms form one stream in the order the one-lane step adds them:
// FENE (nc), native (nn), then the Verlet entries (np).  Thirteen
// wavefronts evaluate the stream in chunks of kSplitK(R) positions (a lane
// per row and position, 64 / R positions of R rows per wavefront), each
// term written to LDS as its three components (-0.0 where a pair lies
// beyond the cut-off or the row has no term at that position: x + (-0.0)
// is x for every float x).  Three owner wavefronts, one per component, add
// a chunk's terms in stream order while the next chunk is evaluated
// (double buffer), starting from the extra force record.  The sums are the
// one-lane step's sums, rounded the same way term by term, so the step is
// bit-identical to k_step_lds.  Wavefront 0 then runs the walls, noise and
// update for the R rows.
constexpr int kSplitOwners = 3;
constexpr int kSplitWorkers = 16 - kSplitOwners;
template <int R>
constexpr int split_k() { return kSplitWorkers * (64 / R); }
// row stride of the term buffers: K, or K + 4 where K / 4 is even, so that
// the owners' 16-byte reads of R rows fall in distinct bank groups
template <int R>
constexpr int split_kp() { return (split_k<R>() / 4) % 2 ? split_k<R>() : split_k<R>() + 4; }

// Term p of the stream of Verlet row r (bead i, nc + nn + np terms), in two
// parts so that the list reads of a chunk go out ahead of its evaluation:
// split_fetch reads the partner and the pair's parameters (type 0 FENE with
// par.x = r0, 1 native, 2 repulsive, 3 no term), split_eval gathers the
// partner from the staged coordinates and forms (tx, ty, tz), -0.0 each where
// nothing is added.
struct SplitSrc {
    int j, type;
    float2 par;
};
__device__ __forceinline__ SplitSrc split_fetch(const DevJob& J, int p, int r, int i, int nc, int nn, int np) {
    SplitSrc f;
    f.j = 0;
    f.type = 3;
    f.par = make_float2(0.0f, 0.0f);
    if (p < nc) {
        f.type = 0;
        if (J.sort_rows) {
            f.j = gmem(J.rb_cov_j)[(size_t)p * J.pair_ld + r];
            f.par.x = gmem(J.rb_cov_r0)[(size_t)p * J.pair_ld + r];
        } else {
            f.j = J.nat_j16 ? (int)gmem(J.cov_j16)[(size_t)p * J.n + i] : gmem(J.cov_j)[(size_t)p * J.n + i];
            f.par.x = gmem(J.cov_r0)[(size_t)p * J.n + i];
        }
    } else if (p < nc + nn) {
        const int k = p - nc;
        f.type = 1;
        if (J.sort_rows) {
            f.j = gmem(J.rb_nat_j)[(size_t)k * J.pair_ld + r];
            f.par = J.rb_nat_par[(size_t)k * J.pair_ld + r];
        } else {
            f.j = J.nat_j16 ? (int)gmem(J.nat_j16)[(size_t)k * J.n + i] : gmem(J.nat_j)[(size_t)k * J.n + i];
            f.par = J.nat_par[(size_t)k * J.n + i];
        }
    } else if (p < nc + nn + np) {
        const size_t e = (size_t)(p - nc - nn) * J.pair_ld + r;
        f.type = 2;
        f.j = J.idx16 ? (int)__builtin_nontemporal_load(gmem(static_cast<const unsigned short*>(J.pair_j)) + e)
                      : __builtin_nontemporal_load(gmem(static_cast<const int*>(J.pair_j)) + e);
    }
    return f;
}
template <typename CP>
__device__ __forceinline__ void split_eval(const DevJob& J, const SplitSrc& f, CP cin, const float4 c, bool fast,
                                           float& tx, float& ty, float& tz) {
    tx = -0.0f, ty = -0.0f, tz = -0.0f;
    if (f.type == 3) return;
    const float4 o = cin[f.j];
    const float x = o.x - c.x, y = o.y - c.y, z = o.z - c.z;
    if (f.type == 0) {
        const float w = fene_w(J.ks, J.R2, x, y, z, f.par.x);
        tx = w * x;
        ty = w * y;
        tz = w * z;
    } else if (f.type == 1) {
        const float q = native_q(x, y, z, f.par);
        tx = x * q;
        ty = y * q;
        tz = z * q;
    } else {
        const float d2 = x * x + y * y + z * z;
        if (d2 < J.rc2) {
            const float q = pair_q(d2, J.a2, J.m6, fast);
            tx = x * q;
            ty = y * q;
            tz = z * q;
        }
    }
}

// Grid: ntiles * (64 / R) workgroups; workgroup b steps rows
// 64 * tile + R * (b % (64 / R)) .. + R - 1 of tile b / (64 / R) when the
// tile's job is an LDS-step job (G: a job above the LDS limit, whose
// partners are gathered from global memory, L2-resident; the other jobs are
// stepped by the other kernels).  Every lane works for row lane % R: it
// reads that row's bead and counts itself, and a worker reads its first two
// chunks' list entries, while the coordinates are staged; wavefront 0 also
// reads the extra force record and the state its tail needs.
template <bool FO, int R, bool G>
__global__ __launch_bounds__(1024) void k_step_split(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                                     const int2* __restrict__ tiles, int ntiles, int kstep,
                                                     float4* __restrict__ fout) {
    constexpr int K = split_k<R>(), KP = split_kp<R>(), S = 64 / R;
    extern __shared__ float4 lds_coord[];
    __shared__ __align__(16) float term[2][3][R][KP];
    __shared__ float fsum[3][R];
    const int tile = blockIdx.x / S;
    if (tile >= ntiles) return;
    const int2 td = tiles[tile];
    const DevJob& J = jobs[td.x];
    if (G ? J.lds_step : !J.lds_step) return;
    const long long s = st[td.x].step + kstep;
    if (!FO && s >= J.end_step) return;
    const int n = J.n;
    const int row0 = td.y + R * (blockIdx.x % S);
    if (row0 >= n) return;  // every exit above is uniform over the workgroup
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    if constexpr (!G) {
        const float4* __restrict__ g = J.coord[s & 1];
        for (int k = tid; k < n; k += 1024) lds_coord[k] = g[k];
    }
    const float4* __restrict__ cin = G ? J.coord[s & 1] : lds_coord;
    const bool owner = w < kSplitOwners;
    // this lane's row (loads from a clamped row where the tile ends early)
    const int q = lane % R, r = row0 + q, rc = min(r, n - 1);
    const int i = gmem(J.perm)[rc];
    int nc, nn;
    if (J.sort_rows) {
        const int cnt = gmem(J.rb_cnt)[rc];
        nc = cnt & 0xffff;
        nn = cnt >> 16;
    } else {
        nc = gmem(J.cov_cnt)[i];
        nn = gmem(J.nat_cnt)[i];
    }
    const int np = gmem(J.pair_cnt)[rc];
    const int len = r < n ? nc + nn + np : 0;
    // the longest row: lanes 0 .. R-1 of each wavefront hold every row
    int lmax = len;
#pragma unroll
    for (int o = R / 2; o > 0; o >>= 1) lmax = max(lmax, __shfl_xor(lmax, o));
    lmax = __shfl(lmax, 0);
    const int nch = (lmax + K - 1) / K;
    // a worker evaluates position pos of every chunk
    const int pos = (w - kSplitOwners) * S + lane / R;
    SplitSrc f0, f1;
    if (!owner) {
        f0 = split_fetch(J, pos < len ? pos : INT_MAX, r, i, nc, nn, np);
        f1 = split_fetch(J, K + pos < len ? K + pos : INT_MAX, r, i, nc, nn, np);
    }
    float acc = 0.0f;
    TailPre pre;
    if (owner && lane < R && r < n) {
        if (J.flags & JF_EXTRA_FORCE) {
            const float4 e = J.fextra[i];
            acc = w == 0 ? e.x : w == 1 ? e.y : e.z;
        }
        if (w == 0 && !FO) pre = tail_prefetch(J, i);
    }
    __syncthreads();  // coordinates staged; extra forces read by all three owners
    if constexpr (!FO) {
        if (w == 0 && lane < R && r < n && (J.flags & JF_EXTRA_FORCE))
            J.fextra[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    const bool fast = div_fast_numerator(J.a2) && J.rc2 <= kDivFastMaxB;
    const float4 c = cin[i];
    for (int ch = 0; ch <= nch; ++ch) {
        if (!owner && ch < nch) {
            float tx, ty, tz;
            split_eval(J, f0, cin, c, fast, tx, ty, tz);
            f0 = f1;
            const int p = (ch + 2) * K + pos;
            if (ch + 2 < nch) f1 = split_fetch(J, p < len ? p : INT_MAX, r, i, nc, nn, np);
            term[ch & 1][0][q][pos] = tx;
            term[ch & 1][1][q][pos] = ty;
            term[ch & 1][2][q][pos] = tz;
        }
        if (owner && ch > 0 && lane < R) {
            // chunk ch - 1 in stream order; positions up to the next multiple
            // of 4 past the longest row hold -0.0
            const float* __restrict__ t = &term[(ch - 1) & 1][w][lane][0];
            const int m = min(K, lmax - (ch - 1) * K);
            for (int k = 0; k < m; k += 4) {
                const float4 v = *reinterpret_cast<const float4*>(t + k);
                acc += v.x;
                acc += v.y;
                acc += v.z;
                acc += v.w;
            }
        }
        __syncthreads();
    }
    if (owner && lane < R) fsum[w][lane] = acc;
    __syncthreads();
    if (w == 0 && lane < R && r < n)
        step_tail<FO>(J, st[td.x], s, i, c, fsum[0][lane], fsum[1][lane], fsum[2][lane],
                      FO ? fout : J.coord[(s + 1) & 1], FO ? nullptr : &pre);
}

// ---------------------------------------------------------------------------
// pairlists
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool due(const DevJob& J, long long s, int freq) {
    return s < J.end_step && (s % J.stride == 0) && (s % freq == 0);
}

__device__ __forceinline__ unsigned long long lanes_below(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Possible pairs by the all-pairs scan, one wavefront per 64-bead tile:
// the job's beads are visited in chunks of 64 (one coordinate per lane,
// loaded once), and every bead i of the tile tests the chunk with its own
// coordinates read from the lane holding them.  The order-preserving
// ballot compaction appends the accepted j to bead i's row (bead-major,
// pp_j[i*pp_cap + k]); lane i keeps bead i's running count.  Exclusions
// (the sorted covalent/native partners) become, per chunk, a 64-bit mask
// built by lane i from its own list.  The predicate sqrtf(d^2) < threshold
// is evaluated as d^2 <= pp_d2 (the same set of floats).
__global__ __launch_bounds__(256) void k_possible(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st,
                                                  const int2* __restrict__ tiles, int ntiles, int force,
                                                  int use_cells, int skip_lds) {
    const int lane = threadIdx.x & 63;
    const TileRange tr = xcd_tiles(ntiles);
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int2 td = tiles[t];
        const DevJob& J = jobs[td.x];
        if (use_cells && J.cell_cnt) continue;  // built by k_possible_cells
        if (skip_lds && J.lds_step && J.idx16) continue;  // built by k_possible_lds
        const long long s = st[td.x].step;
        if (!force && !due(J, s, J.pp_freq)) continue;
        const int n = J.n;
        const CoordGlobal cur{J.coord[s & 1]};
        const float d2max = J.pp_d2;
        const int cap = J.pp_cap;
        const int i0 = td.y, nrow = min(64, n - i0);
        const bool own = lane < nrow;  // lane owns bead i0 + lane
        const float4 ci = own ? cur[i0 + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
        const int ne = own ? J.excl_cnt[i0 + lane] : 0;
        int pe = 0;
        int e_next = pe < ne ? J.excl_j[(size_t)pe * n + i0 + lane] : INT_MAX;
        int cnt = 0;  // lane i: entries of bead i0 + i so far
        for (int base = 0; base < n; base += 64) {
            const int j = base + lane;
            const float4 cj = j < n ? cur[j] : make_float4(0.f, 0.f, 0.f, 0.f);
            // this lane's bead: exclusions falling in [base, base + 64)
            unsigned long long ex = 0ull;
            while (e_next < base + 64) {
                ex |= 1ull << (e_next - base);
                ++pe;
                e_next = pe < ne ? J.excl_j[(size_t)pe * n + i0 + lane] : INT_MAX;
            }
            const int exlo = (int)(unsigned)ex, exhi = (int)(unsigned)(ex >> 32);
            for (int b = 0; b < nrow; ++b) {
                // bead b's coordinates, exclusion mask and count: v_readlane with
                // a uniform lane index
                const float xi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ci.x), b));
                const float yi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ci.y), b));
                const float zi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ci.z), b));
                const unsigned long long exb =
                    ((unsigned long long)(unsigned)__builtin_amdgcn_readlane(exhi, b) << 32) |
                    (unsigned)__builtin_amdgcn_readlane(exlo, b);
                const float x = cj.x - xi, y = cj.y - yi, z = cj.z - zi;
                const bool ok = j < n && j != i0 + b && !((exb >> lane) & 1ull) && x * x + y * y + z * z <= d2max;
                const unsigned long long m = __builtin_amdgcn_ballot_w64(ok);
                const int c0 = __builtin_amdgcn_readlane(cnt, b);
                if (ok) {
                    const int pos = c0 + __popcll(m & lanes_below(lane));
                    if (pos < cap) list_put(J.pp_j, J.idx16, (size_t)(i0 + b) * J.pp_ld + pos, j);
                }
                if (lane == b) cnt = c0 + __popcll(m);
            }
        }
        if (own) {
            J.pp_cnt[i0 + lane] = cnt < cap ? cnt : cap;
            if (cnt > cap) atomicOr(&st[td.x].overflow, 1);
        }
    }
}

// Possible pairs of the LDS-step jobs (16-bit entries; the LDS-step bound
// keeps them far below 65,536 beads), over the step kernel's units: the
// workgroup stages the coordinates of each due job in LDS, and wavefront w
// builds the lists of the beads of the unit's rows 64w..64w+63, one bead at
// a time.  The bead is wave-uniform, so its coordinates are one broadcast
// read, its sorted exclusions (and the bead itself) become a 64-bit mask
// per chunk of candidates in scalar registers, and the accepted count is a
// scalar; the candidates j come 64 per chunk from LDS in ascending order.
// The same list in the same order as k_possible.
// Possible-list rows r0 .. r1-1 of job J (r1 > r0) by one wavefront, from
// the job's coordinates `cur` staged in LDS: each row scans the job's beads
// in chunks of 64, ballot-compacts the accepted ones in ascending order, and
// drops the bead's exclusions and itself.  Overflow sets bit 0 of *ovf.
__device__ __forceinline__ void possible_rows(const DevJob& J, const float4* __restrict__ cur, int r0, int r1,
                                              int* ovf) {
    const int lane = threadIdx.x & 63;
    const int n = J.n;
    const float d2max = J.pp_d2;
    const int cap = J.pp_cap, ecap = J.excl_cap;
    const auto ecnt = gmem(J.excl_cnt);
    const auto ej = gmem(J.excl_j);
    const auto pp = gmem(static_cast<unsigned short*>(J.pp_j));
    const auto pcnt = gmem(J.pp_cnt);
    // bead i's exclusion count and its first 64 exclusions (lane k holds
    // the k-th), loaded one bead ahead from clamped addresses
    auto load_ex = [&](int i, int& ne, int& ev) {
        ne = ecnt[i];
        ev = ej[(size_t)min(lane, ecap - 1) * n + i];
    };
    int ne_n, ev_n;
    load_ex(r0, ne_n, ev_n);
    for (int i = r0; i < r1; ++i) {
        const int ne = ne_n;
        int ev = lane < ne ? ev_n : INT_MAX;
        load_ex(min(i + 1, r1 - 1), ne_n, ev_n);
        const float4 ci = cur[i];
        const auto row = pp + (size_t)i * J.pp_ld;
        int eg = 0, pe = 0;  // exclusion group (of 64) and index within it
        int e_next = ne > 0 ? __builtin_amdgcn_readlane(ev, 0) : INT_MAX;
        int self = i;  // the bead itself, until its chunk is passed
        int cnt = 0;
        // One chunk of 64 candidates from `base` with coordinates c; TAIL:
        // the chunk may run past the job's end.  The exclusion mask is
        // formed only for chunks holding an excluded partner or the bead.
        auto chunk = [&](int base, const float4 c, bool tail) {
            const int lim = base + 64;
            const float x = c.x - ci.x, y = c.y - ci.y, z = c.z - ci.z;
            bool acc = x * x + y * y + z * z <= d2max;
            if (tail) acc = acc & (base + lane < n);
            unsigned long long m = __builtin_amdgcn_ballot_w64(acc);
            if (min(e_next, self) < lim) {
                unsigned long long drop = 0ull;
                while (e_next < lim) {
                    drop |= 1ull << (e_next - base);
                    if (++pe == 64) {
                        eg += 64;
                        pe = 0;
                        ev = eg + lane < ne ? ej[(size_t)(eg + lane) * n + i] : INT_MAX;
                    }
                    e_next = eg + pe < ne ? __builtin_amdgcn_readlane(ev, pe) : INT_MAX;
                }
                if (self < lim) {
                    drop |= 1ull << (self - base);
                    self = INT_MAX;
                }
                m &= ~drop;
                acc = acc && !((drop >> lane) & 1ull);
            }
            if (m) {
                if (acc) {
                    const unsigned pos = __builtin_amdgcn_mbcnt_hi(
                        (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, (unsigned)cnt));
                    if ((int)pos < cap)
                        *reinterpret_cast<__attribute__((address_space(1))) unsigned short*>(
                            reinterpret_cast<__attribute__((address_space(1))) char*>(row) + (pos << 1)) =
                            (unsigned short)(base + lane);
                }
                cnt += __popcll(m);
            }
        };
        int b4 = 0;
        for (; b4 + 256 <= n; b4 += 256) {  // four whole chunks, their LDS reads in flight together
            float4 c[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) c[v] = cur[b4 + 64 * v + lane];
#pragma unroll
            for (int v = 0; v < 4; ++v) chunk(b4 + 64 * v, c[v], false);
        }
        if (b4 < n) {  // the last one to four chunks
            float4 c[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) c[v] = cur[min(b4 + 64 * v + lane, n - 1)];
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if (b4 + 64 * v < n) chunk(b4 + 64 * v, c[v], true);
        }
        if (lane == 0) {
            pcnt[i] = cnt < cap ? cnt : cap;
            if (cnt > cap) atomicOr(ovf, 1);
        }
    }
}

__device__ __forceinline__ bool possible_lds_wanted(const DevJob& J, long long s, int force, int use_cells) {
    return J.lds_step && J.idx16 && !(use_cells && J.cell_cnt) && (force || due(J, s, J.pp_freq));
}

__global__ __launch_bounds__(1024) void k_possible_lds(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st,
                                                       const int2* __restrict__ units, const int4* __restrict__ segs,
                                                       int nunits, int force, int use_cells) {
    extern __shared__ float4 lds_coord[];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int2 ud = units[u];  // first segment, segments
        __syncthreads();           // the previous unit's reads are done
        bool any = false;
        for (int q = 0; q < ud.y; ++q) {
            const int4 sg = segs[ud.x + q];
            const DevJob& J = jobs[sg.x];
            const long long s = st[sg.x].step;
            if (!possible_lds_wanted(J, s, force, use_cells)) continue;
            any = true;
            const float4* __restrict__ g = J.coord[s & 1];
            for (int k = threadIdx.x; k < J.n; k += blockDim.x) lds_coord[sg.w + k] = g[k];
        }
        if (!any) continue;  // the same for every thread of the block
        __syncthreads();
        int t = 64 * w;
        int q = 0;
        int4 sg = make_int4(0, 0, 0, 0);
        for (; q < ud.y; ++q) {  // the segment holding this wavefront's rows
            sg = segs[ud.x + q];
            const int span = (sg.z + 63) & ~63;
            if (t < span) break;
            t -= span;
        }
        if (q == ud.y) continue;
        const DevJob& J = jobs[sg.x];
        if (!possible_lds_wanted(J, st[sg.x].step, force, use_cells)) continue;
        possible_rows(J, lds_coord + sg.w, sg.y + t, sg.y + min(t + 64, sg.z), &st[sg.x].overflow);
    }
}

// The same lists spread for a few small jobs (the step's split_rows
// shape): four 1024-thread workgroups per 64-row tile, one wavefront per
// row, from the job's coordinates staged in LDS.  Grid: ntiles * 4.
__global__ __launch_bounds__(1024) void k_possible_split(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st,
                                                         const int2* __restrict__ tiles, int ntiles, int force,
                                                         int use_cells) {
    extern __shared__ float4 lds_coord[];
    const int tile = blockIdx.x >> 2, quarter = blockIdx.x & 3;
    if (tile >= ntiles) return;
    const int2 td = tiles[tile];
    const DevJob& J = jobs[td.x];
    const long long s = st[td.x].step;
    if (!possible_lds_wanted(J, s, force, use_cells)) return;
    const int n = J.n, row0 = td.y + 16 * quarter;
    if (row0 >= n) return;  // every exit above is uniform over the workgroup
    {
        const float4* __restrict__ g = J.coord[s & 1];
        for (int k = threadIdx.x; k < n; k += 1024) lds_coord[k] = g[k];
    }
    __syncthreads();
    const int i = row0 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (i < n) possible_rows(J, lds_coord, i, i + 1, &st[td.x].overflow);
}

// ---------------------------------------------------------------------------
// cell-binned possible pairs (large jobs)
// ---------------------------------------------------------------------------
// The list of k_possible -- the same predicate on the same operands, the
// same ascending-j order -- from the candidates in the 27 cells around the
// bead instead of the whole job.  Cells are cubes of edge thr*(1+1e-4) +
// 0.01 A: a pair with sqrtf(d^2) < thr has every |coordinate difference|
// within thr*(1 + 2^-21), and the cell index (x - x0) / edge is computed
// with an error far below the margin, so such a pair always lies in
// adjacent cells.  Accepted j are set in a per-wave LDS bitmap over the
// job's beads, the bead's exclusions cleared, and the bitmap read back in
// index order with a wave prefix sum of popcounts.

__device__ __forceinline__ int cell_axis(float x, float org, float inv, int dim) {
    const float q = (x - org) * inv;
    if (!(q >= 0.0f)) return 0;  // below the grid, or NaN
    if (q >= (float)dim) return dim - 1;
    return min((int)q, dim - 1);
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_inclusive_sum(int v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o);
        if (lane >= o) v += u;
    }
    return v;
}

// Per job (one 1024-thread block each): bounding box of the coordinates,
// the grid (edge grown until the cell count fits cell_cap), zeroed counts.
__global__ __launch_bounds__(1024) void k_cells_setup(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st,
                                                      int njobs, int force) {
    const int jb = blockIdx.x;
    if (jb >= njobs) return;
    const DevJob& J = jobs[jb];
    if (!J.cell_cnt) return;
    DevJobState& S = st[jb];
    const long long s = S.step;
    if (!force && !due(J, s, J.pp_freq)) return;
    const float4* __restrict__ cur = J.coord[s & 1];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = threadIdx.x; i < J.n; i += blockDim.x) {
        const float4 r = cur[i];
        lo[0] = fminf(lo[0], r.x), hi[0] = fmaxf(hi[0], r.x);
        lo[1] = fminf(lo[1], r.y), hi[1] = fmaxf(hi[1], r.y);
        lo[2] = fminf(lo[2], r.z), hi[2] = fmaxf(hi[2], r.z);
    }
    __shared__ float red[6][16];
    __shared__ int dims[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float a = lo[d], b = hi[d];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            a = fminf(a, __shfl_xor(a, o));
            b = fmaxf(b, __shfl_xor(b, o));
        }
        if (lane == 0) {
            red[d][w] = a;
            red[3 + d][w] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        float org[3], ext[3];
        for (int d = 0; d < 3; ++d) {
            float a = red[d][0], b = red[3 + d][0];
            for (int k = 1; k < nw; ++k) a = fminf(a, red[d][k]), b = fmaxf(b, red[3 + d][k]);
            org[d] = a;
            ext[d] = b - a;
        }
        float edge = J.pp_thr * 1.0001f + 0.01f;
        int dm[3] = {1, 1, 1};
        const bool finite = isfinite(ext[0]) && isfinite(ext[1]) && isfinite(ext[2]);
        if (finite) {
            for (int it = 0; it < 200; ++it) {
                double cells = 1.0;
                for (int d = 0; d < 3; ++d) cells *= floor((double)ext[d] / edge) + 1.0;
                if (cells <= (double)J.cell_cap) break;
                edge *= 1.1f;
            }
            for (int d = 0; d < 3; ++d) dm[d] = (int)floor((double)ext[d] / edge) + 1;
            if ((long long)dm[0] * dm[1] * dm[2] > J.cell_cap) dm[0] = dm[1] = dm[2] = 1;
        } else {
            org[0] = org[1] = org[2] = 0.0f;
        }
        for (int d = 0; d < 3; ++d) {
            S.cell_org[d] = org[d];
            S.cell_dim[d] = dm[d];
        }
        S.cell_inv = 1.0f / edge;
        dims[0] = dm[0] * dm[1] * dm[2];
    }
    __syncthreads();
    const int nc = dims[0];
    for (int c = threadIdx.x; c < nc; c += blockDim.x) J.cell_cnt[c] = 0;
}

// Each bead's cell and its slot there (the order within a cell is free:
// the bitmap restores index order).
__global__ __launch_bounds__(256) void k_cells_count(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st,
                                                     const int2* __restrict__ tiles, int ntiles, int force) {
    const int lane = threadIdx.x & 63;
    const TileRange tr = xcd_tiles(ntiles);
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int2 td = tiles[t];
        const DevJob& J = jobs[td.x];
        if (!J.cell_cnt) continue;
        const DevJobState& S = st[td.x];
        const long long s = S.step;
        if (!force && !due(J, s, J.pp_freq)) continue;
        const int i = td.y + lane;
        if (i >= J.n) continue;
        const float4 r = J.coord[s & 1][i];
        const int cx = cell_axis(r.x, S.cell_org[0], S.cell_inv, S.cell_dim[0]);
        const int cy = cell_axis(r.y, S.cell_org[1], S.cell_inv, S.cell_dim[1]);
        const int cz = cell_axis(r.z, S.cell_org[2], S.cell_inv, S.cell_dim[2]);
        const int c = (cz * S.cell_dim[1] + cy) * S.cell_dim[0] + cx;
        J.bead_cell[i] = c;
        J.bead_slot[i] = atomicAdd(&J.cell_cnt[c], 1);
    }
}

// Exclusive scan of the cell counts (one 1024-thread block per job).
__global__ __launch_bounds__(1024) void k_cells_scan(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st,
                                                     int njobs, int force) {
    const int jb = blockIdx.x;
    if (jb >= njobs) return;
    const DevJob& J = jobs[jb];
    if (!J.cell_cnt) return;
    const DevJobState& S = st[jb];
    if (!force && !due(J, S.step, J.pp_freq)) return;
    const int nc = S.cell_dim[0] * S.cell_dim[1] * S.cell_dim[2];
    const int per = (nc + blockDim.x - 1) / blockDim.x;
    const int b = threadIdx.x * per, e = min(b + per, nc);
    int sum = 0;
    for (int c = b; c < e; ++c) sum += J.cell_cnt[c];
    __shared__ int part[1024];
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = 0; k < (int)blockDim.x; ++k) {
            const int v = part[k];
            part[k] = acc;
            acc += v;
        }
    }
    __syncthreads();
    int acc = part[threadIdx.x];
    for (int c = b; c < e; ++c) {
        J.cell_start[c] = acc;
        acc += J.cell_cnt[c];
    }
    if (threadIdx.x == 0) J.cell_start[nc] = J.n;
}

__global__ __launch_bounds__(256) void k_cells_fill(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st,
                                                    const int2* __restrict__ tiles, int ntiles, int force) {
    const int lane = threadIdx.x & 63;
    const TileRange tr = xcd_tiles(ntiles);
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int2 td = tiles[t];
        const DevJob& J = jobs[td.x];
        if (!J.cell_cnt) continue;
        if (!force && !due(J, st[td.x].step, J.pp_freq)) continue;
        const int i = td.y + lane;
        if (i >= J.n) continue;
        J.cell_beads[J.cell_start[J.bead_cell[i]] + J.bead_slot[i]] = i;
    }
}

// One 64-thread workgroup (one wavefront) per tile; dynamic LDS holds the
// bitmap of ceil(n/64) words for the largest cell-binned job.
// One wavefront per half tile (32 beads), its beads one after another.  A bead's
// candidates (the 3 x 3 rows of three x-adjacent cells, each row one
// contiguous range of cell_beads) are read eight 64-entry chunks at a time,
// all index loads and then all coordinate gathers in flight together; the
// next bead's coordinates and first 64 exclusions are loaded while this one
// is scanned.  Accepted j set bit j of the LDS bitmap bm and bit j / 64 of
// bm2, so the read-back in index order visits only the 64-word blocks of bm
// that hold an entry.
constexpr int kCellChunks = 8;
__global__ __launch_bounds__(64) void k_possible_cells(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st,
                                                       const int2* __restrict__ tiles, int ntiles, int force,
                                                       int cell_words) {
    extern __shared__ unsigned long long bm[];
    unsigned long long* const bm2 = bm + cell_words;  // one bit per word of bm
    const int lane = threadIdx.x;
    const int nb = gridDim.x, nxcd = nb >= 8 ? 8 : 1;
    const int nhalf = 2 * ntiles;  // work units: the two halves of every tile
    const int slice = (nhalf + nxcd - 1) / nxcd;
    const int xcd = blockIdx.x % nxcd;
    const int hend = min((xcd + 1) * slice, nhalf);
    for (int h = xcd * slice + blockIdx.x / nxcd; h < hend; h += nb / nxcd) {
        const int2 td = tiles[h >> 1];
        const DevJob& J = jobs[td.x];
        if (!J.cell_cnt) continue;
        const DevJobState& S = st[td.x];
        const long long s = S.step;
        if (!force && !due(J, s, J.pp_freq)) continue;
        const int n = J.n;
        const int i0 = td.y + 32 * (h & 1);
        if (i0 >= n) continue;
        const int nwords = (n + 63) >> 6;
        for (int w = lane; w < nwords; w += 64) bm[w] = 0ull;
        for (int w = lane; w < (nwords + 63) >> 6; w += 64) bm2[w] = 0ull;
        __syncthreads();
        const CoordGlobal cur{J.coord[s & 1]};
        const auto cbeads = gmem(J.cell_beads);
        const auto cstart = gmem(J.cell_start);
        const auto ecnt = gmem(J.excl_cnt);
        const auto ej = gmem(J.excl_j);
        const float d2max = J.pp_d2;
        const int cap = J.pp_cap, ecap = J.excl_cap;
        const int dx = S.cell_dim[0], dy = S.cell_dim[1], dz = S.cell_dim[2];
        const int iend = min(i0 + 32, n);
        float4 r1n = cur[i0];
        int nen = ecnt[i0];
        int evn = ej[(size_t)min(lane, ecap - 1) * n + i0];
        for (int i = i0; i < iend; ++i) {
            const float4 r1 = r1n;
            const int ne = nen, ev = evn;
            if (i + 1 < iend) {
                r1n = cur[i + 1];
                nen = ecnt[i + 1];
                evn = ej[(size_t)min(lane, ecap - 1) * n + i + 1];
            }
            const int cx = cell_axis(r1.x, S.cell_org[0], S.cell_inv, dx);
            const int cy = cell_axis(r1.y, S.cell_org[1], S.cell_inv, dy);
            const int cz = cell_axis(r1.z, S.cell_org[2], S.cell_inv, dz);
            int jlo = INT_MAX, jhi = -1;
            for (int z = max(cz - 1, 0); z <= min(cz + 1, dz - 1); ++z)
                for (int y = max(cy - 1, 0); y <= min(cy + 1, dy - 1); ++y) {
                    const int row = (z * dy + y) * dx;
                    // the cells x-1..x+1 of one row are contiguous in cell_beads
                    const int b = cstart[row + max(cx - 1, 0)];
                    const int e = cstart[row + min(cx + 1, dx - 1) + 1];
                    for (int k0 = b; k0 < e; k0 += 64 * kCellChunks) {
                        int jj[kCellChunks];
#pragma unroll
                        for (int u = 0; u < kCellChunks; ++u) {
                            const int k = k0 + 64 * u + lane;
                            jj[u] = k < e ? cbeads[k] : i;  // past the range: the bead itself, skipped
                        }
                        float4 r2[kCellChunks];
#pragma unroll
                        for (int u = 0; u < kCellChunks; ++u) r2[u] = cur[jj[u]];
#pragma unroll
                        for (int u = 0; u < kCellChunks; ++u) {
                            const int j = jj[u];
                            const float x = r2[u].x - r1.x, y2 = r2[u].y - r1.y, z2 = r2[u].z - r1.z;
                            if (j != i && x * x + y2 * y2 + z2 * z2 <= d2max) {
                                atomicOr(&bm[j >> 6], 1ull << (j & 63));
                                atomicOr(&bm2[j >> 12], 1ull << ((j >> 6) & 63));
                                jlo = min(jlo, j);
                                jhi = max(jhi, j);
                            }
                        }
                    }
                }
            __syncthreads();
            for (int k = lane; k < ne; k += 64) {
                const int e = k < 64 ? ev : ej[(size_t)k * n + i];
                atomicAnd(&bm[e >> 6], ~(1ull << (e & 63)));
            }
            __syncthreads();
            jlo = wave_min(jlo);
            jhi = wave_max(jhi);
            int cnt = 0;
            const size_t rowp = (size_t)i * J.pp_ld;
            if (jhi >= 0) {
                const int wlo = jlo >> 6, whi = jhi >> 6;
                for (int wb = wlo & ~63; wb <= whi; wb += 64) {
                    const unsigned long long m2 = bm2[wb >> 6];
                    if (m2 == 0ull) continue;  // no entry in these 64 words
                    const int w = wb + lane;
                    unsigned long long word = 0ull;
                    if ((m2 >> lane) & 1ull) {
                        word = bm[w];
                        bm[w] = 0ull;
                    }
                    const int pc = __popcll(word);
                    const int incl = wave_inclusive_sum(pc, lane);
                    int pos = cnt + incl - pc;
                    while (word) {
                        const int j = (w << 6) + __builtin_ctzll(word);
                        if (pos < cap) list_put(J.pp_j, J.idx16, rowp + pos, j);
                        ++pos;
                        word &= word - 1ull;
                    }
                    cnt += __shfl(incl, 63);
                    if (lane == 0) bm2[wb >> 6] = 0ull;
                }
            }
            __syncthreads();
            if (lane == 0) {
                J.pp_cnt[i] = cnt < cap ? cnt : cap;
                if (cnt > cap) atomicOr(&st[td.x].overflow, 1);
            }
        }
    }
}

// Verlet list: one 1024-thread workgroup per tile of 64 list rows (the
// beads perm[row]).  Wave w filters the possible lists of the beads of rows
// 4w..4w+3 in stored order with the ballot compaction; the accepted
// entries of the tile collect in LDS rows, and the workgroup then writes
// them out slot-major (pair_j[k*n + row]) so each store covers the 64 rows
// of one slot.  A bead's row is read eight 64-entry
// chunks at a time: all eight index loads, then all eight coordinate
// gathers, are in flight before the first compare.  Entries past the LDS
// row (VerletLds::entries) are rare and stored to their slot directly.
template <typename IdxT>
struct VerletLds;
template <>
struct VerletLds<unsigned short> {  // 256 entries per bead, rows padded by one bank
    static constexpr int entries = 256, stride = 258;
};
template <>
struct VerletLds<int> {
    static constexpr int entries = 255, stride = 256;
};

// w: this wavefront's index among the WAVES wavefronts that build the tile.
// CH: chunks of 64 possible-list entries whose loads are in flight together.
// cur: the job's coordinates (staged in LDS, or a CoordGlobal).
template <typename IdxT, int WAVES, int CH = 8, typename CP = const float4*>
__device__ __forceinline__ void verlet_tile(const DevJob& J, DevJobState& S, CP cur, int i0,
                                            IdxT* __restrict__ rows, int* __restrict__ cnt_s, int w) {
    constexpr int E = VerletLds<IdxT>::entries, RS = VerletLds<IdxT>::stride;
    constexpr int RPW = 64 / WAVES;  // rows per wave
    const int lane = threadIdx.x & 63;
    const int n = J.n;
    const int nrow = min(64, n - i0);
    w = __builtin_amdgcn_readfirstlane(w);
    const int b_lo = RPW * w, b_hi = min(b_lo + RPW, nrow);
    const float d2max = J.pl_d2;  // sqrtf(d2) < pl_cut, exactly
    const int cap = J.pair_cap, ppcap = J.pp_cap, ppld = J.pp_ld;
    const auto pp = gmem(static_cast<const IdxT*>(J.pp_j));
    const auto pj = gmem(static_cast<IdxT*>(J.pair_j));
    // beads of the tile's rows and their possible counts (loads issued
    // unconditionally, from clamped addresses, so that no branch leaves a
    // load pending on one path only)
    const int bl0 = gmem(J.perm)[i0 + min(lane, nrow - 1)];
    const int cl0 = gmem(J.pp_cnt)[bl0];
    const int bl = lane < nrow ? bl0 : 0;
    const int cl = lane < nrow ? cl0 : 0;
    // row b's bead and count (b is wave-uniform: a lane read, not a shuffle)
    auto bead_of = [&](int b) { return __builtin_amdgcn_readlane(bl, b); };
    auto count_of = [&](int b) { return __builtin_amdgcn_readlane(cl, b); };
    // Raw words of 64 CH entries of row b from `base`, loaded without a
    // branch from addresses clamped into the row; the filter replaces the
    // words past the row's count by 0 (a valid bead) and rejects them.
    auto load_idx = [&](int b, int base, int (&j)[CH]) {
        const auto row = pp + (size_t)bead_of(b) * ppld;
#pragma unroll
        for (int u = 0; u < CH; ++u) j[u] = (int)row[(unsigned)min(base + 64 * u + lane, ppcap - 1)];
    };
    // One chunk: gather, compare, and compact the accepted entries in order
    // into the row's LDS entries (or, past them, their slots).  Gathers go
    // out four at a time.
    auto filter = [&](int b, const float4 r1, int npp, int base, int (&j)[CH], int& cnt) {
#pragma unroll
        for (int u = 0; u < CH; ++u) j[u] = base + 64 * u + lane < npp ? j[u] : 0;
#pragma unroll
        for (int h = 0; h < CH; h += 4) {
            float4 c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) c[u] = cur[j[h + u]];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k0 = base + 64 * (h + u);
                if (k0 < npp) {
                    const float x = c[u].x - r1.x, y = c[u].y - r1.y, z = c[u].z - r1.z;
                    const bool ok = (k0 + lane < npp) & (x * x + y * y + z * z <= d2max);
                    const unsigned long long m = __builtin_amdgcn_ballot_w64(ok);
                    const int total = cnt + __popcll(m);
                    if (ok) {
                        const int pos = (int)__builtin_amdgcn_mbcnt_hi(
                            (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, (unsigned)cnt));
                        if (total <= E || pos < E) rows[b * RS + pos] = (IdxT)j[h + u];
                        else if (pos < cap) pj[(size_t)pos * J.pair_ld + i0 + b] = (IdxT)j[h + u];
                    }
                    cnt = total;
                }
            }
        }
    };
    // Row b from its first chunk of indices j (loaded ahead); a second chunk
    // is loaded before the first is filtered, further ones as they come.
    auto do_row = [&](int b, int (&j)[CH], const float4 r1) {
        const int npp = count_of(b);
        int cnt = 0;
        int j2[CH];
        const bool two = npp > 64 * CH;
        if (two) load_idx(b, 64 * CH, j2);
        filter(b, r1, npp, 0, j, cnt);
        if (two) {
            filter(b, r1, npp, 64 * CH, j2, cnt);
            for (int base = 128 * CH; base < npp; base += 64 * CH) {
                int j3[CH];
                load_idx(b, base, j3);
                filter(b, r1, npp, base, j3, cnt);
            }
        }
        if (lane == 0) {
            cnt_s[b] = cnt;
            if (cnt > cap) atomicOr(&S.overflow, 2);
        }
    };
    // first chunk and centre of row b (rows past the wave's last load that
    // row's again, unused: no branch)
    auto ahead = [&](int b, int (&j)[CH], float4& r1) {
        b = min(b, b_hi - 1);
        load_idx(b, 0, j);
        r1 = cur[bead_of(b)];
    };
    // The wave's rows as straight-line code (the row loop fully unrolled),
    // a software pipeline two rows deep: while row b is filtered the first
    // chunk of row b+1 is in flight, and row b+2's is issued as soon as row
    // b is done with its registers.  Unrolled, every prefetch has registers
    // of its own and the loads can be waited for one by one.
    if (b_lo >= b_hi) return;
    constexpr int D = RPW < 2 ? 1 : 2;
    int js[D][CH];
    float4 rs[D];
#pragma unroll
    for (int q = 0; q < D; ++q) ahead(b_lo + q, js[q], rs[q]);
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int b = b_lo + q;
        if (b < b_hi) {
            do_row(b, js[q % D], rs[q % D]);
            if (q + D < RPW) ahead(b + D, js[q % D], rs[q % D]);
        }
    }
}

// t: this thread's index among the 64 WAVES threads that store the tile.
template <typename IdxT, int WAVES>
__device__ __forceinline__ void verlet_store(const DevJob& J, int i0, const IdxT* __restrict__ rows,
                                             const int* __restrict__ cnt_s, int t) {
    constexpr int E = VerletLds<IdxT>::entries, RS = VerletLds<IdxT>::stride;
    const int n = J.n, cap = J.pair_cap;
    const int nrow = min(64, n - i0);
    const int b = t & 63, kq = t >> 6;
    const int c = b < nrow ? cnt_s[b] : 0;
    if (b < nrow) {
        if (kq == 0) J.pair_cnt[i0 + b] = c < cap ? c : cap;
    }
    int maxc = min(c, E);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) maxc = max(maxc, __shfl_xor(maxc, o));
    // the four waves' maxima differ only in which beads they saw: all waves
    // see all 64 beads here (b = lane), so maxc is the tile's
    const auto pj = gmem(static_cast<IdxT*>(J.pair_j));
    // a row that outgrew its capacity (flag bit 2 is set) stores only its
    // first cap entries: pair_j holds pair_cap slots
    const int lim = min(c, min(E, cap));
    for (int k = kq; k < maxc; k += WAVES)
        if (k < lim) pj[(size_t)k * J.pair_ld + i0 + b] = rows[b * RS + k];
}

// Re-sorts the rows of each kLdsUnit-row unit of a job by Verlet list
// length (longest first, ties by bead), just before the rebuild that fills
// the rows again, with the lengths of the last rebuild as keys: the 64 rows
// of a wavefront then hold lists of about the same length, so its loop
// over the longest list idles few lanes and reads few padding slots.  Only
// which row and lane serve a bead changes, never the arithmetic, so the
// trajectories are the same bits in any row order.  Bitonic sort in LDS.
// The sort also moves each bead's bonded lists to its new row (the step
// reads them in row order), so it runs at every kSortEvery-th Verlet
// rebuild only: list lengths drift little in between.
__global__ __launch_bounds__(1024) void k_sort_rows(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                                    const int2* __restrict__ units, int nunits, int sort_every,
                                                    int force) {
    __shared__ unsigned long long key[kLdsUnit];
    const int t = threadIdx.x;
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int2 ud = units[u];
        const DevJob& J = jobs[ud.x];
        if (!J.sort_rows) continue;
        if (!force && !due(J, st[ud.x].step, J.pairs_freq * sort_every)) continue;  // a Verlet rebuild step
        const int m = min(kLdsUnit, J.n - ud.y), r = ud.y + t;
        __syncthreads();  // the previous unit's keys are read
        key[t] = t < m ? ((unsigned long long)(unsigned)(J.pair_cap - J.pair_cnt[r]) << 32) | (unsigned)J.perm[r]
                       : ~0ull;
        __syncthreads();
        for (int k = 2; k <= kLdsUnit; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1) {
                const int q = t ^ j;
                if (q > t) {
                    const unsigned long long a = key[t], b = key[q];
                    if ((a > b) == ((t & k) == 0)) {
                        key[t] = b;
                        key[q] = a;
                    }
                }
                __syncthreads();
            }
        if (t < m) {
            // row r now steps bead i: its bonded lists move to row r
            const int i = (int)(unsigned)(key[t] & 0xffffffffull);
            J.perm[r] = i;
            const int n = J.n, ld = J.pair_ld;
            const int nc = J.cov_cnt[i], nn = J.nat_cnt[i];
            J.rb_cnt[r] = nc | (nn << 16);
            for (int k = 0; k < nc; ++k) {
                J.rb_cov_j[(size_t)k * ld + r] = J.cov_j16[(size_t)k * n + i];
                J.rb_cov_r0[(size_t)k * ld + r] = J.cov_r0[(size_t)k * n + i];
            }
            for (int k = 0; k < nn; ++k) {
                J.rb_nat_j[(size_t)k * ld + r] = J.nat_j16[(size_t)k * n + i];
                J.rb_nat_par[(size_t)k * ld + r] = J.nat_par[(size_t)k * n + i];
            }
        }
    }
}

// Wavefronts per 64-row tile of k_verlet (four rows each): sixteen, so that
// the latency of the L2 gathers is hidden by the other wavefronts of the CU.
constexpr int kVerletWaves = 16;
__global__ __launch_bounds__(1024) void k_verlet(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st,
                                                 const int2* __restrict__ tiles, int ntiles, int force, int skip_lds) {
    extern __shared__ __attribute__((aligned(16))) int lds_rows[];  // 64 rows of uint16 or int32 entries
    __shared__ int cnt_s[64];
    const int nb = gridDim.x, nxcd = nb >= 8 ? 8 : 1;
    const int slice = (ntiles + nxcd - 1) / nxcd;
    const int xcd = blockIdx.x % nxcd;
    const int tend = min((xcd + 1) * slice, ntiles);
    for (int t = xcd * slice + blockIdx.x / nxcd; t < tend; t += nb / nxcd) {
        const int2 td = tiles[t];
        const DevJob& J = jobs[td.x];
        DevJobState& S = st[td.x];
        const long long s = S.step;
        if (!force && !due(J, s, J.pairs_freq)) continue;
        if (skip_lds && J.lds_step) continue;  // built by k_verlet_lds
        const CoordGlobal cur{J.coord[s & 1]};
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        if (J.idx16)
            verlet_tile<unsigned short, kVerletWaves>(J, S, cur, td.y, reinterpret_cast<unsigned short*>(lds_rows),
                                                      cnt_s, w);
        else
            verlet_tile<int, kVerletWaves>(J, S, cur, td.y, lds_rows, cnt_s, w);
        __syncthreads();
        if (J.idx16)
            verlet_store<unsigned short, kVerletWaves>(J, td.y, reinterpret_cast<const unsigned short*>(lds_rows),
                                                       cnt_s, threadIdx.x);
        else
            verlet_store<int, kVerletWaves>(J, td.y, static_cast<const int*>(lds_rows), cnt_s, threadIdx.x);
        __syncthreads();
    }
}

// The Verlet build of the LDS-step jobs, over the step kernel's units: the
// workgroup stages the coordinates of each job it serves in LDS, so the
// candidate gathers read LDS instead of L2, then builds the unit's rows two
// 64-row tiles at a time (wavefronts 0-7 one tile, 8-15 the other, eight
// rows each) and writes each tile slot-major as k_verlet does.  Same
// lists, same order.
constexpr int kVerletRowsBytes = 64 * VerletLds<unsigned short>::stride * 2;
template <int CH>
__global__ __launch_bounds__(1024) void k_verlet_lds(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st,
                                                     const int2* __restrict__ units, const int4* __restrict__ segs,
                                                     int nunits, int force) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    float4* coords = reinterpret_cast<float4*>(lds_raw + 2 * kVerletRowsBytes);
    __shared__ int cnt_s[2][64];
    const int grp = threadIdx.x >> 9, gt = threadIdx.x & 511;  // tile group, thread in group
    unsigned short* rows = reinterpret_cast<unsigned short*>(lds_raw + grp * kVerletRowsBytes);
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int2 ud = units[u];
        __syncthreads();  // the previous unit's reads are done
        for (int q = 0; q < ud.y; ++q) {
            const int4 sg = segs[ud.x + q];
            const DevJob& J = jobs[sg.x];
            const long long s = st[sg.x].step;
            if (!force && !due(J, s, J.pairs_freq)) continue;
            const float4* __restrict__ g = J.coord[s & 1];
            for (int k = threadIdx.x; k < J.n; k += blockDim.x) coords[sg.w + k] = g[k];
        }
        __syncthreads();
        for (int q = 0; q < ud.y; ++q) {
            const int4 sg = segs[ud.x + q];
            const DevJob& J = jobs[sg.x];
            DevJobState& S = st[sg.x];
            if (!force && !due(J, S.step, J.pairs_freq)) continue;
            for (int t0 = 0; t0 < sg.z; t0 += 128) {
                const int i0 = sg.y + t0 + 64 * grp;
                const bool mine = t0 + 64 * grp < sg.z;
                if (mine) verlet_tile<unsigned short, 8, CH>(J, S, coords + sg.w, i0, rows, cnt_s[grp], gt >> 6);
                __syncthreads();
                if (mine) verlet_store<unsigned short, 8>(J, i0, rows, cnt_s[grp], gt);
                __syncthreads();
            }
        }
    }
}

// The Verlet build of the LDS-step jobs spread for a few small jobs (the
// step's split_rows shape): four 1024-thread workgroups per 64-row tile,
// each builds 16 rows, one wavefront per row (the same in-order ballot
// compaction as k_verlet_lds, so the same lists), from the job's
// coordinates staged in LDS.  Grid: ntiles * 4.
__global__ __launch_bounds__(1024) void k_verlet_split(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st,
                                                       const int2* __restrict__ tiles, int ntiles, int force) {
    constexpr int E = VerletLds<unsigned short>::entries, RS = VerletLds<unsigned short>::stride;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    unsigned short* rows = reinterpret_cast<unsigned short*>(lds_raw);
    float4* coords = reinterpret_cast<float4*>(lds_raw + kVerletRowsBytes);
    __shared__ int cnt_s[64];
    const int tile = blockIdx.x >> 2, quarter = blockIdx.x & 3;
    if (tile >= ntiles) return;
    const int2 td = tiles[tile];
    const DevJob& J = jobs[td.x];
    if (!J.lds_step) return;  // built by k_verlet
    DevJobState& S = st[td.x];
    const long long s = S.step;
    if (!force && !due(J, s, J.pairs_freq)) return;
    const int n = J.n, i0 = td.y, nrow = min(64, n - i0);
    if (16 * quarter >= nrow) return;  // every exit above is uniform over the workgroup
    {
        const float4* __restrict__ g = J.coord[s & 1];
        for (int k = threadIdx.x; k < n; k += 1024) coords[k] = g[k];
    }
    __syncthreads();
    const int w = threadIdx.x >> 6;
    verlet_tile<unsigned short, 64, 8>(J, S, coords, i0, rows, cnt_s, 16 * quarter + w);
    __syncthreads();
    // this workgroup's 16 rows: thread (row b, entry lane kq)
    const int b = 16 * quarter + (threadIdx.x & 15), kq = threadIdx.x >> 4;
    if (b < nrow) {
        const int c = cnt_s[b], cap = J.pair_cap;
        if (kq == 0) J.pair_cnt[i0 + b] = c < cap ? c : cap;
        const auto pj = gmem(static_cast<unsigned short*>(J.pair_j));
        const int lim = min(c, min(E, cap));
        for (int k = kq; k < lim; k += 64) pj[(size_t)k * J.pair_ld + i0 + b] = rows[b * RS + k];
    }
}

// ---------------------------------------------------------------------------
// per-job updaters at stride boundaries
// ---------------------------------------------------------------------------
// One wavefront per job.  The reaction on the tip is summed in bead order,
// as the host loop of the reference sums it: the wavefront loads the
// beads' forces kTipChunk at a time (coalesced, all in flight together, and
// clears them), stages their components in LDS, and lane c < 3 adds
// component c bead by bead; the components' sums are then gathered to lane 0.
constexpr int kTipChunk = 512;
__device__ void tip_update(const DevJob& J, DevJobState& S, long long s) {
    __shared__ __align__(16) float comp[3][kTipChunk];
    const int lane = threadIdx.x & 63;
    if (s == J.retraction && lane == 0) S.V = -S.V;
    if (s % J.nav != 0) return;
    float acc = 0.0f;  // lane c < 3: component c
    for (int base = 0; base < J.n; base += kTipChunk) {
        float4 v[kTipChunk / 64];
#pragma unroll
        for (int u = 0; u < kTipChunk / 64; ++u) {
            const int i = base + 64 * u + lane;
            v[u] = i < J.n ? J.tipF[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
#pragma unroll
        for (int u = 0; u < kTipChunk / 64; ++u) {
            const int i = base + 64 * u + lane;
            if (i < J.n) J.tipF[i] = make_float4(0.0f, 0.0f, 0.0f, v[u].w);
            comp[0][64 * u + lane] = v[u].x;
            comp[1][64 * u + lane] = v[u].y;
            comp[2][64 * u + lane] = v[u].z;
        }
        __syncthreads();  // the workgroup is this one wavefront
        const int m = min(kTipChunk, J.n - base);
        if (lane < 3) {
            const float* __restrict__ c = comp[lane];
            int b = 0;
            for (; b + 4 <= m; b += 4) {
                const float4 q = *reinterpret_cast<const float4*>(c + b);
                acc += q.x;
                acc += q.y;
                acc += q.z;
                acc += q.w;
            }
            for (; b < m; ++b) acc += c[b];
        }
        __syncthreads();  // the sums' reads are done before the next chunk's writes
    }
    float f0 = __shfl(acc, 0), f1 = __shfl(acc, 1), f2 = __shfl(acc, 2);
    if (lane != 0) return;
    const float fn = (float)J.nav;
    f0 /= fn;
    f1 /= fn;
    f2 /= fn;
    S.tipFav[0] = f0;
    S.tipFav[1] = f1;
    S.tipFav[2] = f2;
    S.tipFav[3] = sqrtf(f0 * f0 + f1 * f1 + f2 * f2);
    S.dx = (float)(s / J.nav) * S.V;
    if (J.moveSurface) {
        for (int d = 0; d < 3; ++d) S.micaR[d] = J.micaR0[d] + J.dir[d] * S.dx;
    } else {
        for (int d = 0; d < 3; ++d) S.chip[d] = J.chip0[d] + J.dir[d] * S.dx;
    }
    const float dr0 = S.chip[0] - S.tip[0], dr1 = S.chip[1] - S.tip[1], dr2 = S.chip[2] - S.tip[2];
    S.tipDr = sqrtf(dr0 * dr0 + dr1 * dr1 + dr2 * dr2);
    const float df0 = -f0 + dr0 * J.indKs, df1 = -f1 + dr1 * J.indKs, df2 = -f2 + dr2 * J.indKs;
    const float mult = ((J.h * ((float)J.nav)) / (J.tipZeta));
    if (J.fixTrans) {
        const float ct = mult * df0 * J.dir[0] + mult * df1 * J.dir[1] + mult * df2 * J.dir[2];
        for (int d = 0; d < 3; ++d) S.tip[d] += ct * J.dir[d];
    } else {
        S.tip[0] += mult * df0;
        S.tip[1] += mult * df1;
        S.tip[2] += mult * df2;
    }
}

__device__ void pull_update(const DevJob& J, DevJobState& S, long long s) {
    const float xt = J.deltax * (float)(s / J.pullFreq);
    const float4 cn = J.coord[s & 1][J.pulledEnd];
    S.extF[0] = J.k_trans * (S.cant[0] - cn.x);
    S.extF[1] = J.k_trans * (S.cant[1] - cn.y);
    S.extF[2] = J.k_trans * (S.cant[2] - cn.z);
    for (int d = 0; d < 3; ++d) S.cant[d] = J.cant0[d] + xt * J.pullVec[d];
    for (int k = 0; k < J.npulled; ++k) {
        J.ext[J.pulled[k]] = make_float4(S.extF[0], S.extF[1], S.extF[2], 2.0f);
        J.bflag[J.pulled[k]] = 2;
    }
}

// mask: 1 indentation tip, 2 pulling cantilever, 4 heating schedule.
// One 64-thread workgroup per job; lane 0 does the scalar work.
__global__ __launch_bounds__(64) void k_job_update(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st,
                                                   int njobs, int mask) {
    const int jb = blockIdx.x;
    if (jb >= njobs) return;
    const DevJob& J = jobs[jb];
    DevJobState& S = st[jb];
    const long long s = S.step;
    if (s >= J.end_step || s % J.stride != 0) return;
    if ((mask & 1) && (J.flags & JF_INDENTATION)) tip_update(J, S, s);
    if (threadIdx.x != 0) return;
    if ((mask & 2) && (J.flags & JF_PULL_UPDATER)) pull_update(J, S, s);
    if ((mask & 4) && (J.flags & JF_HEATING) && s % J.tempFreq == 0) {
        const float temp = J.initialT + (float)(s / J.tempFreq) * J.deltaT;
        S.var = sqrtf(2.0f * temp * J.zeta / J.h);
    }
}

__global__ void k_job_advance(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st, int njobs, int nsteps) {
    const int jb = blockIdx.x * blockDim.x + threadIdx.x;
    if (jb >= njobs) return;
    const long long s = st[jb].step + nsteps;
    st[jb].step = s < jobs[jb].end_step ? s : jobs[jb].end_step;
}

// ---------------------------------------------------------------------------
// energies (output cadence only)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_energies(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                                  int job) {
    const DevJob& J = jobs[job];
    const int n = J.n;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;  // Verlet row of bead i
    if (r >= n) return;
    const int i = J.perm[r];
    const long long s = st[job].step;
    const float4* __restrict__ cur = J.coord[s & 1];
    const float4 c = cur[i];
    float pot = 0.0f;
    for (int k = 0; k < J.cov_cnt[i]; ++k) {
        const int j = J.cov_j[(size_t)k * n + i];
        const float r0 = J.cov_r0[(size_t)k * n + i];
        const float4 o = cur[j];
        const float x = o.x - c.x, y = o.y - c.y, z = o.z - c.z;
        float w = sqrtf(x * x + y * y + z * z);
        w = w - r0;
        pot = (float)((double)pot - ((double)J.ks / 2.0) * (double)J.R2 * (double)logf_series(1.0f - w * w / (J.R2)));
    }
    J.energies[i] = pot;
    pot = 0.0f;
    for (int k = 0; k < J.nat_cnt[i]; ++k) {
        const int j = J.nat_j[(size_t)k * n + i];
        const float2 par = J.nat_par[(size_t)k * n + i];
        const float4 o = cur[j];
        const float x = o.x - c.x, y = o.y - c.y, z = o.z - c.z;
        const float q = par.x / (x * x + y * y + z * z);
        float p6 = q * q * q;
        const float p12 = p6 * p6;
        p6 = -2.0f * p6;
        const float eh = (float)((double)(-par.y * par.x) / 12.0);
        pot += eh * (p6 + p12);
    }
    J.energies[n + i] = pot;
    pot = 0.0f;
    for (int k = 0; k < J.pair_cnt[r]; ++k) {
        const int j = list_at(J.pair_j, J.idx16, (size_t)k * J.pair_ld + r);
        const float4 o = cur[j];
        const float x = o.x - c.x, y = o.y - c.y, z = o.z - c.z;
        const float d2 = x * x + y * y + z * z;
        if (d2 < J.rc2) {
            const float q = J.a2 / d2;
            pot += J.el * q * q * q;
        }
    }
    J.energies[2 * n + i] = pot;
}

// Self-check of div_rn_fast against the IEEE division: every float bit
// pattern b in [lo, hi) (inside the fast range) with each numerator a[k];
// counts the results that differ.
__global__ void k_division_check(unsigned lo, unsigned hi, const float* __restrict__ a, int na,
                                 unsigned long long* __restrict__ bad) {
    unsigned long long mine = 0;
    for (unsigned u = lo + blockIdx.x * blockDim.x + threadIdx.x; u < hi; u += gridDim.x * blockDim.x) {
        const float b = __uint_as_float(u);
        for (int k = 0; k < na; ++k)
            mine += __float_as_uint(div_rn_fast(a[k], b)) != __float_as_uint(a[k] / b) ? 1ull : 0ull;
    }
    if (mine) atomicAdd(bad, mine);
}

hipError_t launch_division_check(unsigned lo, unsigned hi, const float* a, int na, unsigned long long* bad,
                                 hipStream_t stream) {
    hipLaunchKernelGGL(k_division_check, dim3(8192), dim3(256), 0, stream, lo, hi, a, na, bad);
    return hipGetLastError();
}

// Advances `count` draws of the Hybrid Taus state of beads [0, n) and
// writes either the raw uint32 outputs or the Gaussian triples.
__global__ void k_rng_probe(uint4* __restrict__ state, int n, int count, int gaussian, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 s = state[i];
    for (int k = 0; k < count; ++k) {
        if (gaussian) {
            const float3 g = gauss3(s);
            out[((size_t)i * count + k) * 3 + 0] = g.x;
            out[((size_t)i * count + k) * 3 + 1] = g.y;
            out[((size_t)i * count + k) * 3 + 2] = g.z;
        } else {
            const uint32_t u = hybrid_taus(s);
            reinterpret_cast<uint32_t*>(out)[(size_t)i * count + k] = u;
        }
    }
    state[i] = s;
}

// Plain streaming copy (16 B per lane per iteration): the reference point
// for HBM bandwidth and for scaling memory counters.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_copy(const float4* src, float4* dst, size_t n, hipStream_t stream) {
    hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, stream, src, dst, n);
    return hipGetLastError();
}

static int grid_for(int ntiles, int max_blocks) {
    int b = (ntiles + kTilesPerBlock - 1) / kTilesPerBlock;
    if (max_blocks > 0 && b > max_blocks) b = max_blocks;
    if (b >= 8) b = (b + 7) / 8 * 8;  // whole XCD groups
    return b < 1 ? 1 : b;
}

// k_step_split over every tile (its workgroups of other jobs exit at once);
// G: the jobs above the LDS limit, gathering from global memory.
template <bool FO, bool G>
static void launch_split(int R, const DevJob* jobs, const DevJobState* st, const int2* tiles, int ntiles,
                         int lds_beads, int kstep, float4* out, hipStream_t stream) {
    const size_t lds = G ? 0 : (size_t)lds_beads * sizeof(float4);
    if (R == 4)
        hipLaunchKernelGGL((k_step_split<FO, 4, G>), dim3(ntiles * 16), dim3(1024), lds, stream, jobs, st, tiles,
                           ntiles, kstep, out);
    else if (R == 8)
        hipLaunchKernelGGL((k_step_split<FO, 8, G>), dim3(ntiles * 8), dim3(1024), lds, stream, jobs, st, tiles,
                           ntiles, kstep, out);
    else
        hipLaunchKernelGGL((k_step_split<FO, 16, G>), dim3(ntiles * 4), dim3(1024), lds, stream, jobs, st, tiles,
                           ntiles, kstep, out);
}

hipError_t launch_step(const DevJob* jobs, const DevJobState* st, const int2* tiles, int ntiles, int any_global,
                       const int2* units, const int4* segs, int nunits, int lds_beads, int kstep, int max_blocks,
                       hipStream_t stream, const StepShape& shape) {
    if (ntiles <= 0) return hipSuccess;
    if (any_global && shape.split_global)
        launch_split<false, true>(shape.split_global, jobs, st, tiles, ntiles, 0, kstep, nullptr, stream);
    else if (any_global)
        hipLaunchKernelGGL(k_step<false>, dim3(grid_for(ntiles, max_blocks)), dim3(256), 0, stream, jobs, st, tiles,
                           ntiles, kstep, nullptr);
    if (nunits > 0 && shape.split_rows) {
        launch_split<false, false>(shape.split_rows, jobs, st, tiles, ntiles, lds_beads, kstep, nullptr, stream);
    } else if (nunits > 0) {
        const dim3 g(nunits), b(shape.unit_rows);
        const size_t lds = (size_t)lds_beads * sizeof(float4);
        if (shape.pair_batch == 4)
            hipLaunchKernelGGL((k_step_lds<false, 4>), g, b, lds, stream, jobs, st, units, segs, nunits, kstep, nullptr);
        else if (shape.pair_batch == 8)
            hipLaunchKernelGGL((k_step_lds<false, 8>), g, b, lds, stream, jobs, st, units, segs, nunits, kstep, nullptr);
        else
            hipLaunchKernelGGL((k_step_lds<false, 6>), g, b, lds, stream, jobs, st, units, segs, nunits, kstep, nullptr);
    }
    return hipGetLastError();
}

hipError_t launch_forces(const DevJob* jobs, const DevJobState* st, const int2* tiles, int ntiles, int any_global,
                         const int2* units, const int4* segs, int nunits, int lds_beads, float4* out,
                         hipStream_t stream, const StepShape& shape) {
    if (ntiles <= 0) return hipSuccess;
    if (any_global && shape.split_global)
        launch_split<true, true>(shape.split_global, jobs, st, tiles, ntiles, 0, 0, out, stream);
    else if (any_global)
        hipLaunchKernelGGL(k_step<true>, dim3(grid_for(ntiles, 0)), dim3(256), 0, stream, jobs, st, tiles, ntiles, 0,
                           out);
    if (nunits > 0 && shape.split_rows)
        launch_split<true, false>(shape.split_rows, jobs, st, tiles, ntiles, lds_beads, 0, out, stream);
    else if (nunits > 0)
        hipLaunchKernelGGL((k_step_lds<true, 6>), dim3(nunits), dim3(kLdsUnit), (size_t)lds_beads * sizeof(float4),
                           stream, jobs, st, units, segs, nunits, 0, out);
    return hipGetLastError();
}

hipError_t launch_possible(const DevJob* jobs, DevJobState* st, const int2* tiles, int ntiles, int njobs,
                           int cell_words, int any_global, const int2* sunits, const int4* segs, int nsunits,
                           int lds_beads, int force, hipStream_t stream, const StepShape& shape) {
    if (ntiles <= 0) return hipSuccess;
    const int use_cells = cell_words > 0;
    if (any_global || nsunits <= 0)
        hipLaunchKernelGGL(k_possible, dim3(grid_for(ntiles, 0)), dim3(256), 0, stream, jobs, st, tiles, ntiles,
                           force, use_cells, nsunits > 0);
    if (nsunits > 0 && shape.split_rows)
        hipLaunchKernelGGL(k_possible_split, dim3(ntiles * 4), dim3(1024), (size_t)lds_beads * sizeof(float4), stream,
                           jobs, st, tiles, ntiles, force, use_cells);
    else if (nsunits > 0)
        hipLaunchKernelGGL(k_possible_lds, dim3(nsunits), dim3(kLdsUnit), (size_t)lds_beads * sizeof(float4), stream,
                           jobs, st, sunits, segs, nsunits, force, use_cells);
    if (use_cells) {
        hipLaunchKernelGGL(k_cells_setup, dim3(njobs), dim3(1024), 0, stream, jobs, st, njobs, force);
        hipLaunchKernelGGL(k_cells_count, dim3(grid_for(ntiles, 0)), dim3(256), 0, stream, jobs, st, tiles, ntiles,
                           force);
        hipLaunchKernelGGL(k_cells_scan, dim3(njobs), dim3(1024), 0, stream, jobs, st, njobs, force);
        hipLaunchKernelGGL(k_cells_fill, dim3(grid_for(ntiles, 0)), dim3(256), 0, stream, jobs, st, tiles, ntiles,
                           force);
        int blocks = 2 * ntiles;  // one wavefront per half tile
        if (blocks >= 8) blocks = (blocks + 7) / 8 * 8;
        hipLaunchKernelGGL(k_possible_cells, dim3(blocks), dim3(64), (size_t)(cell_words + (cell_words + 63) / 64) * 8,
                           stream, jobs, st, tiles, ntiles, force, cell_words);
    }
    return hipGetLastError();
}

hipError_t launch_verlet(const DevJob* jobs, DevJobState* st, const int2* tiles, int ntiles, int any_idx32,
                         int any_global, const int2* units, int nunits, const int2* sunits, const int4* segs,
                         int nsunits, int lds_beads, int force, hipStream_t stream, const StepShape& shape) {
    if (ntiles <= 0) return hipSuccess;
    if (nunits > 0)
        hipLaunchKernelGGL(k_sort_rows, dim3(nunits), dim3(kLdsUnit), 0, stream, jobs, st, units, nunits,
                           shape.sort_every, force);
    int blocks = ntiles;
    if (blocks >= 8) blocks = (blocks + 7) / 8 * 8;
    const size_t lds = any_idx32 ? 64 * VerletLds<int>::stride * sizeof(int)
                                 : 64 * VerletLds<unsigned short>::stride * sizeof(unsigned short);
    if (any_global || nsunits <= 0)
        hipLaunchKernelGGL(k_verlet, dim3(blocks), dim3(64 * kVerletWaves), lds, stream, jobs, st, tiles, ntiles,
                           force, nsunits > 0);
    if (nsunits > 0 && shape.split_rows) {
        const size_t lds = (size_t)kVerletRowsBytes + (size_t)lds_beads * 16;
        hipLaunchKernelGGL(k_verlet_split, dim3(ntiles * 4), dim3(1024), lds, stream, jobs, st, tiles, ntiles, force);
    } else if (nsunits > 0) {
        const size_t lds = 2 * (size_t)kVerletRowsBytes + (size_t)lds_beads * 16;
        if (shape.verlet_chunk == 4)
            hipLaunchKernelGGL(k_verlet_lds<4>, dim3(nsunits), dim3(1024), lds, stream, jobs, st, sunits, segs, nsunits,
                               force);
        else
            hipLaunchKernelGGL(k_verlet_lds<8>, dim3(nsunits), dim3(1024), lds, stream, jobs, st, sunits, segs, nsunits,
                               force);
    }
    return hipGetLastError();
}

hipError_t launch_job_update(const DevJob* jobs, DevJobState* st, int njobs, hipStream_t stream, int mask) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_job_update, dim3(njobs), dim3(64), 0, stream, jobs, st, njobs, mask);
    return hipGetLastError();
}

hipError_t launch_job_advance(const DevJob* jobs, DevJobState* st, int njobs, int nsteps, hipStream_t stream) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_job_advance, dim3((njobs + 63) / 64), dim3(64), 0, stream, jobs, st, njobs, nsteps);
    return hipGetLastError();
}

hipError_t launch_energies(const DevJob* jobs, const DevJobState* st, int job, int n, hipStream_t stream) {
    hipLaunchKernelGGL(k_energies, dim3((n + 255) / 256), dim3(256), 0, stream, jobs, st, job);
    return hipGetLastError();
}

hipError_t launch_rng_probe(uint4* state, int n, int count, int gaussian, float* out, hipStream_t stream) {
    hipLaunchKernelGGL(k_rng_probe, dim3((n + 255) / 256), dim3(256), 0, stream, state, n, count, gaussian, out);
    return hipGetLastError();
}

const void* step_kernel_symbol() { return reinterpret_cast<const void*>(&k_step<false>); }

}  // namespace sop
