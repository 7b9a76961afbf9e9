// kernels.hip -- CDNA4 kernels of the SOP Langevin hot path.
//
// k_step       one launch per step for every admitted job: FENE + native
//              LJ 12-6 + repulsive LJ (+ indentation walls) gathered per bead
//              over slot-major lists, then the Hybrid-Taus noise, the
//              fixed/pulled-bead rules and the overdamped Euler update.
//              Reference bodies: src/Potentials/covalent_kernel.cu:28-51,
//              native_kernel.cu:27-52, pairs_kernel.cu:33-59,
//              indentation_kernel.cu:71-128, src/Integrators/langevin_kernel.cu:17-58.
// k_possible   possible-pairs list: same-job beads j != i, ascending j, with
//              sqrtf(d^2) < threshold, excluding covalent and native partners
//              (src/Updaters/possiblepairlist_kernel.cu:18-71).
// k_verlet     Verlet list: the possible list filtered by sqrtf(d^2) <
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

// ---------------------------------------------------------------------------
// fused step
// ---------------------------------------------------------------------------

// Repulsive LJ over a bead's slot-major Verlet entries, in list order.  The
// list words are streamed with non-temporal loads (each is used once per
// step) and fetched one batch of eight ahead: while batch b is computed the
// indices of batch b+1 and the coordinates of batch b are in flight.
template <typename IdxT>
__device__ __forceinline__ void pair_loop(const IdxT* __restrict__ pj, int np, int n, const float4* __restrict__ cin,
                                          const float4 c, float rc2, float a2, float m6, float& fx, float& fy,
                                          float& fz) {
#define SOP_PAIR(o)                                                  \
    {                                                                \
        const float x = (o).x - c.x, y = (o).y - c.y, z = (o).z - c.z; \
        const float d2 = x * x + y * y + z * z;                      \
        if (d2 < rc2) {                                              \
            float q = a2 / d2;                                       \
            q = q * q;                                               \
            q = m6 * q * q;                                          \
            fx += x * q;                                             \
            fy += y * q;                                             \
            fz += z * q;                                             \
        }                                                            \
    }
    int k = 0;
    if (np >= 8) {
        int jn[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) jn[u] = (int)__builtin_nontemporal_load(pj + (size_t)u * n);
        for (; k + 8 <= np; k += 8) {
            float4 o[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) o[u] = cin[jn[u]];
            if (k + 16 <= np) {
#pragma unroll
                for (int u = 0; u < 8; ++u) jn[u] = (int)__builtin_nontemporal_load(pj + (size_t)(k + 8 + u) * n);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) SOP_PAIR(o[u]);
        }
    }
    for (; k < np; ++k) {
        const float4 o = cin[(int)__builtin_nontemporal_load(pj + (size_t)k * n)];
        SOP_PAIR(o);
    }
#undef SOP_PAIR
}
__global__ __launch_bounds__(256) void k_step(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                              const int2* __restrict__ tiles, int ntiles, int kstep) {
    const int lane = threadIdx.x & 63;
    const TileRange tr = xcd_tiles(ntiles);
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int2 td = tiles[t];
        const DevJob& J = jobs[td.x];
        const long long s = st[td.x].step + kstep;
        if (s >= J.end_step) continue;
        const int i = td.y + lane;
        const int n = J.n;
        if (i >= n) continue;
        const float4* __restrict__ cin = J.coord[s & 1];
        float4* __restrict__ cout = J.coord[(s + 1) & 1];
        const float4 c = cin[i];
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        if (J.flags & JF_EXTRA_FORCE) {
            const float4 e = J.fextra[i];
            fx = e.x; fy = e.y; fz = e.z;
            J.fextra[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        // FENE (covalent_kernel.cu:36-48)
        {
            const int nc = J.cov_cnt[i];
            const int* __restrict__ cj = J.cov_j + i;
            const float* __restrict__ cr = J.cov_r0 + i;
            for (int k = 0; k < nc; ++k) {
                const int j = cj[(size_t)k * n];
                const float r0 = cr[(size_t)k * n];
                const float4 o = cin[j];
                const float x = o.x - c.x, y = o.y - c.y, z = o.z - c.z;
                float w = sqrtf(x * x + y * y + z * z);
                w = w - r0;
                w = J.ks * w / ((w + r0) * (1.0f - w * w / J.R2));
                fx += w * x;
                fy += w * y;
                fz += w * z;
            }
        }
        // native LJ 12-6 (native_kernel.cu:36-49), gathers four ahead
        {
            const int nn = J.nat_cnt[i];
            const int* __restrict__ nj = J.nat_j + i;
            const float2* __restrict__ npar = J.nat_par + i;
#define SOP_NAT(o, par)                                                 \
    {                                                                   \
        const float x = (o).x - c.x, y = (o).y - c.y, z = (o).z - c.z;     \
        float q = (par).x / (x * x + y * y + z * z);                    \
        const float p6 = q * q * q;                                     \
        q = q * (p6 * p6 - p6) * (par).y;                               \
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
                    pp[u] = npar[(size_t)(k + u) * n];
                }
                float4 o[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) o[u] = cin[jj[u]];
#pragma unroll
                for (int u = 0; u < 4; ++u) SOP_NAT(o[u], pp[u]);
            }
            for (; k < nn; ++k) {
                const float2 par = npar[(size_t)k * n];
                const float4 o = cin[nj[(size_t)k * n]];
                SOP_NAT(o, par);
            }
#undef SOP_NAT
        }
        // repulsive LJ over the Verlet list (pairs_kernel.cu:41-56)
        if (J.idx16)
            pair_loop(static_cast<const unsigned short*>(J.pair_j) + i, J.pair_cnt[i], n, cin, c, J.rc2, J.a2, J.m6,
                      fx, fy, fz);
        else
            pair_loop(static_cast<const int*>(J.pair_j) + i, J.pair_cnt[i], n, cin, c, J.rc2, J.a2, J.m6, fx, fy,
                      fz);
        const DevJobState& S = st[td.x];
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
            J.tipF[i] = fin;
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
        // hydrodynamic jobs: noise and forces for the HI integrator
        // (integrateTea_prepare, bdhitea_kernel.cu:33-65); the update itself
        // runs in hi.hip once every bead's forces are known
        if (J.flags & JF_HI) {
            uint4 rs = J.rng[i];
            const float var = S.var;
            const float3 g = gauss3(rs);
            J.rng[i] = rs;
            J.rforce[i] = make_float4(g.x * var, g.y * var, g.z * var, 0.0f);
            if (J.flags & JF_PULLING) {
                const float4 e = J.ext[i];
                if (e.w == 1.0f) {
                    fx = 0.0f; fy = 0.0f; fz = 0.0f;
                }
                if (e.w == 2.0f) {
                    fx += e.x; fy += e.y; fz += e.z;
                }
            }
            J.mforce[i] = make_float4(fx, fy, fz, 0.0f);
            continue;
        }
        // Langevin step (langevin_kernel.cu:24-56)
        if (!(J.flags & JF_MINIMIZATION)) {
            uint4 rs = J.rng[i];
            const float var = S.var;
            const float3 g = gauss3(rs);
            J.rng[i] = rs;
            fx += g.x * var;
            fy += g.y * var;
            fz += g.z * var;
        }
        if (J.flags & JF_PULLING) {
            const float4 e = J.ext[i];
            if (e.w == 1.0f) {
                fx = 0.0f; fy = 0.0f; fz = 0.0f;
            }
            if (e.w == 2.0f) {
                fx += e.x; fy += e.y; fz += e.z;
            }
        }
        const float m = J.hOverZeta;
        const float dx = m * fx, dy = m * fy, dz = m * fz;
        float4 cn = c;
        cn.x += dx;
        cn.y += dy;
        cn.z += dz;
        // the temperature accumulator is cleared at updater passes that fall
        // on a multiple of nav (output_manager.cu:124-126)
        const bool reset = (s % J.stride == 0) && (s % J.nav == 0);
        const float T0 = reset ? 0.0f : J.T[i];
        J.T[i] = T0 + J.tempNorm * (dx * dx + dy * dy + dz * dz);
        cout[i] = cn;
    }
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

// Possible pairs, one wavefront per bead: lanes test 64 consecutive j at a
// time; an order-preserving ballot/popcount compaction appends the accepted
// j to the bead's row (bead-major, pp_j[i*pp_cap + k]).  Exclusions are the
// bead's sorted covalent/native partners, turned into a 64-bit mask per
// chunk by a scalar walk.
__global__ __launch_bounds__(256) void k_possible(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st,
                                                  const int2* __restrict__ tiles, int ntiles, int force) {
    const int lane = threadIdx.x & 63;
    const TileRange tr = xcd_tiles(ntiles);
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int2 td = tiles[t];
        const DevJob& J = jobs[td.x];
        const long long s = st[td.x].step;
        if (!force && !due(J, s, J.pp_freq)) continue;
        const int n = J.n;
        const float4* __restrict__ cur = J.coord[s & 1];
        const float thr = J.pp_thr;
        const int cap = J.pp_cap;
        const int iend = min(td.y + 64, n);
        for (int i = td.y; i < iend; ++i) {
            const float4 r1 = cur[i];
            const int ne = J.excl_cnt[i];
            int pe = 0;
            int cnt = 0;
            const size_t row = (size_t)i * cap;
            for (int base = 0; base < n; base += 64) {
                const int j = base + lane;
                bool ok = j < n && j != i;
                if (ok) {
                    const float4 r2 = cur[j];
                    const float x = r2.x - r1.x, y = r2.y - r1.y, z = r2.z - r1.z;
                    ok = sqrtf(x * x + y * y + z * z) < thr;
                }
                unsigned long long ex = 0ull;
                while (pe < ne) {
                    const int e = J.excl_j[(size_t)pe * n + i];
                    if (e >= base + 64) break;
                    ex |= 1ull << (e - base);
                    ++pe;
                }
                ok = ok && !((ex >> lane) & 1ull);
                const unsigned long long m = __ballot(ok);
                if (ok) {
                    const int pos = cnt + __popcll(m & lanes_below(lane));
                    if (pos < cap) list_put(J.pp_j, J.idx16, row + pos, j);
                }
                cnt += __popcll(m);
            }
            if (lane == 0) {
                J.pp_cnt[i] = cnt < cap ? cnt : cap;
                if (cnt > cap) atomicOr(&st[td.x].overflow, 1);
            }
        }
    }
}

// Verlet list, one wavefront per bead: the bead's possible row filtered in
// stored order by the same ballot compaction into a bead-major scratch row;
// k_pairs_transpose then lays the rows out slot-major for k_step.
__global__ __launch_bounds__(256) void k_verlet(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st,
                                                const int2* __restrict__ tiles, int ntiles, int force) {
    const int lane = threadIdx.x & 63;
    const TileRange tr = xcd_tiles(ntiles);
    for (int t = tr.begin; t < tr.end; t += tr.step) {
        const int2 td = tiles[t];
        const DevJob& J = jobs[td.x];
        const long long s = st[td.x].step;
        if (!force && !due(J, s, J.pairs_freq)) continue;
        const int n = J.n;
        const float4* __restrict__ cur = J.coord[s & 1];
        const float cut = J.pl_cut;
        const int cap = J.pair_cap;
        const int iend = min(td.y + 64, n);
        for (int i = td.y; i < iend; ++i) {
            const float4 r1 = cur[i];
            const int npp = J.pp_cnt[i];
            const size_t inrow = (size_t)i * J.pp_cap, outrow = (size_t)i * cap;
            int cnt = 0;
            for (int base = 0; base < npp; base += 64) {
                const int k = base + lane;
                bool ok = k < npp;
                int j = 0;
                if (ok) {
                    j = list_at(J.pp_j, J.idx16, inrow + k);
                    const float4 r2 = cur[j];
                    const float x = r2.x - r1.x, y = r2.y - r1.y, z = r2.z - r1.z;
                    ok = sqrtf(x * x + y * y + z * z) < cut;
                }
                const unsigned long long m = __ballot(ok);
                if (ok) {
                    const int pos = cnt + __popcll(m & lanes_below(lane));
                    if (pos < cap) list_put(J.pair_tmp, J.idx16, outrow + pos, j);
                }
                cnt += __popcll(m);
            }
            if (lane == 0) {
                J.pair_cnt[i] = cnt < cap ? cnt : cap;
                if (cnt > cap) atomicOr(&st[td.x].overflow, 2);
            }
        }
    }
}

// Bead-major scratch rows -> slot-major Verlet list (pair_j[k*n + i]), one
// 64-bead tile per 64-thread workgroup, 64x64 blocks through LDS so both the
// reads and the writes are whole rows.
__global__ __launch_bounds__(64) void k_pairs_transpose(const DevJob* __restrict__ jobs,
                                                        const DevJobState* __restrict__ st,
                                                        const int2* __restrict__ tiles, int ntiles, int force) {
    __shared__ int buf[64][65];
    const int lane = threadIdx.x;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int2 td = tiles[t];
        const DevJob& J = jobs[td.x];
        const long long s = st[td.x].step;
        if (!force && !due(J, s, J.pairs_freq)) continue;
        const int n = J.n;
        const int cap = J.pair_cap;
        const int rows = min(64, n - td.y);
        int c = lane < rows ? J.pair_cnt[td.y + lane] : 0;
        int maxc = c;
        for (int off = 32; off > 0; off >>= 1) maxc = max(maxc, __shfl_xor(maxc, off));
        for (int kc = 0; kc < maxc; kc += 64) {
            __syncthreads();
            for (int b = 0; b < rows; ++b) {
                const int k = kc + lane;
                const int cb = __shfl(c, b);
                buf[b][lane] = k < cb ? list_at(J.pair_tmp, J.idx16, (size_t)(td.y + b) * cap + k) : 0;
            }
            __syncthreads();
            const int kend = min(64, maxc - kc);
            for (int k = 0; k < kend; ++k)
                if (lane < rows && kc + k < c) list_put(J.pair_j, J.idx16, (size_t)(kc + k) * n + td.y + lane, buf[lane][k]);
        }
    }
}

// ---------------------------------------------------------------------------
// per-job updaters at stride boundaries
// ---------------------------------------------------------------------------
__device__ void tip_update(const DevJob& J, DevJobState& S, long long s) {
    if (s == J.retraction) S.V = -S.V;
    if (s % J.nav != 0) return;
    float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
    for (int i = 0; i < J.n; ++i) {
        float4 v = J.tipF[i];
        f0 += v.x;
        f1 += v.y;
        f2 += v.z;
        v.x = 0.0f;
        v.y = 0.0f;
        v.z = 0.0f;
        J.tipF[i] = v;
    }
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
    for (int k = 0; k < J.npulled; ++k) J.ext[J.pulled[k]] = make_float4(S.extF[0], S.extF[1], S.extF[2], 2.0f);
}

// mask: 1 indentation tip, 2 pulling cantilever, 4 heating schedule
__global__ void k_job_update(const DevJob* __restrict__ jobs, DevJobState* __restrict__ st, int njobs, int mask) {
    const int jb = blockIdx.x * blockDim.x + threadIdx.x;
    if (jb >= njobs) return;
    const DevJob& J = jobs[jb];
    DevJobState& S = st[jb];
    const long long s = S.step;
    if (s >= J.end_step || s % J.stride != 0) return;
    if ((mask & 1) && (J.flags & JF_INDENTATION)) tip_update(J, S, s);
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
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
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
    for (int k = 0; k < J.pair_cnt[i]; ++k) {
        const int j = list_at(J.pair_j, J.idx16, (size_t)k * n + i);
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

// Per-bead forces only (no integration, no state change): the parity hook
// for the force gather.  Same arithmetic as k_step.
__global__ __launch_bounds__(256) void k_forces(const DevJob* __restrict__ jobs, const DevJobState* __restrict__ st,
                                                int job, float4* __restrict__ out) {
    const DevJob& J = jobs[job];
    const int n = J.n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long s = st[job].step;
    const float4* __restrict__ cin = J.coord[s & 1];
    const float4 c = cin[i];
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    for (int k = 0; k < J.cov_cnt[i]; ++k) {
        const int j = J.cov_j[(size_t)k * n + i];
        const float r0 = J.cov_r0[(size_t)k * n + i];
        const float4 o = cin[j];
        const float x = o.x - c.x, y = o.y - c.y, z = o.z - c.z;
        float w = sqrtf(x * x + y * y + z * z);
        w = w - r0;
        w = J.ks * w / ((w + r0) * (1.0f - w * w / J.R2));
        fx += w * x;
        fy += w * y;
        fz += w * z;
    }
    for (int k = 0; k < J.nat_cnt[i]; ++k) {
        const int j = J.nat_j[(size_t)k * n + i];
        const float2 par = J.nat_par[(size_t)k * n + i];
        const float4 o = cin[j];
        const float x = o.x - c.x, y = o.y - c.y, z = o.z - c.z;
        float q = par.x / (x * x + y * y + z * z);
        const float p6 = q * q * q;
        q = q * (p6 * p6 - p6) * par.y;
        fx += x * q;
        fy += y * q;
        fz += z * q;
    }
    for (int k = 0; k < J.pair_cnt[i]; ++k) {
        const int j = list_at(J.pair_j, J.idx16, (size_t)k * n + i);
        const float4 o = cin[j];
        const float x = o.x - c.x, y = o.y - c.y, z = o.z - c.z;
        const float d2 = x * x + y * y + z * z;
        if (d2 < J.rc2) {
            float q = J.a2 / d2;
            q = q * q;
            q = J.m6 * q * q;
            fx += x * q;
            fy += y * q;
            fz += z * q;
        }
    }
    if (J.flags & JF_INDENTATION) {
        const DevJobState& S = st[job];
        float x = c.x - S.tip[0], y = c.y - S.tip[1], z = c.z - S.tip[2];
        const float rr = sqrtf(x * x + y * y + z * z);
        float r2 = rr - J.tipR;
        r2 = 1.0f / r2;
        r2 = r2 * r2;
        float r6 = r2 * r2 * r2;
        float dw = J.tipAp * r6 + J.tipBp;
        dw = dw * r6 / (rr * (rr - J.tipR));
        const float dfx = x * dw, dfy = y * dw, dfz = z * dw;
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
    out[i] = make_float4(fx, fy, fz, 0.0f);
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

hipError_t launch_step(const DevJob* jobs, const DevJobState* st, const int2* tiles, int ntiles, int kstep,
                       int max_blocks, hipStream_t stream) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_step, dim3(grid_for(ntiles, max_blocks)), dim3(256), 0, stream, jobs, st, tiles, ntiles, kstep);
    return hipGetLastError();
}

hipError_t launch_possible(const DevJob* jobs, DevJobState* st, const int2* tiles, int ntiles, int force,
                           hipStream_t stream) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_possible, dim3(grid_for(ntiles, 0)), dim3(256), 0, stream, jobs, st, tiles, ntiles, force);
    return hipGetLastError();
}

hipError_t launch_verlet(const DevJob* jobs, DevJobState* st, const int2* tiles, int ntiles, int force,
                         hipStream_t stream) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_verlet, dim3(grid_for(ntiles, 0)), dim3(256), 0, stream, jobs, st, tiles, ntiles, force);
    hipLaunchKernelGGL(k_pairs_transpose, dim3(ntiles), dim3(64), 0, stream, jobs, st, tiles, ntiles, force);
    return hipGetLastError();
}

hipError_t launch_job_update(const DevJob* jobs, DevJobState* st, int njobs, hipStream_t stream, int mask) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_job_update, dim3((njobs + 63) / 64), dim3(64), 0, stream, jobs, st, njobs, mask);
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

hipError_t launch_forces(const DevJob* jobs, const DevJobState* st, int job, int n, float4* out, hipStream_t stream) {
    hipLaunchKernelGGL(k_forces, dim3((n + 255) / 256), dim3(256), 0, stream, jobs, st, job, out);
    return hipGetLastError();
}

hipError_t launch_rng_probe(uint4* state, int n, int count, int gaussian, float* out, hipStream_t stream) {
    hipLaunchKernelGGL(k_rng_probe, dim3((n + 255) / 256), dim3(256), 0, stream, state, n, count, gaussian, out);
    return hipGetLastError();
}

const void* step_kernel_symbol() { return reinterpret_cast<const void*>(&k_step); }

}  // namespace sop
