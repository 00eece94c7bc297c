// dsmc_kernels.hip -- the DSMC step on gfx950 (CDNA4), one wave per cell.
//
// Per step (System::step, system.cpp:22-47):
//   k_move            accelerate + move + walls + periodic, in place, in the
//                     current cell-sorted order; classifies every particle as
//                     staying in its cell or moving to another (mover list)
//   k_scan_*          new cell counts and their exclusive scans
//   k_bucket          movers into per-destination-cell buckets
//   k_gather_collide  per destination cell: stable gather (arrivals from lower
//                     old indices, stayers, arrivals from higher ones) into
//                     the other state buffer, NTC collisions in LDS,
//                     per-cell FP64 moments, write-out
// The stable order is the one of a stable counting sort keyed by cell
// (matlab/3d/sortParticles.m:11-63): inside a cell, particles keep the order
// of their indices before the step.
#include <hip/hip_runtime.h>
#include "dsmc_device.h"
#include "dsmc_kernels.h"

namespace dsmc {

static constexpr int WAVE = 64;

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ void cell_coords(const Member &M, uint32_t lc, int &cx, int &cy, int &cz) {
    cz = (int)(lc % (uint32_t)M.cpa[2]);
    cy = (int)((lc / (uint32_t)M.cpa[2]) % (uint32_t)M.cpa[1]);
    cx = (int)(lc / ((uint32_t)M.cpa[1] * (uint32_t)M.cpa[2]));
}

__device__ __forceinline__ void load_tile(const Member &M, const uint8_t *types, uint8_t *tile,
                                          int cx, int cy, int cz, VoxelView &V) {
    V.types = types;
    V.tile = tile;
    V.tx0 = cx * M.b[0] - 1;
    V.ty0 = cy * M.b[1] - 1;
    V.tz0 = cz * M.b[2] - 1;
    V.TX = M.b[0] + 2;
    V.TY = M.b[1] + 2;
    V.TZ = M.b[2] + 2;
    V.n0 = M.n[0]; V.n1 = M.n[1]; V.n2 = M.n[2];
    int tot = V.TX * V.TY * V.TZ;
    for (int i = (int)threadIdx.x; i < tot; i += WAVE) {
        int c = i % V.TZ, b = (i / V.TZ) % V.TY, a = i / (V.TY * V.TZ);
        int x = wrapi(V.tx0 + a, M.n[0]), y = wrapi(V.ty0 + b, M.n[1]), z = wrapi(V.tz0 + c, M.n[2]);
        tile[i] = types[((size_t)x * M.n[1] + y) * M.n[2] + z];
    }
}

// ------------------------------------------------------------------ move
__global__ __launch_bounds__(64) void k_move(MoveArgs a) {
    const uint32_t c = blockIdx.x + a.cell_begin;
    const uint32_t lane = threadIdx.x;
    const uint32_t s0 = a.start[c], s1 = a.start[c + 1];
    if (s0 == s1) {
        if (lane == 0) a.leave[c] = 0;
        return;
    }
    const Member M = a.members[a.cell_member[c]];
    const uint8_t *types = a.types + M.vox0;
    const float *frames = a.frames + 9 * M.vox0;
    __shared__ uint8_t tile[MAX_TILE];
    int cx, cy, cz;
    cell_coords(M, c - M.cell0, cx, cy, cz);
    VoxelView V;
    load_tile(M, types, tile, cx, cy, cz, V);
    __syncthreads();
    const uint32_t step = *a.step;
    MoveCounters cnt = {0, 0, 0, {0, 0, 0}};
    uint32_t leavers = 0;
    // software pipeline: the next chunk's loads are in flight while this
    // chunk is moved
    uint32_t p = s0 + lane;
    float ox = 0, oy = 0, oz = 0, vx = 0, vy = 0, vz = 0;
    uint32_t vp = 0;
    if (p < s1) {
        ox = a.ox[p]; oy = a.oy[p]; oz = a.oz[p];
        vx = a.vx[p]; vy = a.vy[p]; vz = a.vz[p];
        vp = a.vox[p];
    }
    for (uint32_t base = s0; base < s1; base += WAVE) {
        const uint32_t pn = p + WAVE;
        float nox = 0, noy = 0, noz = 0, nvx = 0, nvy = 0, nvz = 0;
        uint32_t nvp = 0;
        if (pn < s1) {
            nox = a.ox[pn]; noy = a.oy[pn]; noz = a.oz[pn];
            nvx = a.vx[pn]; nvy = a.vy[pn]; nvz = a.vz[pn];
            nvp = a.vox[pn];
        }
        const bool valid = p < s1;
        uint32_t newc = c;
        if (valid) {
            int ix, iy, iz;
            unpack_vox(vp, ix, iy, iz);
            uint32_t nhit = 0;
            int32_t last = move_particle(M, V, frames, step, p, ix, iy, iz, ox, oy, oz, vx, vy, vz,
                                         cnt, nhit);
            a.ox[p] = ox; a.oy[p] = oy; a.oz[p] = oz;
            a.vx[p] = vx; a.vy[p] = vy; a.vz[p] = vz;
            a.vox[p] = pack_vox(ix, iy, iz);
            newc = cell_of(M, ix, iy, iz);
            if (a.rec_nhit) {
                a.rec_nhit[p] = (int32_t)nhit;
                a.rec_last[p] = last;
            }
        }
        // movers are listed in the cell's own index range, leavers first:
        // no global counter, no returning atomic
        const bool mover = valid && newc != c;
        const uint64_t mm = __ballot(mover);
        if (mover) {
            const uint32_t slot = s0 + leavers + mbcnt(mm);
            a.mv_idx[slot] = p;
            a.mv_cell[slot] = newc;
            atomicAdd(&a.arrive[newc], 1u);
        }
        leavers += (uint32_t)__popcll(mm);
        p = pn;
        ox = nox; oy = noy; oz = noz; vx = nvx; vy = nvy; vz = nvz; vp = nvp;
    }
    // wave-uniform already; reduce the per-lane counters
    uint32_t wh = wave_sum(cnt.wall_hits), rs = wave_sum(cnt.reseeds), er = wave_sum(cnt.errors);
    int32_t p0 = wave_sum(cnt.periodic[0]), p1 = wave_sum(cnt.periodic[1]), p2 = wave_sum(cnt.periodic[2]);
    if (lane == 0) {
        a.leave[c] = leavers;
        if (wh) atomicAdd(&a.counters[CNT_WALL], (unsigned long long)wh);
        if (rs) atomicAdd(&a.counters[CNT_RESEED], (unsigned long long)rs);
        if (er) atomicAdd(&a.counters[CNT_ERROR], (unsigned long long)er);
        if (p0) atomicAdd(&a.counters[CNT_PER0], (unsigned long long)(long long)p0);
        if (p1) atomicAdd(&a.counters[CNT_PER1], (unsigned long long)(long long)p1);
        if (p2) atomicAdd(&a.counters[CNT_PER2], (unsigned long long)(long long)p2);
    }
}

// Every particle of an unsorted buffer becomes a mover into its cell (used
// once after seeding or loading a state: the start array is all zero).
__global__ __launch_bounds__(256) void k_classify_all(ClassifyArgs a) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < a.n; p += gridDim.x * blockDim.x) {
        const Member M = a.members[a.part_member[p]];
        int ix, iy, iz;
        unpack_vox(a.vox[p], ix, iy, iz);
        uint32_t c = cell_of(M, ix, iy, iz);
        a.mv_idx[p] = p;
        a.mv_cell[p] = c;
        atomicAdd(&a.arrive[c], 1u);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.mv_count = a.n;
}

// ------------------------------------------------------------------ scan
// Per block of SCAN_TILE cells: sum of new counts and arrivals.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(ScanArgs a) {
    __shared__ unsigned long long red[SCAN_THREADS / WAVE];
    const uint32_t base = blockIdx.x * SCAN_TILE;
    unsigned long long s = 0;
    for (uint32_t i = threadIdx.x; i < SCAN_TILE; i += SCAN_THREADS) {
        uint32_t c = base + i;
        if (c < a.ncells) {
            uint32_t n_old = a.start[c + 1] - a.start[c];
            uint32_t n_new = n_old - a.leave[c] + a.arrive[c];
            s += ((unsigned long long)a.arrive[c] << 32) | n_new;
        }
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < SCAN_THREADS / WAVE; w++) t += red[w];
        a.block_sums[blockIdx.x] = t;
    }
}

// One block: exclusive scan of the block sums (n_blocks <= SCAN_THREADS*8).
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_top(ScanArgs a) {
    __shared__ unsigned long long buf[SCAN_THREADS];
    const uint32_t nb = a.n_blocks;
    const uint32_t per = (nb + SCAN_THREADS - 1) / SCAN_THREADS;
    const uint32_t b0 = threadIdx.x * per;
    unsigned long long local = 0;
    for (uint32_t i = 0; i < per; i++)
        if (b0 + i < nb) local += a.block_sums[b0 + i];
    buf[threadIdx.x] = local;
    __syncthreads();
    for (uint32_t off = 1; off < SCAN_THREADS; off <<= 1) {
        unsigned long long v = threadIdx.x >= off ? buf[threadIdx.x - off] : 0ull;
        __syncthreads();
        buf[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned long long run = threadIdx.x ? buf[threadIdx.x - 1] : 0ull;
    for (uint32_t i = 0; i < per; i++)
        if (b0 + i < nb) {
            unsigned long long v = a.block_sums[b0 + i];
            a.block_sums[b0 + i] = run;
            run += v;
        }
    if (threadIdx.x == SCAN_THREADS - 1) {
        a.start_new[a.ncells] = (uint32_t)(buf[SCAN_THREADS - 1] & 0xffffffffull);
        a.bstart[a.ncells] = (uint32_t)(buf[SCAN_THREADS - 1] >> 32);
    }
}

// Per block: exclusive scan of its cells with the block offset.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(ScanArgs a) {
    __shared__ unsigned long long buf[SCAN_THREADS];
    const uint32_t base = blockIdx.x * SCAN_TILE;
    constexpr int PER = SCAN_TILE / SCAN_THREADS;
    unsigned long long vals[PER];
    unsigned long long local = 0;
#pragma unroll
    for (int i = 0; i < PER; i++) {
        uint32_t c = base + threadIdx.x * PER + i;
        unsigned long long v = 0;
        if (c < a.ncells) {
            uint32_t n_old = a.start[c + 1] - a.start[c];
            uint32_t n_new = n_old - a.leave[c] + a.arrive[c];
            v = ((unsigned long long)a.arrive[c] << 32) | n_new;
        }
        vals[i] = v;
        local += v;
    }
    buf[threadIdx.x] = local;
    __syncthreads();
    for (uint32_t off = 1; off < SCAN_THREADS; off <<= 1) {
        unsigned long long v = threadIdx.x >= off ? buf[threadIdx.x - off] : 0ull;
        __syncthreads();
        buf[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned long long run = a.block_sums[blockIdx.x] + (threadIdx.x ? buf[threadIdx.x - 1] : 0ull);
#pragma unroll
    for (int i = 0; i < PER; i++) {
        uint32_t c = base + threadIdx.x * PER + i;
        if (c < a.ncells) {
            a.start_new[c] = (uint32_t)(run & 0xffffffffull);
            a.bstart[c] = (uint32_t)(run >> 32);
            a.bfill[c] = 0;
        }
        run += vals[i];
    }
}

// Flat mover list (all particles of an unsorted buffer).
__global__ __launch_bounds__(256) void k_bucket(BucketArgs a) {
    const uint32_t n = *a.mv_count;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t c = a.mv_cell[i];
        uint32_t slot = atomicAdd(&a.bfill[c], 1u);
        a.bucket[a.bstart[c] + slot] = a.mv_idx[i];
    }
}

// Movers listed per source cell (k_move): cell c's leavers occupy
// mv_idx/mv_cell[start[c] .. start[c] + leave[c]).
__global__ __launch_bounds__(256) void k_bucket_cells(BucketArgs a) {
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < a.ncells; c += gridDim.x * blockDim.x) {
        const uint32_t s0 = a.start[c], nl = a.leave[c];
        for (uint32_t k = 0; k < nl; k++) {
            uint32_t d = a.mv_cell[s0 + k];
            uint32_t slot = atomicAdd(&a.bfill[d], 1u);
            a.bucket[a.bstart[d] + slot] = a.mv_idx[s0 + k];
        }
    }
}

// ------------------------------------------------------------------ gather + collide
// Velocity storage for one cell: LDS when it fits, else the cell's range of
// the output arrays (global memory, owned by this wave).
struct VelStore {
    float *x, *y, *z;
    uint32_t *mark;   // LDS conflict marks (nullptr on the global path)
};

__device__ __forceinline__ void ntc_collide(const Member &M, uint32_t cell_local, uint32_t step,
                                            uint32_t n, int pairs, double vrm, VelStore S,
                                            double &crm_out, uint32_t &accepted_out,
                                            uint32_t &cand_out) {
    const uint32_t lane = threadIdx.x;
    double crm = vrm;
    uint32_t acc = 0, cand = 0;
    if (S.mark)
        for (uint32_t i = lane; i < n; i += WAVE) S.mark[i] = 0xffffffffu;
    __syncthreads();
    while ((int)cand < pairs) {
        const uint32_t k = cand + lane;
        const bool valid = (int)k < pairs;
        const uint4 o = philox(TAG_SELECT, step, cell_local, k, M.key_lo, M.key_hi);
        const uint32_t i0 = __umulhi(o.x, n);
        const uint32_t i1 = (i0 + 1u + __umulhi(o.y, n - 1u)) % n;
        const float dx = S.x[i0] - S.x[i1];
        const float dy = S.y[i0] - S.y[i1];
        const float dz = S.z[i0] - S.z[i1];
        const float vrel = sqrtf(dx * dx + dy * dy + dz * dz);
        const bool accept = valid && (double)vrel > u01d(o.z) * vrm;
        bool conflict;
        if (S.mark) {
            if (accept) {
                atomicMin(&S.mark[i0], lane);
                atomicMin(&S.mark[i1], lane);
            }
            __syncthreads();
            conflict = valid && (S.mark[i0] < lane || S.mark[i1] < lane);
        } else {
            conflict = false;
            for (uint32_t j = 0; j < WAVE; j++) {
                uint32_t a0 = __shfl(i0, (int)j, WAVE), a1 = __shfl(i1, (int)j, WAVE);
                int acj = __shfl((int)accept, (int)j, WAVE);
                if (j < lane && acj && (a0 == i0 || a0 == i1 || a1 == i0 || a1 == i1)) conflict = true;
            }
            conflict = conflict && valid;
        }
        const uint64_t cmask = __ballot(conflict);
        const uint32_t remaining = (uint32_t)(pairs - (int)cand);
        const uint32_t f = cmask ? (uint32_t)(__ffsll((long long)cmask) - 1) :
                                   (remaining < (uint32_t)WAVE ? remaining : (uint32_t)WAVE);
        const bool commit = lane < f;
        if (commit && (double)vrel > crm) crm = (double)vrel;
        const bool do_scatter = commit && accept;
        acc += (uint32_t)__popcll(__ballot(do_scatter));
        if (do_scatter) {
            const uint4 q = philox(TAG_SCATTER, step, cell_local, k, M.key_lo, M.key_hi);
            const float cos_th = 1.0f - 2.0f * u01f(q.x);
            const float sin_th = sqrtf(1.0f - cos_th * cos_th);
            float sp, cp;
            sincospif(2.0f * u01f(q.y), &sp, &cp);
            const float vrx = vrel * cos_th;
            const float vry = vrel * sin_th * cp;
            const float vrz = vrel * sin_th * sp;
            const float cmx = 0.5f * (S.x[i0] + S.x[i1]);
            const float cmy = 0.5f * (S.y[i0] + S.y[i1]);
            const float cmz = 0.5f * (S.z[i0] + S.z[i1]);
            S.x[i0] = cmx + 0.5f * vrx; S.y[i0] = cmy + 0.5f * vry; S.z[i0] = cmz + 0.5f * vrz;
            S.x[i1] = cmx - 0.5f * vrx; S.y[i1] = cmy - 0.5f * vry; S.z[i1] = cmz - 0.5f * vrz;
        }
        __syncthreads();
        if (S.mark && accept) {
            S.mark[i0] = 0xffffffffu;
            S.mark[i1] = 0xffffffffu;
        }
        __syncthreads();
        cand += f;
    }
    crm_out = wave_max(crm);
    accepted_out = acc;
    cand_out = cand;
}

extern __shared__ float dyn_lds[];

__global__ __launch_bounds__(64) void k_gather_collide(GatherArgs a) {
    const uint32_t c = blockIdx.x + a.cell_begin;
    const uint32_t lane = threadIdx.x;
    const uint32_t ns = a.start_new[c], ne = a.start_new[c + 1];
    const uint32_t n = ne - ns;
    const uint32_t os = a.start[c], oe = a.start[c + 1];
    const uint32_t bs = a.bstart[c], be = a.bstart[c + 1];
    const uint32_t nb = be - bs;
    const uint32_t n_stay = n - nb;
    const uint32_t K = a.lds_cap;
    const bool in_lds = n <= K;
    float *lx = dyn_lds, *ly = dyn_lds + K, *lz = dyn_lds + 2 * K;
    uint32_t *lmark = (uint32_t *)(dyn_lds + 3 * K);   // also the arrival-key buffer
    VelStore S;
    if (in_lds) { S.x = lx; S.y = ly; S.z = lz; S.mark = lmark; }
    else { S.x = a.vx2 + ns; S.y = a.vy2 + ns; S.z = a.vz2 + ns; S.mark = nullptr; }

    // ---- arrivals: rank by index before the step (stable order)
    uint32_t n_low = 0;
    const bool keys_lds = nb <= K;
    if (nb > WAVE && keys_lds) {
        for (uint32_t i = lane; i < nb; i += WAVE) lmark[i] = a.bucket[bs + i];
        __syncthreads();
    }
    for (uint32_t i = lane; i < nb; i += WAVE) n_low += a.bucket[bs + i] < os ? 1u : 0u;
    n_low = wave_sum(n_low);

    const uint32_t step = *a.step;
    const Member M = a.members[a.cell_member[c]];

    // ---- stayers: stable compaction of the old range by wave ballot; all
    // seven arrays of a chunk are loaded together and the next chunk is in
    // flight while this one is placed
    uint32_t run = 0;
    uint32_t p = os + lane;
    float ox = 0, oy = 0, oz = 0, vx = 0, vy = 0, vz = 0;
    uint32_t vp = 0;
    if (p < oe) {
        vp = a.vox[p];
        ox = a.ox[p]; oy = a.oy[p]; oz = a.oz[p];
        vx = a.vx[p]; vy = a.vy[p]; vz = a.vz[p];
    }
    for (uint32_t base = os; base < oe; base += WAVE) {
        const uint32_t pn = p + WAVE;
        float nox = 0, noy = 0, noz = 0, nvx = 0, nvy = 0, nvz = 0;
        uint32_t nvp = 0;
        if (pn < oe) {
            nvp = a.vox[pn];
            nox = a.ox[pn]; noy = a.oy[pn]; noz = a.oz[pn];
            nvx = a.vx[pn]; nvy = a.vy[pn]; nvz = a.vz[pn];
        }
        const bool valid = p < oe;
        bool stay = false;
        if (valid) {
            int ix, iy, iz;
            unpack_vox(vp, ix, iy, iz);
            stay = cell_of(M, ix, iy, iz) == c;
        }
        const uint64_t sm = __ballot(stay);
        if (stay) {
            const uint32_t L = n_low + run + mbcnt(sm);
            const uint32_t d = ns + L;
            a.ox2[d] = ox; a.oy2[d] = oy; a.oz2[d] = oz;
            a.vox2[d] = vp;
            S.x[L] = vx; S.y[L] = vy; S.z[L] = vz;
            if (a.rec_perm) a.rec_perm[d] = (int32_t)p;
        }
        run += (uint32_t)__popcll(sm);
        p = pn;
        ox = nox; oy = noy; oz = noz; vx = nvx; vy = nvy; vz = nvz; vp = nvp;
    }
    // ---- arrivals
    for (uint32_t base = 0; base < nb; base += WAVE) {
        const uint32_t i = base + lane;
        const bool valid = i < nb;
        uint32_t key = valid ? a.bucket[bs + i] : 0xffffffffu;
        uint32_t rank = 0;
        if (nb <= WAVE) {
            for (uint32_t j = 0; j < nb; j++) {
                uint32_t kj = (uint32_t)__shfl((int)key, (int)j, WAVE);
                rank += kj < key ? 1u : 0u;
            }
        } else if (keys_lds) {
            for (uint32_t j = 0; j < nb; j++) rank += lmark[j] < key ? 1u : 0u;
        } else {
            for (uint32_t j = 0; j < nb; j++) rank += a.bucket[bs + j] < key ? 1u : 0u;
        }
        if (valid) {
            const uint32_t L = rank < n_low ? rank : n_stay + rank;
            const uint32_t d = ns + L;
            const uint32_t q = key;
            a.ox2[d] = a.ox[q]; a.oy2[d] = a.oy[q]; a.oz2[d] = a.oz[q];
            a.vox2[d] = a.vox[q];
            S.x[L] = a.vx[q]; S.y[L] = a.vy[q]; S.z[L] = a.vz[q];
            if (a.rec_perm) a.rec_perm[d] = (int32_t)q;
        }
    }
    __syncthreads();

    // ---- collide (Cell::prepare + Cell::collide, cell.cpp:42-111)
    const bool active = a.coeff[c] > 0.0;
    if (a.do_collide && active) {
        const double vrm = a.vr_max[c];
        const double select = a.coeff[c] * (double)n * (double)((int)n - 1) * vrm + a.rest[c];
        const int pairs = (int)round(select);
        const double rest = select - (double)pairs;
        double crm = vrm;
        uint32_t acc = 0, cand = 0;
        if (n >= 2) ntc_collide(M, c - M.cell0, step, n, pairs, vrm, S, crm, acc, cand);
        if (lane == 0) {
            a.rest[c] = rest;
            if (n >= 2) a.vr_max[c] = crm;
            a.stat_pairs[c] = pairs;
            a.stat_acc[c] = (int32_t)acc;
            if (acc) atomicAdd(&a.counters[CNT_COLL], (unsigned long long)acc);
            if (cand) atomicAdd(&a.counters[CNT_CAND], (unsigned long long)cand);
            if (!in_lds) atomicAdd(&a.counters[CNT_OVERFLOW], 1ull);
        }
        __syncthreads();
    }

    // ---- moments (FP64) and write-out
    double m[7] = {0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = lane; i < n; i += WAVE) {
        const float x = S.x[i], y = S.y[i], z = S.z[i];
        if (!in_lds) { /* already in place */ }
        else { a.vx2[ns + i] = x; a.vy2[ns + i] = y; a.vz2[ns + i] = z; }
        const double dx = x, dy = y, dz = z;
        m[0] += 1.0; m[1] += dx; m[2] += dy; m[3] += dz;
        m[4] += dx * dx; m[5] += dy * dy; m[6] += dz * dz;
    }
    if (a.do_sample) {
#pragma unroll
        for (int j = 0; j < 7; j++) m[j] = wave_sum(m[j]);
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < 7; j++) a.moments[(size_t)j * a.ncells + c] += m[j];
        }
    }
}

// ------------------------------------------------------------------ seeding
__global__ __launch_bounds__(256) void k_seed(SeedArgs a) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < a.n; p += gridDim.x * blockDim.x) {
        const uint32_t m = a.part_member[p];
        const Member M = a.members[m];
        const uint32_t i = p - a.member_first[m];      // creation index inside the member
        const uint8_t *types = a.types + M.vox0;
        uint4 o = philox(TAG_INIT_VEL, 0u, i, 0u, M.key_lo, M.key_hi);
        float r01 = sqrtf(-2.0f * logf(u01f(o.x)));
        float s01, c01, s2, c2;
        sincospif(2.0f * u01f(o.y), &s01, &c01);
        float r2 = sqrtf(-2.0f * logf(u01f(o.z)));
        sincospif(2.0f * u01f(o.w), &s2, &c2);
        a.vx[p] = (r01 * c01) * M.s_gas;
        a.vy[p] = (r01 * s01) * M.s_gas;
        a.vz[p] = (r2 * c2) * M.s_gas;
        int ix = 0, iy = 0, iz = 0;
        float ox = 0.5f, oy = 0.5f, oz = 0.5f;
        bool ok = false;
        for (uint32_t attempt = 0; attempt < 100000u && !ok; attempt++) {
            uint4 q = philox(TAG_INIT_POS, 0u, i, attempt, M.key_lo, M.key_hi);
            uint32_t lin = __umulhi(q.x, M.nvox);
            if (types[lin] != V_EMPTY) continue;
            uint32_t nyz = (uint32_t)M.n[1] * (uint32_t)M.n[2];
            ix = (int)(lin / nyz);
            iy = (int)((lin / (uint32_t)M.n[2]) % (uint32_t)M.n[1]);
            iz = (int)(lin % (uint32_t)M.n[2]);
            ox = u01f(q.y); oy = u01f(q.z); oz = u01f(q.w);
            ok = true;
        }
        if (!ok) atomicAdd(&a.counters[CNT_ERROR], 1ull);
        a.ox[p] = ox; a.oy[p] = oy; a.oz[p] = oz;
        a.vox[p] = pack_vox(ix, iy, iz);
    }
}

// ------------------------------------------------------------------ small kernels
__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_begin_step(uint32_t *step, uint32_t *mv_count) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *step += 1u;
        *mv_count = 0u;
    }
}

// Deterministic FP64 totals over all particles: per-block partials.
__global__ __launch_bounds__(256) void k_totals_partial(const float *vx, const float *vy, const float *vz,
                                                        uint32_t n, double *partial) {
    double t[4] = {0, 0, 0, 0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double x = vx[i], y = vy[i], z = vz[i];
        t[0] += x; t[1] += y; t[2] += z; t[3] += x * x + y * y + z * z;
    }
    __shared__ double red[4][4];
#pragma unroll
    for (int j = 0; j < 4; j++) t[j] = wave_sum(t[j]);
    if ((threadIdx.x & 63) == 0)
        for (int j = 0; j < 4; j++) red[threadIdx.x >> 6][j] = t[j];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int j = 0; j < 4; j++) partial[blockIdx.x * 4 + j] = red[0][j] + red[1][j] + red[2][j] + red[3][j];
}

// ------------------------------------------------------------------ launchers
#define LAUNCH_CHECK() do { hipError_t e = hipGetLastError(); if (e != hipSuccess) return e; } while (0)

hipError_t launch_move(const MoveArgs &a, uint32_t ncells, hipStream_t s) {
    if (ncells == 0) return hipSuccess;
    hipLaunchKernelGGL(k_move, dim3(ncells), dim3(64), 0, s, a);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_classify_all(const ClassifyArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_classify_all, dim3(2048), dim3(256), 0, s, a);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_scan(const ScanArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_scan_reduce, dim3(a.n_blocks), dim3(SCAN_THREADS), 0, s, a);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SCAN_THREADS), 0, s, a);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_down, dim3(a.n_blocks), dim3(SCAN_THREADS), 0, s, a);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_bucket(const BucketArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_bucket, dim3(2048), dim3(256), 0, s, a);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_bucket_cells(const BucketArgs &a, hipStream_t s) {
    uint32_t blocks = (a.ncells + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bucket_cells, dim3(blocks), dim3(256), 0, s, a);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_gather_collide(const GatherArgs &a, uint32_t ncells, hipStream_t s) {
    if (ncells == 0) return hipSuccess;
    size_t lds = (size_t)a.lds_cap * 16;
    hipLaunchKernelGGL(k_gather_collide, dim3(ncells), dim3(64), lds, s, a);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_seed(const SeedArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_seed, dim3(2048), dim3(256), 0, s, a);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_begin_step(uint32_t *step, uint32_t *mv_count, hipStream_t s) {
    hipLaunchKernelGGL(k_begin_step, dim3(1), dim3(64), 0, s, step, mv_count);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_totals_partial(const float *vx, const float *vy, const float *vz, uint32_t n,
                                 double *partial, uint32_t nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_totals_partial, dim3(nblocks), dim3(256), 0, s, vx, vy, vz, n, partial);
    LAUNCH_CHECK();
    return hipSuccess;
}

}  // namespace dsmc
