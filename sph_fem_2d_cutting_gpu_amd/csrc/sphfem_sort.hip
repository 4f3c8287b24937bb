// Cell list of the locked Green grid (grid.cu:25-42, grid_gpu_green.cu:13-213) for gfx950:
//  * FP64 cell hash with C truncation; blanked particles keep their previous hash;
//  * stable LSD radix sort of (hash, index), 8-bit digits, 4096-key tiles, ranking inside a
//    tile by 64-lane ballot multisplit (stable: wave order, round order, lane order);
//  * cell boundaries from neighbouring sorted keys via a lane shuffle, and the gather of the
//    particle arrays (all 20 for the reference-form sort, or fused with init/predict/EOS/
//    artificial stress for the step).
#include "sphfem_physics.h"

#pragma clang fp contract(off)

namespace sphfem {

constexpr int SORT_BS = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_BS * SORT_ITEMS;   // 4096 keys per block
constexpr int RADIX = 256;
constexpr int WAVE = 64;

__device__ __forceinline__ int hash_of(const DevConsts &c, sphfem_double2 pos) {
    int ix = (int)((pos.x - c.geom.bbmin_x) / c.geom.dx);
    int iy = (int)((pos.y - c.geom.bbmin_y) / c.geom.dx);
    return ix * c.geom.ny + iy;
}

// grid.cu:25-42; also clears the cells the previous step occupied (their starts were the only
// non -1 entries) and, optionally, counts digit 0 of the new keys per 4096-key tile.
__global__ void __launch_bounds__(SORT_BS) k_hash(DevConsts c, const double *__restrict__ blanked,
                                                  const sphfem_double2 *__restrict__ pos,
                                                  const int *__restrict__ old_hash, int *__restrict__ hash_out,
                                                  int *__restrict__ cell_start, int clear_old, int N,
                                                  int num_cell, int *__restrict__ hist0, int nblocks,
                                                  int *__restrict__ flags) {
    __shared__ int lh[RADIX];
    for (int d = threadIdx.x; d < RADIX; d += SORT_BS) lh[d] = 0;
    __syncthreads();
    const int base = blockIdx.x * SORT_TILE;
    for (int it = 0; it < SORT_ITEMS; it++) {
        int i = base + it * SORT_BS + threadIdx.x;
        if (i >= N) break;
        int oh = old_hash[i];
        if (clear_old && oh >= 0 && oh < num_cell) cell_start[oh] = -1;
        int h = oh;
        if (blanked[i] != 1.) h = hash_of(c, pos[i]);
        if (h < 0 || h >= num_cell) {
            atomicOr(&flags[0], 1);
            h = h < 0 ? 0 : num_cell - 1;
        }
        hash_out[i] = h;
        if (hist0) atomicAdd(&lh[h & (RADIX - 1)], 1);
    }
    if (hist0) {
        __syncthreads();
        for (int d = threadIdx.x; d < RADIX; d += SORT_BS) hist0[d * nblocks + blockIdx.x] = lh[d];
    }
}

// `mode`, when given, points at the device re-sort mode: the full-sort kernels run only in mode 1
__global__ void __launch_bounds__(SORT_BS) k_upsweep(const int *__restrict__ keys, int N, int shift,
                                                     int *__restrict__ hist, int nblocks, const int *mode) {
    if (mode && *mode != 1) return;
    __shared__ int lh[RADIX];
    for (int d = threadIdx.x; d < RADIX; d += SORT_BS) lh[d] = 0;
    __syncthreads();
    const int base = blockIdx.x * SORT_TILE;
    for (int it = 0; it < SORT_ITEMS; it++) {
        int i = base + it * SORT_BS + threadIdx.x;
        if (i >= N) break;
        atomicAdd(&lh[(keys[i] >> shift) & (RADIX - 1)], 1);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < RADIX; d += SORT_BS) hist[d * nblocks + blockIdx.x] = lh[d];
}

// one block per digit: exclusive scan of that digit's per-tile counts, total into digit_tot
__global__ void __launch_bounds__(SORT_BS) k_scan_digits(int *__restrict__ hist, int nblocks,
                                                         int *__restrict__ digit_tot, const int *mode) {
    if (mode && *mode != 1) return;
    __shared__ int part[SORT_BS];
    int *row = hist + (long)blockIdx.x * nblocks;
    int per = (nblocks + SORT_BS - 1) / SORT_BS;
    int lo = threadIdx.x * per;
    int hi = min(lo + per, nblocks);
    int s = 0;
    for (int j = lo; j < hi; j++) s += row[j];
    part[threadIdx.x] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan over 256 partials
    for (int off = 1; off < SORT_BS; off <<= 1) {
        int v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = threadIdx.x ? part[threadIdx.x - 1] : 0;
    for (int j = lo; j < hi; j++) {
        int v = row[j];
        row[j] = run;
        run += v;
    }
    if (threadIdx.x == SORT_BS - 1) digit_tot[blockIdx.x] = part[SORT_BS - 1];
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const unsigned lane = __lane_id();
    return lane ? ((~0ull) >> (64 - lane)) : 0ull;
}

// stable scatter of one pass; vals_in == nullptr means the values are the input positions
__global__ void __launch_bounds__(SORT_BS) k_downsweep(const int *__restrict__ keys_in,
                                                       const int *__restrict__ vals_in,
                                                       int *__restrict__ keys_out, int *__restrict__ vals_out,
                                                       int N, int shift, const int *__restrict__ hist,
                                                       int nblocks, const int *__restrict__ digit_tot,
                                                       int *__restrict__ flags, const int *mode) {
    if (mode && *mode != 1) return;
    __shared__ int block_off[RADIX];
    __shared__ int wcount[SORT_BS / WAVE][RADIX];
    __shared__ int scan[RADIX];
    const int tid = threadIdx.x;
    const int w = tid / WAVE;
    const int lane = tid % WAVE;
    // digit bases: exclusive scan over the 256 digit totals
    scan[tid] = digit_tot[tid];
    for (int k = 0; k < SORT_BS / WAVE; k++) wcount[k][tid] = 0;
    __syncthreads();
    for (int off = 1; off < RADIX; off <<= 1) {
        int v = tid >= off ? scan[tid - off] : 0;
        __syncthreads();
        scan[tid] += v;
        __syncthreads();
    }
    block_off[tid] = (tid ? scan[tid - 1] : 0) + hist[(long)tid * nblocks + blockIdx.x];
    __syncthreads();

    int key[SORT_ITEMS], val[SORT_ITEMS], rank[SORT_ITEMS];
    const int base = blockIdx.x * SORT_TILE + w * (SORT_ITEMS * WAVE);
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; r++) {
        int i = base + r * WAVE + lane;
        bool valid = i < N;
        int k = valid ? keys_in[i] : 0;
        key[r] = k;
        val[r] = valid ? (vals_in ? vals_in[i] : i) : 0;
        int d = (k >> shift) & (RADIX - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint64_t m = __ballot((d >> b) & 1);
            peers &= ((d >> b) & 1) ? m : ~m;
        }
        int before = __popcll(peers & lt);
        int cnt = __popcll(peers);
        int old = 0;
        if (valid) old = wcount[w][d];
        __builtin_amdgcn_wave_barrier();
        if (valid && before == 0) wcount[w][d] = old + cnt;
        __builtin_amdgcn_wave_barrier();
        rank[r] = valid ? old + before : -1;
    }
    __syncthreads();
    // prefix of each digit over the waves of this block
    {
        int d = tid;
        int run = 0;
        for (int k = 0; k < SORT_BS / WAVE; k++) {
            int v = wcount[k][d];
            wcount[k][d] = run;
            run += v;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; r++) {
        if (rank[r] < 0) continue;
        int d = (key[r] >> shift) & (RADIX - 1);
        int dst = block_off[d] + wcount[w][d] + rank[r];
        if ((unsigned)dst >= (unsigned)N) {   // cannot happen for a consistent histogram
            atomicOr(&flags[0], 2);
            continue;
        }
        keys_out[dst] = key[r];
        vals_out[dst] = val[r];
    }
}

// The full stable LSD radix sort of the step's fallback (mode 1) by one workgroup, for up to
// SPHFEM_SORT_ONE_BLOCK_MAX particles: one launch, which returns at once on the incremental path,
// in place of the per-pass upsweep / digit-scan / downsweep launches.  The passes, the digit
// offsets and the stable order within a pass (tiles in increasing index; within a tile waves, item
// rows and lanes in index order, as k_downsweep ranks them) are those of the multi-block sort, so
// the permutation is the same.  Unsorted keys in keys_in (values: their positions); passes ping-pong
// between (ka, va) and (kb, vb) and the result ends in (ka, va).  kb may be keys_in: pass 0 reads
// keys_in and writes ka, later passes no longer need it.  Every lane runs every loop (bounds are
// uniform), so the ballots see whole wavefronts.
// It also does k_resort_scan's work first (the changed counts of the hash pass -> re-sort mode) and,
// on the incremental path, k_resort_compact's and k_resort_rank's (the changed particles gathered
// into LDS from the changed bits, then ranked): the step's sort is then three launches (hash pass,
// this, scatter) instead of five.
__device__ __forceinline__ int changed_before(const unsigned long long *__restrict__ cmask,
                                              const int *__restrict__ wlocal, const int *__restrict__ bprefix,
                                              int p, int N, int m);

__global__ void __launch_bounds__(SORT_BS) k_sort_one_block(const int *keys_in, int *ka, int *va, int *kb, int *vb,
                                                            int N, int passes, int *flags, int *bprefix, int nb,
                                                            int force_full, int cap,
                                                            const unsigned long long *cmask, const int *wlocal,
                                                            unsigned long long *csorted, int4 *cdest) {
    __shared__ int off[RADIX];
    __shared__ int cnt[RADIX];
    __shared__ int wcount[SORT_BS / WAVE][RADIX];
    __shared__ unsigned long long comp[SPHFEM_RESORT_CAP];
    __shared__ int s_mode, s_m;
    const int tid = threadIdx.x;
    {   // exclusive scan of the tiles' changed counts, as k_resort_scan
        int carry = 0;
        for (int c0 = 0; c0 < nb; c0 += SORT_BS) {
            const int j = c0 + tid;
            const int v = j < nb ? bprefix[j] : 0;
            cnt[tid] = v;
            __syncthreads();
            for (int o = 1; o < SORT_BS; o <<= 1) {
                const int u = tid >= o ? cnt[tid - o] : 0;
                __syncthreads();
                cnt[tid] += u;
                __syncthreads();
            }
            if (j < nb) bprefix[j] = carry + cnt[tid] - v;
            carry += cnt[SORT_BS - 1];
            __syncthreads();
        }
        if (tid == 0) {
            bprefix[nb] = carry;
            flags[4] = carry;
            s_mode = (force_full || carry > cap) ? 1 : 2;
            s_m = carry;
            flags[3] = s_mode;
        }
        __threadfence();
        __syncthreads();
    }
    if (s_mode == 2) {
        // the changed particles in position order (k_resort_compact), as (new key, position)
        const int m = s_m;
        const int words = (N + 63) >> 6;
        for (int w = tid; w < words; w += SORT_BS) {
            const unsigned long long bits = cmask[w];
            unsigned long long rest = bits;
            while (rest) {
                const int b = __ffsll((long long)rest) - 1;
                rest &= rest - 1;
                const int i = (w << 6) + b;
                const int k = changed_before(cmask, wlocal, bprefix, i, N, m);
                if ((unsigned)k < (unsigned)m)
                    comp[k] = ((unsigned long long)(unsigned)keys_in[i] << 32) | (unsigned)i;
                else
                    atomicOr(&flags[0], 2);
            }
        }
        __syncthreads();
        // their ranks and destinations (k_resort_rank), the old keys being ka (the sorted keys)
        for (int k = tid; k < m; k += SORT_BS) {
            const unsigned long long ck = comp[k];
            int r = 0;
            for (int q = 0; q < m; q++) r += comp[q] < ck;
            int lo = 0, hi = N;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                const unsigned long long cm = ((unsigned long long)(unsigned)ka[mid] << 32) | (unsigned)mid;
                if (cm < ck) lo = mid + 1;
                else hi = mid;
            }
            const int dest = r + lo - changed_before(cmask, wlocal, bprefix, lo, N, m);
            if ((unsigned)dest >= (unsigned)N || r >= m) {
                atomicOr(&flags[0], 2);
                continue;
            }
            csorted[r] = ck;
            cdest[r] = make_int4(dest, (int)(ck >> 32), (int)(ck & 0xffffffffu), 0);
        }
        return;
    }
    if (s_mode != 1) return;
    const int w = tid / WAVE;
    const int lane = tid % WAVE;
    const uint64_t lt = lanemask_lt();
    const int *kin = keys_in, *vin = nullptr;
    int *kout = ka, *vout = va;
    for (int ps = 0; ps < passes; ps++) {
        const int shift = 8 * ps;
        // digit totals of the pass, then their exclusive scan
        cnt[tid] = 0;
        __syncthreads();
        for (int i = tid; i < N; i += SORT_BS) atomicAdd(&cnt[(kin[i] >> shift) & (RADIX - 1)], 1);
        __syncthreads();
        for (int o = 1; o < RADIX; o <<= 1) {
            const int v = tid >= o ? cnt[tid - o] : 0;
            __syncthreads();
            cnt[tid] += v;
            __syncthreads();
        }
        off[tid] = tid ? cnt[tid - 1] : 0;
        __syncthreads();
        for (int t0 = 0; t0 < N; t0 += SORT_TILE) {
            for (int k = 0; k < SORT_BS / WAVE; k++) wcount[k][tid] = 0;
            __syncthreads();
            int key[SORT_ITEMS], val[SORT_ITEMS], rank[SORT_ITEMS];
            const int base = t0 + w * (SORT_ITEMS * WAVE);
#pragma unroll
            for (int r = 0; r < SORT_ITEMS; r++) {
                const int i = base + r * WAVE + lane;
                const bool valid = i < N;
                const int k = valid ? kin[i] : 0;
                key[r] = k;
                val[r] = valid ? (vin ? vin[i] : i) : 0;
                const int d = (k >> shift) & (RADIX - 1);
                uint64_t peers = __ballot(valid);
#pragma unroll
                for (int b = 0; b < 8; b++) {
                    const uint64_t m = __ballot((d >> b) & 1);
                    peers &= ((d >> b) & 1) ? m : ~m;
                }
                const int before = __popcll(peers & lt);
                const int c = __popcll(peers);
                int old = 0;
                if (valid) old = wcount[w][d];
                __builtin_amdgcn_wave_barrier();
                if (valid && before == 0) wcount[w][d] = old + c;
                __builtin_amdgcn_wave_barrier();
                rank[r] = valid ? old + before : -1;
            }
            __syncthreads();
            {   // per digit: prefix over the waves of the tile, and the tile's count
                int run = 0;
                for (int k = 0; k < SORT_BS / WAVE; k++) {
                    const int v = wcount[k][tid];
                    wcount[k][tid] = run;
                    run += v;
                }
                cnt[tid] = run;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < SORT_ITEMS; r++) {
                if (rank[r] < 0) continue;
                const int d = (key[r] >> shift) & (RADIX - 1);
                const int dst = off[d] + wcount[w][d] + rank[r];
                if ((unsigned)dst >= (unsigned)N) {   // cannot happen for consistent counts
                    atomicOr(&flags[0], 2);
                    continue;
                }
                kout[dst] = key[r];
                vout[dst] = val[r];
            }
            __syncthreads();
            off[tid] += cnt[tid];
            __syncthreads();
        }
        // this pass's output, written by all waves, is the next pass's input
        __threadfence();
        __syncthreads();
        kin = kout;
        vin = vout;
        kout = kout == ka ? kb : ka;
        vout = vout == va ? vb : va;
    }
    if (kin != ka) {
        for (int i = tid; i < N; i += SORT_BS) {
            ka[i] = kin[i];
            va[i] = vin[i];
        }
    }
}

__global__ void k_fill_int(int *p, long n, int v) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

void launch_fill_int(sphfem_ctx *ctx, int *ptr, long n, int value) {
    if (n <= 0) return;
    int blocks = (int)std::min<long>(nblk(n, 256), 8192);
    KTimer t(ctx, "fill_int");
    k_fill_int<<<blocks, 256, 0, ctx->stream>>>(ptr, n, value);
}

void launch_hash(sphfem_ctx *ctx, sphfem_grid *g, const sphfem_particles &p, bool clear_old, int *hist0) {
    // writes into p.hash (the reference's in-place layout); the fused step passes hist0
    (void)hist0;
    int N = (int)p.N;
    int nb = nblk(N, SORT_TILE);
    KTimer t(ctx, "hash");
    k_hash<<<nb, SORT_BS, 0, ctx->stream>>>(ctx->c, p.blanked, p.pos, p.hash, p.hash, g->cell_start,
                                            clear_old ? 1 : 0, N, g->num_cell, nullptr, nb, ctx->d_flags);
}

// Radix sort of (p.hash, index) leaving sorted keys in p.hash and the permutation in p.idx
// (in vals_out when given).
// If hash_with_hist0, the hash pass (which computes the keys) is run here with the digit-0
// histogram fused and writes the keys into the buffer the pass parity requires.
void launch_radix_sort(sphfem_ctx *ctx, sphfem_grid *g, sphfem_particles &p, bool hash_with_hist0, int *vals_out) {
    const int N = (int)p.N;
    const int nb = nblk(N, SORT_TILE);
    const int passes = (g->key_bits + 7) / 8;
    int *kin, *vin = nullptr;
    // choose buffers so that the last pass writes p.hash / p.idx
    int *kbuf[2] = {p.hash, g->keys_alt};
    int *vbuf[2] = {vals_out ? vals_out : p.idx, g->vals_alt};
    int src = (passes % 2 == 1) ? 1 : 0;   // buffer holding the unsorted keys
    if (hash_with_hist0) {
        KTimer t(ctx, "hash");
        k_hash<<<nb, SORT_BS, 0, ctx->stream>>>(ctx->c, p.blanked, p.pos, p.hash, kbuf[src], g->cell_start,
                                                g->first_sort ? 0 : 1, N, g->num_cell, g->hist, nb,
                                                ctx->d_flags);
    } else if (src == 1) {
        (void)hipMemcpyAsync(g->keys_alt, p.hash, sizeof(int) * (size_t)N, hipMemcpyDeviceToDevice, ctx->stream);
    }
    kin = kbuf[src];
    for (int ps = 0; ps < passes; ps++) {
        int shift = 8 * ps;
        int dst = 1 - src;
        if (!(ps == 0 && hash_with_hist0)) {
            KTimer t(ctx, "radix_upsweep");
            k_upsweep<<<nb, SORT_BS, 0, ctx->stream>>>(kin, N, shift, g->hist, nb, nullptr);
        }
        {
            KTimer t(ctx, "radix_scan");
            k_scan_digits<<<RADIX, SORT_BS, 0, ctx->stream>>>(g->hist, nb, g->digit_tot, nullptr);
        }
        {
            KTimer t(ctx, "radix_downsweep");
            k_downsweep<<<nb, SORT_BS, 0, ctx->stream>>>(kin, ps == 0 ? nullptr : vin, kbuf[dst], vbuf[dst], N,
                                                         shift, g->hist, nb, g->digit_tot, ctx->d_flags, nullptr);
        }
        kin = kbuf[dst];
        vin = vbuf[dst];
        src = dst;
    }
}

// ---------------------------------------------------------------------------------------------
// Incremental re-sort for the step.  After a step the particles are stored in the previous sorted
// order, so `old` (the previous sorted keys) is ascending.  A particle keeps its cell unless it
// crossed a cell boundary (rare at dt ~ 1e-4 of the cell crossing time).  The stable sort of the
// new keys with the previous position as tie-break is then the merge of
//   U: unchanged positions, already in (key, position) order, and
//   C: changed positions sorted by (new key, position):
//   dest(u) = u - changed_before(u) + |{c in C : (key_c, c) < (key_u, u)}|
//   dest(C[r]) = r + P - changed_before(P),  P = first p with (old[p], p) >= (key_c, c)
// which is exactly the result of the full stable radix sort (same keys, same order).
// The path is chosen on the device: the full radix sort runs on the first step and whenever more
// than `cap` particles changed cells.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ int changed_before(const unsigned long long *__restrict__ cmask,
                                              const int *__restrict__ wlocal, const int *__restrict__ bprefix,
                                              int p, int N, int m) {
    if (p >= N) return m;
    const unsigned long long bits = cmask[p >> 6];
    const int b = p & 63;
    return bprefix[p / SORT_TILE] + wlocal[p >> 6] + __popcll(b ? (bits & ((~0ull) >> (64 - b))) : 0ull);
}

// grid.cu:25-42 hash (as k_hash, with the digit-0 histogram for the full-sort fallback), plus the
// changed bit of every position, the changed count before each 64 positions within the tile
// (wlocal) and the tile's total (bcount)
// Workgroups past nblocks (tool_blocks of them, when a FEM tool is given) do the tool's step start
// instead (tool_begin_item), which shares no data with the hash pass: one launch fewer per step.
constexpr int HASH_BS = 1024;                       // 16 wavefronts per tile, 4 items per lane
constexpr int HASH_ITEMS = SORT_TILE / HASH_BS;
__global__ void __launch_bounds__(HASH_BS) k_hash_mark(DevConsts c, const double *__restrict__ blanked,
                                                       const sphfem_double2 *__restrict__ pos,
                                                       const int *__restrict__ old_hash, int *__restrict__ hash_out,
                                                       int *__restrict__ cell_start, int clear_old, int N,
                                                       int num_cell, int *__restrict__ hist0, int nblocks,
                                                       int *flags, unsigned long long *__restrict__ cmask,
                                                       int *__restrict__ wlocal, int *__restrict__ bcount,
                                                       sphfem_tool tool, int tool_blocks, double half_dt) {
    if ((int)blockIdx.x >= nblocks) {
        if (tool_blocks > 0)
            tool_begin_item(tool, half_dt, flags, ((int)blockIdx.x - nblocks) * HASH_BS + (int)threadIdx.x);
        return;
    }
    __shared__ int lh[RADIX];
    __shared__ int wc[SORT_TILE / 64];
    for (int d = threadIdx.x; d < RADIX; d += HASH_BS) lh[d] = 0;
    __syncthreads();
    const int base = blockIdx.x * SORT_TILE;
    // every load of the lane's items first, then the hashes: the pass waits for memory once
    int ohs[HASH_ITEMS];
    double bls[HASH_ITEMS];
    sphfem_double2 pss[HASH_ITEMS];
#pragma unroll
    for (int it = 0; it < HASH_ITEMS; it++) {
        const int i = base + it * HASH_BS + threadIdx.x;
        ohs[it] = i < N ? old_hash[i] : 0;
        bls[it] = i < N ? blanked[i] : 1.;
        pss[it] = i < N ? pos[i] : sphfem_double2{0., 0.};
    }
#pragma unroll
    for (int it = 0; it < HASH_ITEMS; it++) {
        const int i = base + it * HASH_BS + threadIdx.x;
        bool changed = false;
        if (i < N) {
            const int oh = ohs[it];
            int h = oh;
            if (bls[it] != 1.) h = hash_of(c, pss[it]);
            if (h < 0 || h >= num_cell) {
                atomicOr(&flags[0], 1);
                h = h < 0 ? 0 : num_cell - 1;
            }
            hash_out[i] = h;
            atomicAdd(&lh[h & (RADIX - 1)], 1);
            changed = h != oh;
            // a cell that keeps one of its particles stays occupied (the reorder rewrites its start);
            // one that all of them left is cleared by each of them
            if (clear_old && changed && oh >= 0 && oh < num_cell) cell_start[oh] = -1;
        }
        const unsigned long long bits = __ballot(changed);
        if (__lane_id() == 0) {
            wc[(i - base) >> 6] = __popcll(bits);
            if (i < N) cmask[i >> 6] = bits;
        }
    }
    __syncthreads();
    for (int d = threadIdx.x; d < RADIX; d += HASH_BS) hist0[d * nblocks + blockIdx.x] = lh[d];
    if (threadIdx.x == 0) {
        int run = 0;
        for (int k = 0; k < SORT_TILE / 64; k++) {
            const int g = (base >> 6) + k;
            if ((long)g * 64 < N) wlocal[g] = run;
            run += wc[k];
        }
        bcount[blockIdx.x] = run;
    }
}

// one block: exclusive prefix of the tile totals (in place, coalesced chunks), the total and the
// path decision
constexpr int RESORT_SCAN_BS = 1024;
__global__ void __launch_bounds__(RESORT_SCAN_BS) k_resort_scan(int *__restrict__ bprefix, int nb,
                                                                int *__restrict__ flags, int force_full, int cap) {
    __shared__ int part[RESORT_SCAN_BS];
    int carry = 0;
    for (int c0 = 0; c0 < nb; c0 += RESORT_SCAN_BS) {
        const int j = c0 + threadIdx.x;
        const int v = j < nb ? bprefix[j] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < RESORT_SCAN_BS; off <<= 1) {
            const int u = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += u;
            __syncthreads();
        }
        if (j < nb) bprefix[j] = carry + part[threadIdx.x] - v;
        carry += part[RESORT_SCAN_BS - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bprefix[nb] = carry;
        flags[4] = carry;
        flags[3] = (force_full || carry > cap) ? 1 : 2;
    }
}

// changed positions, in position order
__global__ void __launch_bounds__(256) k_resort_compact(const int *__restrict__ keys, int N,
                                                        const unsigned long long *__restrict__ cmask,
                                                        const int *__restrict__ wlocal,
                                                        const int *__restrict__ bprefix, int2 *__restrict__ clist,
                                                        const int *flags) {
    if (flags[3] != 2) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const unsigned long long bits = cmask[i >> 6];
    if (!((bits >> (i & 63)) & 1ull)) return;
    const int k = changed_before(cmask, wlocal, bprefix, i, N, 0);
    clist[k] = make_int2(keys[i], i);
}

// one block: rank of each changed position among the changed ones (counting in LDS; the composite
// keys are distinct) and its destination
__global__ void __launch_bounds__(RESORT_SCAN_BS) k_resort_rank(const int2 *__restrict__ clist,
                                                                const int *__restrict__ old_keys, int N,
                                                                const unsigned long long *__restrict__ cmask,
                                                                const int *__restrict__ wlocal,
                                                                const int *__restrict__ bprefix,
                                                                unsigned long long *__restrict__ csorted,
                                                                int4 *__restrict__ cdest, int *__restrict__ flags) {
    __shared__ unsigned long long comp[SPHFEM_RESORT_CAP];
    if (flags[3] != 2) return;
    const int m = flags[4];
    for (int k = threadIdx.x; k < m; k += RESORT_SCAN_BS) {
        const int2 e = clist[k];
        comp[k] = ((unsigned long long)(unsigned)e.x << 32) | (unsigned)e.y;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < m; k += RESORT_SCAN_BS) {
        const unsigned long long ck = comp[k];
        int r = 0;
        for (int q = 0; q < m; q++) r += comp[q] < ck;
        // P: first position p with (old[p], p) >= (key, position); (old[p], p) increases with p
        int lo = 0, hi = N;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const unsigned long long cm = ((unsigned long long)(unsigned)old_keys[mid] << 32) | (unsigned)mid;
            if (cm < ck) lo = mid + 1;
            else hi = mid;
        }
        const int dest = r + lo - changed_before(cmask, wlocal, bprefix, lo, N, m);
        if ((unsigned)dest >= (unsigned)N || r >= m) {
            atomicOr(&flags[0], 2);
            continue;
        }
        csorted[r] = ck;
        cdest[r] = make_int4(dest, (int)(ck >> 32), (int)(ck & 0xffffffffu), 0);
    }
}

// scatter: unchanged positions by the merge formula, changed ones from cdest
__global__ void __launch_bounds__(256) k_resort_scatter(const int *__restrict__ keys, int N,
                                                        const unsigned long long *__restrict__ cmask,
                                                        const int *__restrict__ wlocal,
                                                        const int *__restrict__ bprefix,
                                                        const unsigned long long *__restrict__ csorted,
                                                        const int4 *__restrict__ cdest, int *__restrict__ keys_out,
                                                        int *__restrict__ idx_out, int *__restrict__ flags) {
    if (flags[3] != 2) return;
    const int m = flags[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < m) {
        const int4 e = cdest[i];
        if ((unsigned)e.x < (unsigned)N) {
            keys_out[e.x] = e.y;
            idx_out[e.x] = e.z;
        }
    }
    if (i >= N) return;
    if ((cmask[i >> 6] >> (i & 63)) & 1ull) return;
    const int key = keys[i];
    const unsigned long long ck = ((unsigned long long)(unsigned)key << 32) | (unsigned)i;
    int lo = 0, hi = m;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (csorted[mid] < ck) lo = mid + 1;
        else hi = mid;
    }
    const int dest = i - changed_before(cmask, wlocal, bprefix, i, N, m) + lo;
    if ((unsigned)dest >= (unsigned)N) {
        atomicOr(&flags[0], 2);
        return;
    }
    keys_out[dest] = key;
    idx_out[dest] = i;
}

__global__ void k_copy_keys_if_full(const int *__restrict__ src, int *__restrict__ dst, int N, const int *mode) {
    if (*mode != 1) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) dst[i] = src[i];
}

static void launch_resort_radix(sphfem_ctx *ctx, sphfem_grid *g, sphfem_particles &p, int passes);

// Re-sort of the step: new keys and the permutation end in p.hash / p.idx (as launch_radix_sort);
// the full-sort fallback by one workgroup (one launch) when asked and N allows, else the multi-block
// radix sort
void launch_resort(sphfem_ctx *ctx, sphfem_grid *g, sphfem_particles &p, bool one_block_fallback,
                   sphfem_tool *tool) {
    const int N = (int)p.N;
    const int nb = nblk(N, SORT_TILE);
    int *flags = ctx->d_flags;
    {
        KTimer t(ctx, tool ? "hash_tool_begin" : "hash");
        const int tb = tool ? nblk(std::max(std::max(tool->n_nodes, tool->n_seg), 1), HASH_BS) : 0;
        k_hash_mark<<<nb + tb, HASH_BS, 0, ctx->stream>>>(ctx->c, p.blanked, p.pos, p.hash, g->keys_alt,
                                                          g->cell_start, g->first_sort ? 0 : 1, N, g->num_cell,
                                                          g->hist, nb, flags, g->cmask, g->wprefix, g->bprefix,
                                                          tool ? *tool : sphfem_tool{}, tb, 0.5 * ctx->c.dt);
    }
    // re-sort mode (full or incremental), then the full stable radix sort in mode 1: unsorted keys
    // are in keys_alt
    const int cap = std::min(g->resort_limit, SPHFEM_RESORT_CAP);
    const int force_full = (g->first_sort || cap <= 0) ? 1 : 0;
    const int passes = (g->key_bits + 7) / 8;
    if (one_block_fallback && N <= SPHFEM_SORT_ONE_BLOCK_MAX) {
        KTimer t(ctx, "resort_one_block");
        k_sort_one_block<<<1, SORT_BS, 0, ctx->stream>>>(g->keys_alt, p.hash, p.idx, g->keys_alt, g->vals_alt, N,
                                                         passes, flags, g->bprefix, nb, force_full, cap, g->cmask,
                                                         g->wprefix, g->csorted, g->cdest);
    } else {
        {
            KTimer t(ctx, "resort_scan");
            k_resort_scan<<<1, RESORT_SCAN_BS, 0, ctx->stream>>>(g->bprefix, nb, flags, force_full, cap);
        }
        launch_resort_radix(ctx, g, p, passes);
        // incremental path (mode 2)
        {
            KTimer t(ctx, "resort_compact");
            k_resort_compact<<<nblk(N, 256), 256, 0, ctx->stream>>>(g->keys_alt, N, g->cmask, g->wprefix,
                                                                    g->bprefix, g->clist, flags);
        }
        {
            KTimer t(ctx, "resort_rank");
            k_resort_rank<<<1, RESORT_SCAN_BS, 0, ctx->stream>>>(g->clist, p.hash, N, g->cmask, g->wprefix,
                                                                  g->bprefix, g->csorted, g->cdest, flags);
        }
    }
    {
        KTimer t(ctx, "resort_scatter");
        k_resort_scatter<<<nblk(N, 256), 256, 0, ctx->stream>>>(g->keys_alt, N, g->cmask, g->wprefix, g->bprefix,
                                                                g->csorted, g->cdest, p.hash, p.idx, flags);
    }
}

// the multi-block fallback of launch_resort: per pass an upsweep (pass 0's histogram came with the
// hash), the digit scan and the downsweep, each returning at once on the incremental path
static void launch_resort_radix(sphfem_ctx *ctx, sphfem_grid *g, sphfem_particles &p, int passes) {
    const int N = (int)p.N;
    const int nb = nblk(N, SORT_TILE);
    int *flags = ctx->d_flags;
    const int *mode = flags + 3;
    int *kbuf[2] = {p.hash, g->keys_alt};
    int *vbuf[2] = {p.idx, g->vals_alt};
    int src = 1;
    if (passes % 2 == 0) {
        KTimer t(ctx, "radix_copy");
        k_copy_keys_if_full<<<std::min(nblk(N, 256), 2048), 256, 0, ctx->stream>>>(g->keys_alt, p.hash, N, mode);
        src = 0;
    }
    int *kin = kbuf[src], *vin = nullptr;
    for (int ps = 0; ps < passes; ps++) {
        const int shift = 8 * ps;
        const int dst = 1 - src;
        if (ps > 0) {
            KTimer t(ctx, "radix_upsweep");
            k_upsweep<<<nb, SORT_BS, 0, ctx->stream>>>(kin, N, shift, g->hist, nb, mode);
        }
        {
            KTimer t(ctx, "radix_scan");
            k_scan_digits<<<RADIX, SORT_BS, 0, ctx->stream>>>(g->hist, nb, g->digit_tot, mode);
        }
        {
            KTimer t(ctx, "radix_downsweep");
            k_downsweep<<<nb, SORT_BS, 0, ctx->stream>>>(kin, ps == 0 ? nullptr : vin, kbuf[dst], vbuf[dst], N,
                                                         shift, g->hist, nb, g->digit_tot, flags, mode);
        }
        kin = kbuf[dst];
        vin = vbuf[dst];
        src = dst;
    }
}

// grid_gpu_green.cu:13-115: boundaries + gather of the 20 arrays into the buffer
__global__ void __launch_bounds__(256) k_reorder_full(const sphfem_particles p, sphfem_particles b,
                                                      int *__restrict__ cell_start, int *__restrict__ cell_end,
                                                      int num_cell, int *__restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int N = (int)p.N;
    if (i >= N) return;
    int h = p.hash[i];
    const int s = p.idx[i];
    if ((unsigned)h >= (unsigned)num_cell || (unsigned)s >= (unsigned)N) {
        atomicOr(&flags[0], 4);
        return;
    }
    int hp = __shfl_up(h, 1);
    if (__lane_id() == 0 && i > 0) hp = p.hash[i - 1];
    if (i == 0 || h != hp) {
        cell_start[h] = i;
        if (i > 0) cell_end[hp] = i;
    }
    if (i == N - 1) cell_end[h] = i + 1;
    b.pos[i] = p.pos[s];
    b.vel[i] = p.vel[s];
    b.fc[i] = p.fc[s];
    b.ft[i] = p.ft[s];
    b.h[i] = p.h[s];
    b.n[i] = p.n[s];
    b.rho[i] = p.rho[s];
    b.S[i] = p.S[s];
    b.eps_pl[i] = p.eps_pl[s];
    b.eps_pl_dot[i] = p.eps_pl_dot[s];
    b.T[i] = p.T[s];
    b.fixed[i] = p.fixed[s];
    b.blanked[i] = p.blanked[s];
    b.tool_particle[i] = p.tool_particle[s];
    b.pos_t[i] = p.pos_t[s];
    b.vel_t[i] = p.vel_t[s];
    b.rho_t[i] = p.rho_t[s];
    b.S_t[i] = p.S_t[s];
    b.T_t[i] = p.T_t[s];
    b.on_seg[i] = p.on_seg[s];
}

void launch_reorder_full(sphfem_ctx *ctx, sphfem_grid *g, const sphfem_particles &p) {
    KTimer t(ctx, "reorder");
    k_reorder_full<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(p, g->buf, g->cell_start, g->cell_end,
                                                            g->num_cell, ctx->d_flags);
}

// ---------------------------------------------------------------------------------------------
// grid_gpu_rothlin.cu:129-249: the stable sort's permutation gathers idx (which therefore keeps
// particle labels), pos, vel, fc, h, rho, S, eps_pl, eps_pl_dot, fixed, blanked, tool_particle, T,
// pos_t, vel_t, rho_t, S_t, T_t and on_seg in place; ft and n are not gathered (as in the
// reference).  The reference's on_seg copy takes the gathered idx buffer (:244-245), which would
// leave particle labels where segment numbers are expected; on_seg is gathered here.  The gather
// writes the buffer and a second pass copies it back, so the particle arrays keep their addresses.
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_rothlin_gather(const sphfem_particles p, sphfem_particles b,
                                                        const int *__restrict__ perm, int *__restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = (int)p.N;
    if (i >= N) return;
    const int s = perm[i];
    if ((unsigned)s >= (unsigned)N) {
        atomicOr(&flags[0], 4);
        return;
    }
    b.idx[i] = p.idx[s];
    b.pos[i] = p.pos[s];
    b.vel[i] = p.vel[s];
    b.fc[i] = p.fc[s];
    b.h[i] = p.h[s];
    b.rho[i] = p.rho[s];
    b.S[i] = p.S[s];
    b.eps_pl[i] = p.eps_pl[s];
    b.eps_pl_dot[i] = p.eps_pl_dot[s];
    b.fixed[i] = p.fixed[s];
    b.blanked[i] = p.blanked[s];
    b.tool_particle[i] = p.tool_particle[s];
    b.T[i] = p.T[s];
    b.pos_t[i] = p.pos_t[s];
    b.vel_t[i] = p.vel_t[s];
    b.rho_t[i] = p.rho_t[s];
    b.S_t[i] = p.S_t[s];
    b.T_t[i] = p.T_t[s];
    b.on_seg[i] = p.on_seg[s];
}

__global__ void __launch_bounds__(256) k_rothlin_copy_back(sphfem_particles p, const sphfem_particles b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int)p.N) return;
    p.idx[i] = b.idx[i];
    p.pos[i] = b.pos[i];
    p.vel[i] = b.vel[i];
    p.fc[i] = b.fc[i];
    p.h[i] = b.h[i];
    p.rho[i] = b.rho[i];
    p.S[i] = b.S[i];
    p.eps_pl[i] = b.eps_pl[i];
    p.eps_pl_dot[i] = b.eps_pl_dot[i];
    p.fixed[i] = b.fixed[i];
    p.blanked[i] = b.blanked[i];
    p.tool_particle[i] = b.tool_particle[i];
    p.T[i] = b.T[i];
    p.pos_t[i] = b.pos_t[i];
    p.vel_t[i] = b.vel_t[i];
    p.rho_t[i] = b.rho_t[i];
    p.S_t[i] = b.S_t[i];
    p.T_t[i] = b.T_t[i];
    p.on_seg[i] = b.on_seg[i];
}

// grid_gpu_rothlin.cu:19-126: unique_by_key + |differences| + scatter + inclusive scan with the
// last cell fixed up yields offsets[c] = number of sorted keys below c, so every cell gets
// cell_start[c] = lower_bound(c) and cell_end[c] = lower_bound(c + 1) (start == end when empty).
// One thread per cell, two binary searches over the sorted keys (L2-resident for these sizes).
__device__ __forceinline__ int lower_bound_keys(const int *__restrict__ keys, int N, int c) {
    int lo = 0, n = N;
    while (n > 0) {
        const int half = n >> 1;
        if (keys[lo + half] < c) {
            lo += half + 1;
            n -= half + 1;
        } else {
            n = half;
        }
    }
    return lo;
}

__global__ void __launch_bounds__(256) k_cells_dense(const int *__restrict__ keys, int N, int num_cell,
                                                     int *__restrict__ cell_start, int *__restrict__ cell_end) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= num_cell) return;
    cell_start[c] = lower_bound_keys(keys, N, c);
    cell_end[c] = lower_bound_keys(keys, N, c + 1);
}

void launch_cells_dense(sphfem_ctx *ctx, const int *sorted_keys, int N, int num_cell, int *cell_start, int *cell_end) {
    KTimer t(ctx, "cells_dense");
    k_cells_dense<<<nblk(num_cell, 256), 256, 0, ctx->stream>>>(sorted_keys, N, num_cell, cell_start, cell_end);
}

__global__ void k_iota(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

void launch_iota(sphfem_ctx *ctx, int *ptr, int n) {
    k_iota<<<nblk(n, 256), 256, 0, ctx->stream>>>(ptr, n);
}

// grid.cu:88-127: smallest and largest x and y over all particle positions and the largest h
// (blanked particles included, as the reference's min/max_element over the whole arrays).
// Positions are finite; min and max of finite values do not depend on the reduction order.
__device__ __forceinline__ void bbox_combine(double *a, const double *b) {
    a[0] = fmin(a[0], b[0]);
    a[1] = fmin(a[1], b[1]);
    a[2] = fmax(a[2], b[2]);
    a[3] = fmax(a[3], b[3]);
    a[4] = fmax(a[4], b[4]);
}

__device__ void bbox_block_reduce(double *v, double *out) {
    __shared__ double red[5][256 / 64];
    for (int off = 32; off > 0; off >>= 1) {
        double o[5];
        for (int k = 0; k < 5; k++) o[k] = __shfl_xor(v[k], off);
        bbox_combine(v, o);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
        for (int k = 0; k < 5; k++) red[k][wave] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc[5];
        for (int k = 0; k < 5; k++) acc[k] = red[k][0];
        for (int w = 1; w < 256 / 64; w++) {
            double o[5];
            for (int k = 0; k < 5; k++) o[k] = red[k][w];
            bbox_combine(acc, o);
        }
        for (int k = 0; k < 5; k++) out[k] = acc[k];
    }
}

__global__ void __launch_bounds__(256) k_bbox_partial(const sphfem_double2 *__restrict__ pos,
                                                      const double *__restrict__ h, int N, double *__restrict__ part) {
    double v[5] = {INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const sphfem_double2 q = pos[i];
        const double o[5] = {q.x, q.y, q.x, q.y, h[i]};
        bbox_combine(v, o);
    }
    bbox_block_reduce(v, part + 5 * blockIdx.x);
}

__global__ void __launch_bounds__(256) k_bbox_final(const double *__restrict__ part, int nb, double *__restrict__ out) {
    double v[5] = {INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nb; b += blockDim.x) bbox_combine(v, part + 5 * b);
    bbox_block_reduce(v, out);
}

void launch_bbox(sphfem_ctx *ctx, const sphfem_particles &p, double *part, double *out) {
    const int nb = std::max(1, std::min(nblk(p.N, 256), SPHFEM_BBOX_BLOCKS));
    KTimer t(ctx, "bbox");
    k_bbox_partial<<<nb, 256, 0, ctx->stream>>>(p.pos, p.h, (int)p.N, part);
    k_bbox_final<<<1, 256, 0, ctx->stream>>>(part, nb, out);
}

// p.hash must hold the keys; leaves the sorted keys in p.hash, the gathered arrays in place and
// the dense cell table in the grid
void launch_rothlin_sort(sphfem_ctx *ctx, sphfem_grid *g, sphfem_particles &p) {
    launch_radix_sort(ctx, g, p, false, g->perm);
    {
        KTimer t(ctx, "rothlin_gather");
        k_rothlin_gather<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(p, g->buf, g->perm, ctx->d_flags);
    }
    {
        KTimer t(ctx, "rothlin_copy_back");
        k_rothlin_copy_back<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(p, g->buf);
    }
    launch_cells_dense(ctx, p.hash, (int)p.N, g->num_cell, g->cell_start, g->cell_end);
}

// a live particle with a non-finite pair input (interactions_gpu.cu:143-379 reads pos, vel, h, rho, p,
// S, R and T of both particles): the pair kernels then evaluate the reference's full 5x5-cell window,
// where such a neighbour turns the sums NaN even outside the kernel support.  A finite sum that
// overflows only sends the step down that (exact, slower) path.
__device__ __forceinline__ void flag_nonfinite(int *__restrict__ flags, double blanked, sphfem_double2 pos,
                                               sphfem_double2 vel, sphfem_double4 S, double rho, double T, double h,
                                               double pr, sphfem_double4 R) {
    const double s = ((pos.x + pos.y) + (vel.x + vel.y)) + ((S.x + S.y) + (S.z + rho)) + ((T + h) + (pr + R.x)) +
                     (R.y + R.z);
    if (blanked != 1. && !isfinite(s)) atomicOr(&flags[SPHFEM_FLAG_NONFINITE], 1);
}

// FP32C neighbour records of one sorted particle: positions relative to the cell corner keep ~1e-7
// of the cell size; the stress terms are the per-particle factors of the pair sum, (S - p)/rho^2 and
// S_xy/rho^2; the temperature is stored as T - T_init, so that the FP32 differences T_i - T_j of a
// nearly uniform field keep ~1e-7 of the deviation instead of ~1e-7 of T; four SoA blocks; the
// first has 4 padding records that absorb four-wide scan loads
__device__ __forceinline__ void write_records(const DevConsts &c, float4 *__restrict__ rec, int N, int i, int h,
                                             sphfem_double2 pos, sphfem_double2 vel, sphfem_double4 S, double rho,
                                             double T, double blanked, double tool, double hh, double pr,
                                             sphfem_double4 R) {
    const int ny = c.geom.ny;
    const int ix = h / ny, iy = h - ix * ny;
    const double ox = c.geom.bbmin_x + ix * c.geom.dx;
    const double oy = c.geom.bbmin_y + iy * c.geom.dx;
    const double rho21 = 1. / (rho * rho);
    const float vol = (float)(c.phys.mass / rho);
    float4 r0, r1, r2, r3;
    r0.x = blanked == 1. ? 1e30f : (float)(pos.x - ox);     // blanked: never inside a support
    r0.y = (float)(pos.y - oy);
    r0.z = (float)iy;
    r0.w = (float)hh;
    r1.x = (float)vel.x;
    r1.y = (float)vel.y;
    r1.z = tool == 0. ? (float)((S.x - pr) * rho21) : 0.f;
    r1.w = tool == 0. ? (float)(S.y * rho21) : 0.f;
    r2.x = tool == 0. ? (float)((S.z - pr) * rho21) : 0.f;
    r2.y = (float)R.x;
    r2.z = (float)R.y;
    r2.w = (float)R.z;
    r3.x = tool == 1. ? -vol : vol;
    r3.y = (float)rho;
    r3.z = (float)sqrtf_d(c.phys.K / rho);
    r3.w = (float)(T - c.trml_wp.T_init);    // only differences enter the sums: offset from T_init
    const size_t n4 = (size_t)N + 4;
    rec[i] = r0;
    rec[n4 + i] = r1;
    rec[n4 + (size_t)N + i] = r2;
    rec[n4 + 2 * (size_t)N + i] = r3;
}

// FP32C records of an already sorted state as it stands (no predictor): the inputs of
// interactions_monaghan, with p and R as material_eos and corrector_artificial_stress left them
// (interactions_gpu.cu:143-379 reads them from the particle arrays)
__global__ void __launch_bounds__(256) k_records_state(DevConsts c, const sphfem_particles p, int num_cell,
                                                       int *__restrict__ flags, float4 *__restrict__ rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = (int)p.N;
    if (i >= N) return;
    int h = p.hash[i];
    if ((unsigned)h >= (unsigned)num_cell) {
        atomicOr(&flags[0], 4);
        h = 0;
    }
    const double tool = p.tool_particle[i];
    const double pr = tool == 0. ? p.p[i] : 0.;
    const sphfem_double4 R = tool == 0. ? p.R[i] : sphfem_double4{0., 0., 0., 0.};
    const sphfem_double2 pos = p.pos[i], vel = p.vel[i];
    const sphfem_double4 S = p.S[i];
    const double rho = p.rho[i], T = p.T[i], blanked = p.blanked[i], hh = p.h[i];
    flag_nonfinite(flags, blanked, pos, vel, S, rho, T, hh, pr, R);
    write_records(c, rec, N, i, h, pos, vel, S, rho, T, blanked, tool, hh, pr, R);
}

void launch_records_state(sphfem_ctx *ctx, const sphfem_particles &p, int num_cell, float4 *rec) {
    KTimer t(ctx, "records_state");
    k_records_state<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(ctx->c, p, num_cell, ctx->d_flags, rec);
}

// cell boundaries of the sorted keys (grid_gpu_green.cu:13-51); false if the key or the source
// index is out of range (flagged)
__device__ __forceinline__ bool cell_bounds(const sphfem_particles &p, int i, int N, int num_cell,
                                            int *__restrict__ cell_start, int *__restrict__ cell_end,
                                            int *__restrict__ flags, int &h, int &s) {
    h = p.hash[i];
    s = p.idx[i];
    if ((unsigned)h >= (unsigned)num_cell || (unsigned)s >= (unsigned)N) {
        atomicOr(&flags[0], 4);
        return false;
    }
    int hp = __shfl_up(h, 1);
    if (__lane_id() == 0 && i > 0) hp = p.hash[i - 1];
    if (i == 0 || h != hp) {
        cell_start[h] = i;
        if (i > 0) cell_end[hp] = i;
    }
    if (i == N - 1) cell_end[h] = i + 1;
    return true;
}

// FP64 form, fused: boundaries, gather (state and old derivatives), leap_frog init
// (leap_frog.cu:25-37), predict (:45-69), EOS (actions_gpu.cu:222-232) and artificial stress
// (:234-278).  Reads the unsorted arrays p via idx, writes the sorted predicted state to b.
__global__ void __launch_bounds__(256) k_reorder_predict(DevConsts c, const sphfem_particles p, sphfem_particles b,
                                                         int *__restrict__ cell_start, int *__restrict__ cell_end,
                                                         sphfem_double2 *__restrict__ pos_init,
                                                         sphfem_double2 *__restrict__ vel_init,
                                                         sphfem_double4 *__restrict__ S_init,
                                                         double *__restrict__ rho_init, double *__restrict__ T_init,
                                                         int num_cell, int *__restrict__ flags,
                                                         double *__restrict__ pre_vol,
                                                         sphfem_double4 *__restrict__ pre_s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = (int)p.N;
    if (i >= N) return;
    int h, s;
    if (!cell_bounds(p, i, N, num_cell, cell_start, cell_end, flags, h, s)) return;
    sphfem_double2 pos = p.pos[s], vel = p.vel[s];
    sphfem_double4 S = p.S[s];
    double rho = p.rho[s], T = p.T[s];
    const double blanked = p.blanked[s], tool = p.tool_particle[s];
    if (blanked != 1.) {
        const double dt = c.dt;
        pos_init[i] = pos;
        vel_init[i] = vel;
        S_init[i] = S;
        rho_init[i] = rho;
        T_init[i] = T;
        const sphfem_double2 pt = p.pos_t[s], vt = p.vel_t[s];
        const sphfem_double4 St = p.S_t[s];
        const double rt = p.rho_t[s], Tt = p.T_t[s];
        pos.x = pos.x + 0.5 * dt * pt.x;
        pos.y = pos.y + 0.5 * dt * pt.y;
        vel.x = vel.x + 0.5 * dt * vt.x;
        vel.y = vel.y + 0.5 * dt * vt.y;
        S.x = S.x + 0.5 * dt * St.x;
        S.y = S.y + 0.5 * dt * St.y;
        S.z = S.z + 0.5 * dt * St.z;
        S.w = S.w + 0.5 * dt * St.w;
        rho = rho + 0.5 * dt * rt;
        T = T + 0.5 * dt * Tt;
    }
    b.pos[i] = pos;
    b.vel[i] = vel;
    b.S[i] = S;
    b.rho[i] = rho;
    b.T[i] = T;
    b.blanked[i] = blanked;
    b.tool_particle[i] = tool;
    b.fc[i] = p.fc[s];
    b.ft[i] = p.ft[s];
    b.h[i] = p.h[s];
    b.n[i] = p.n[s];
    b.eps_pl[i] = p.eps_pl[s];
    b.eps_pl_dot[i] = p.eps_pl_dot[s];
    b.fixed[i] = p.fixed[s];
    b.on_seg[i] = p.on_seg[s];
    // the neighbour factors of interactions_gpu.cu:247-252, 283-285 (m/rho_j in the heat term, :352)
    pre_vol[i] = c.phys.mass / rho;
    if (tool == 0.) {
        const double pr = eos_pressure(c, rho);
        b.p[i] = pr;
        b.R[i] = artificial_stress(c, rho, pr, S);
        const double rho21 = 1. / (rho * rho);
        pre_s[i] = {(S.x - pr) * rho21, S.y * rho21, (S.z - pr) * rho21, sqrtf_d(c.phys.K / rho)};
        flag_nonfinite(flags, blanked, pos, vel, S, rho, T, p.h[s], pr, b.R[i]);
    } else {
        // neither gathered nor computed for tool particles by the reference: they stay in place
        b.p[i] = p.p[i];
        b.R[i] = p.R[i];
        pre_s[i] = {0., 0., 0., 0.};   // the pair loop skips these terms for tool particles
        flag_nonfinite(flags, blanked, pos, vel, sphfem_double4{0., 0., 0., 0.}, rho, T, p.h[s], 0.,
                       sphfem_double4{0., 0., 0., 0.});
    }
}

// FP32C form: boundaries, predict, EOS and artificial stress of the sorted particle, written only
// as neighbour records.  The update kernel gathers the pre-step state through idx itself (and
// repeats the predictor for velocity, stress and density), so no sorted copy of the state is
// written here.
__global__ void __launch_bounds__(256) k_reorder_records(DevConsts c, const sphfem_particles p,
                                                         int *__restrict__ cell_start, int *__restrict__ cell_end,
                                                         int num_cell, int *__restrict__ flags,
                                                         float4 *__restrict__ rec, const RateView rates) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = (int)p.N;
    if (i >= N) return;
    int h, s;
    if (!cell_bounds(p, i, N, num_cell, cell_start, cell_end, flags, h, s)) return;
    sphfem_double2 pos = ld_stream(&p.pos[s]), vel = ld_stream(&p.vel[s]);
    sphfem_double4 S = ld_stream(&p.S[s]);
    double rho = ld_stream(&p.rho[s]), T = ld_stream(&p.T[s]);
    const double blanked = ld_stream(&p.blanked[s]), tool = ld_stream(&p.tool_particle[s]);
    if (blanked != 1.) {
        const double dt = c.dt;
        // the FP32 derivative carry of the FP32C form, in the order of p
        const float2 pt = ld_stream(&rates.pos_t[s]), vt = ld_stream(&rates.vel_t[s]);
        const float4 St = ld_stream(&rates.S_t[s]);
        const double rt = ld_stream(&rates.rho_t[s]), Tt = ld_stream(&rates.T_t[s]);
        pos.x = pos.x + 0.5 * dt * pt.x;
        pos.y = pos.y + 0.5 * dt * pt.y;
        vel.x = vel.x + 0.5 * dt * vt.x;
        vel.y = vel.y + 0.5 * dt * vt.y;
        S.x = S.x + 0.5 * dt * St.x;
        S.y = S.y + 0.5 * dt * St.y;
        S.z = S.z + 0.5 * dt * St.z;
        S.w = S.w + 0.5 * dt * St.w;
        rho = rho + 0.5 * dt * rt;
        T = T + 0.5 * dt * Tt;
    }
    sphfem_double4 R = {0., 0., 0., 0.};
    double pr = 0.;
    if (tool == 0.) {
        pr = eos_pressure(c, rho);
        R = artificial_stress(c, rho, pr, S);
    }
    const double hh = ld_stream(&p.h[s]);
    flag_nonfinite(flags, blanked, pos, vel, tool == 0. ? S : sphfem_double4{0., 0., 0., 0.}, rho, T, hh, pr, R);
    write_records(c, rec, N, i, h, pos, vel, S, rho, T, blanked, tool, hh, pr, R);
}

void launch_reorder_predict(sphfem_ctx *ctx, sphfem_grid *g, const sphfem_particles &p, sphfem_leap_frog *lf,
                            const RateView *rates) {
    if (lf->precision == SPHFEM_FP32C) {
        KTimer t(ctx, "reorder_records");
        k_reorder_records<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(ctx->c, p, g->cell_start, g->cell_end,
                                                                   g->num_cell, ctx->d_flags, lf->rec, *rates);
        return;
    }
    KTimer t(ctx, "reorder_predict");
    k_reorder_predict<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(ctx->c, p, g->buf, g->cell_start, g->cell_end,
                                                               lf->pos_init, lf->vel_init, lf->S_init,
                                                               lf->rho_init, lf->T_init, g->num_cell,
                                                               ctx->d_flags, lf->pre_vol, lf->pre_s);
}

}  // namespace sphfem
