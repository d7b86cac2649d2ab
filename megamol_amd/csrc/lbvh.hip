// lbvh.hip — linear BVH built on the GPU, replacing optixAccelBuild
// (plugins/optix_hpg/src/optix/SphereGeometry.cpp:197-214) and the sphere bounds program
// (plugins/optix_hpg/cuda_resources/sphere.cuh:133-147, fused into the Morton pass here).
//
// Steps, each one kernel (or a few) on the context stream:
//   1. scene reduction: centre bounds, sphere bounds, |coord|max, Morton scale
//   2. Morton codes, 30-bit (10 per axis) or 63-bit (21 per axis), over one uniform scale
//   3. LSD radix sort of (key, particle index), 8-bit digits, stable: a tile's keys are ranked
//      with wavefront ballots (8 ballots give each lane the mask of same-digit lanes)
//   4. hierarchy and boxes together, bottom-up, with the gather of the spheres into Morton order
//      (leaf k) on the way: the second arrival at a parent merges and writes the whole node
//      (Karras 2012 numbering and index tie-break for equal keys, Apetrei 2014 parent rule; LDS
//      meeting places inside 1024-leaf chunks; a node over two leaves is not stored, its parent
//      references the leaf pair)
//   5. the nodes that straddle chunks: one workgroup per span of chunks, then one for the top
#include <algorithm>

#include "ctx.hpp"

namespace mmhrt {

namespace {

constexpr int RB = 256;      // threads per block
// keys per thread per tile: RI = 8, 16 or 32 (a launch option); a tile is RB * RI keys

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

// Mask of the lanes of this wave whose (valid) DB-bit digit equals ours.
template <int DB = 8>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < DB; ++b) {
        uint64_t bm = __ballot((d >> b) & 1u);
        m &= ((d >> b) & 1u) ? bm : ~bm;
    }
    return m;
}

// Barrier for work that shares only LDS: waits for this wave's LDS accesses, not for its global
// loads and stores, so loads and stores issued before it stay in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// ---------------------------------------------------------------- scene reduction
constexpr int NRED = 13;

__global__ __launch_bounds__(256) void k_reduce_partial(const float4* __restrict__ sph, uint64_t n,
                                                        float* __restrict__ partial) {
    float v[NRED];
    for (int i = 0; i < 3; ++i) {
        v[i] = INFINITY;       // centre min
        v[3 + i] = -INFINITY;  // centre max
        v[6 + i] = INFINITY;   // sphere min
        v[9 + i] = -INFINITY;  // sphere max
    }
    v[12] = 0.0f;
    // 8 spheres per thread per step, loads issued first at clamped indices (min and max do not
    // depend on the order, and a repeated sphere changes nothing)
    constexpr int SB = 8;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x * SB;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x * SB; b < n; b += step) {
        float4 sb[SB];
#pragma unroll
        for (int j = 0; j < SB; ++j) {
            const uint64_t i = b + (uint64_t)j * blockDim.x + threadIdx.x;
            sb[j] = sph[i < n ? i : n - 1];
        }
#pragma unroll
        for (int j = 0; j < SB; ++j) {
            const float4 s = sb[j];
            float c[3] = {s.x, s.y, s.z};
            float r = s.w;
            for (int a = 0; a < 3; ++a) {
                v[a] = fminf(v[a], c[a]);
                v[3 + a] = fmaxf(v[3 + a], c[a]);
                v[6 + a] = fminf(v[6 + a], c[a] - r);
                v[9 + a] = fmaxf(v[9 + a], c[a] + r);
                v[12] = fmaxf(v[12], fabsf(c[a]) + fabsf(r));
            }
        }
    }
    __shared__ float red[NRED][256];
    for (int k = 0; k < NRED; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int k = 0; k < NRED; ++k) {
                float a = red[k][threadIdx.x], b = red[k][threadIdx.x + s];
                bool is_min = (k < 3) || (k >= 6 && k < 9);
                red[k][threadIdx.x] = is_min ? fminf(a, b) : fmaxf(a, b);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < NRED) partial[(size_t)blockIdx.x * NRED + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_reduce_final(const float* __restrict__ partial, int nblocks,
                                                      SceneInfo* info, int bits) {
    __shared__ float red[NRED][256];
    float v[NRED];
    for (int k = 0; k < NRED; ++k) {
        bool is_min = (k < 3) || (k >= 6 && k < 9);
        v[k] = k == 12 ? 0.0f : (is_min ? INFINITY : -INFINITY);
    }
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int k = 0; k < NRED; ++k) {
            float x = partial[(size_t)b * NRED + k];
            bool is_min = (k < 3) || (k >= 6 && k < 9);
            v[k] = is_min ? fminf(v[k], x) : fmaxf(v[k], x);
        }
    for (int k = 0; k < NRED; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int k = 0; k < NRED; ++k) {
                float a = red[k][threadIdx.x], b = red[k][threadIdx.x + st];
                bool is_min = (k < 3) || (k >= 6 && k < 9);
                red[k][threadIdx.x] = is_min ? fminf(a, b) : fmaxf(a, b);
            }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    for (int k = 0; k < NRED; ++k) v[k] = red[k][0];
    for (int a = 0; a < 3; ++a) {
        info->clo[a] = v[a];
        info->chi[a] = v[3 + a];
        info->blo[a] = v[6 + a];
        info->bhi[a] = v[9 + a];
    }
    info->absmax = v[12];
    float ext = fmaxf(fmaxf(v[3] - v[0], v[4] - v[1]), v[5] - v[2]);
    float cells = bits == 63 ? 2097152.0f : 1024.0f;
    info->scale = ext > 0.0f ? cells / ext : 0.0f;
}

// ---------------------------------------------------------------- Morton codes
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
__device__ __forceinline__ uint64_t spread21(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x001f00000000ffffull;
    v = (v | (v << 16)) & 0x001f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

// ---------------------------------------------------------------- radix sort
// A tile is RT consecutive keys; wave w of the block owns the contiguous segment
// [w * RT/4, (w+1) * RT/4) and walks it in RI rounds of 64, so the wave order, the round order
// and the lane order together are the input order (what makes the ranking stable). Every
// thread issues all of its RI key (and value) loads before it ranks anything.

// Adds digit d of every active lane to the LDS histogram h; lanes with the first active lane's
// digit add as one (Morton digits of neighbouring keys are often equal, and the LDS unit
// serialises same-address lanes).
__device__ __forceinline__ void hist_add(uint32_t* h, uint32_t d) {
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    const bool eq = d == d0;
    const uint64_t m = __ballot(eq);
    const uint64_t act = __ballot(true);
    if (__lane_id() == (uint32_t)(__ffsll((unsigned long long)act) - 1)) atomicAdd(&h[d0], (uint32_t)__popcll(m));
    if (!eq) atomicAdd(&h[d], 1u);
}

__device__ __forceinline__ uint32_t morton30(const float4 s, const SceneInfo* info, float sc) {
    const float f[3] = {(s.x - info->clo[0]) * sc, (s.y - info->clo[1]) * sc, (s.z - info->clo[2]) * sc};
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const uint32_t u = (uint32_t)f[a];
        q[a] = u > 1023u ? 1023u : u;
    }
    return (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
}
__device__ __forceinline__ uint64_t morton63(const float4 s, const SceneInfo* info, float sc) {
    const float f[3] = {(s.x - info->clo[0]) * sc, (s.y - info->clo[1]) * sc, (s.z - info->clo[2]) * sc};
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const uint32_t u = (uint32_t)f[a];
        q[a] = u > 2097151u ? 2097151u : u;
    }
    return (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
}

// The Morton codes of one radix tile (RT spheres, laid out as the sort's tiles are) and the
// tile's histogram of the first 8-bit digit: the first radix pass then needs no histogram pass,
// and no value array either, since its values are the particle indices themselves.
template <typename K, int RI, int DB>
__global__ __launch_bounds__(RB) void k_morton_tile(const float4* __restrict__ sph, uint64_t n,
                                                    const SceneInfo* __restrict__ info, K* __restrict__ keys,
                                                    uint32_t* __restrict__ hist) {
    constexpr int RT = RB * RI, RW = RT / (RB / 64), SB = RI < 16 ? RI : 16, ND = 1 << DB;
    __shared__ uint32_t h[ND];
    for (int d = threadIdx.x; d < ND; d += RB) h[d] = 0;
    __syncthreads();
    const float sc = info->scale;
    const uint64_t base = (uint64_t)blockIdx.x * RT + (uint64_t)(threadIdx.x >> 6) * RW + __lane_id();
    const bool whole = (uint64_t)(blockIdx.x + 1) * RT <= n;  // uniform
    for (int r0 = 0; r0 < RI; r0 += SB) {
        float4 sb[SB];  // SB sphere loads in flight: a whole tile's at constant offsets, else at clamped indices
        if (whole) {
            const float4* sp = sph + base;
#pragma unroll
            for (int j = 0; j < SB; ++j) sb[j] = sp[(r0 + j) * 64];
        } else {
#pragma unroll
            for (int j = 0; j < SB; ++j) {
                const uint64_t i = base + (uint64_t)(r0 + j) * 64;
                sb[j] = sph[i < n ? i : n - 1];
            }
        }
#pragma unroll
        for (int j = 0; j < SB; ++j) {
            const uint64_t i = base + (uint64_t)(r0 + j) * 64;
            K key;
            if constexpr (sizeof(K) == 8)
                key = morton63(sb[j], info, sc);
            else
                key = morton30(sb[j], info, sc);
            if (whole || i < n) {
                keys[i] = key;
                hist_add(h, (uint32_t)(key & (K)(ND - 1)));
            }
        }
    }
    lds_barrier();  // the histogram is in LDS; the key stores drain meanwhile
    for (int d = threadIdx.x; d < ND; d += RB) hist[(size_t)blockIdx.x * ND + d] = h[d];
}

template <typename K, int RI, int DB>
__global__ __launch_bounds__(RB) void k_radix_hist(const K* __restrict__ keys, uint64_t n, int shift,
                                                   uint32_t* __restrict__ hist) {
    constexpr int RT = RB * RI, RW = RT / (RB / 64), ND = 1 << DB;
    __shared__ uint32_t h[ND];
    for (int d = threadIdx.x; d < ND; d += RB) h[d] = 0;
    const uint64_t base = (uint64_t)blockIdx.x * RT + (uint64_t)(threadIdx.x >> 6) * RW + __lane_id();
    // unconditional loads at clamped indices (a guarded load becomes a branch and a wait per
    // key); keys past n count as digit ND, which is not counted
    uint32_t d[RI];
#pragma unroll
    for (int r = 0; r < RI; ++r) {
        uint64_t i = base + (uint64_t)r * 64;
        d[r] = (uint32_t)((keys[i < n ? i : n - 1] >> shift) & (K)(ND - 1));
    }
#pragma unroll
    for (int r = 0; r < RI; ++r)
        if (base + (uint64_t)r * 64 >= n) d[r] = (uint32_t)ND;
    __syncthreads();
    // LDS atomics; the lanes whose digit equals lane 0's add as one (Morton digits of
    // neighbouring keys are often equal, and the LDS unit serialises same-address lanes)
#pragma unroll
    for (int r = 0; r < RI; ++r) {
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(d[r]);
        const bool eq = d[r] == d0;
        const uint64_t m = __ballot(eq);
        if (__lane_id() == 0 && d0 < (uint32_t)ND) atomicAdd(&h[d0], (uint32_t)__popcll(m));
        if (!eq && d[r] < (uint32_t)ND) atomicAdd(&h[d[r]], 1u);
    }
    __syncthreads();
    for (int dd = threadIdx.x; dd < ND; dd += RB) hist[(size_t)blockIdx.x * ND + dd] = h[dd];
}

// Stable scatter of one tile for one digit (vin null: the values are the key indices). Each wave
// ranks its segment with ballots and a running per-(wave, digit) count in LDS (no barrier between
// rounds), one pass over the digits turns the per-wave counts into offsets, the keys are placed in
// digit-sorted order in LDS and finally written out in digit-contiguous runs, so that consecutive
// threads write consecutive addresses. The tile's digit counts and global offsets are loaded with
// its keys, and the barriers order LDS only (no global memory passes between the threads).
template <typename K, int RI, int DB>
__global__ __launch_bounds__(RB) void k_radix_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                      K* __restrict__ kout, uint32_t* __restrict__ vout, uint64_t n,
                                                      int shift, const uint32_t* __restrict__ hist,
                                                      const uint32_t* __restrict__ offs, int xcd) {
    constexpr int RT = RB * RI, RW = RT / (RB / 64), ND = 1 << DB, DPT = ND / RB;  // digits per thread
    __shared__ K s_key[RT];
    __shared__ uint32_t s_val[RT];
    __shared__ uint16_t wc[RB / 64][ND];  // per-wave running digit counts, then per-wave offsets (< RT)
    __shared__ uint32_t goff[ND];         // global offset of the digit's run minus its local start
    __shared__ uint32_t wsum[RB / 64];
    // xcd: workgroups are dealt to the 8 XCDs in turn, so workgroup w takes tile (w mod 8) * per +
    // w / 8 and each XCD scatters one contiguous range of tiles: the runs that consecutive tiles
    // write for a digit are adjacent, and their shared cache lines meet in the same L2
    uint32_t tile = blockIdx.x;
    if (xcd) {
        const uint32_t per = gridDim.x / 8;
        tile = (blockIdx.x % 8) * per + blockIdx.x / 8;
        if ((uint64_t)tile * RT >= n) return;
    }
    const int wave = threadIdx.x >> 6;
    const uint32_t lane = __lane_id();
    const uint64_t tbase = (uint64_t)tile * RT;
    const uint64_t base = tbase + (uint64_t)wave * RW + lane;
    // loads: a whole tile from one base address at constant offsets, the last one at clamped
    // indices (validity is applied where the keys are used)
    K key[RI];
    uint32_t val[RI];
    if (tbase + RT <= n) {
        const K* kp = kin + base;
#pragma unroll
        for (int r = 0; r < RI; ++r) key[r] = kp[r * 64];
        if (vin) {
            const uint32_t* vp = vin + base;
#pragma unroll
            for (int r = 0; r < RI; ++r) val[r] = vp[r * 64];
        }
    } else {
#pragma unroll
        for (int r = 0; r < RI; ++r) {
            const uint64_t i = base + (uint64_t)r * 64;
            key[r] = kin[i < n ? i : n - 1];
        }
        if (vin) {
#pragma unroll
            for (int r = 0; r < RI; ++r) {
                const uint64_t i = base + (uint64_t)r * 64;
                val[r] = vin[i < n ? i : n - 1];
            }
        }
    }
    if (!vin) {  // the first pass: the values are the indices
#pragma unroll
        for (int r = 0; r < RI; ++r) val[r] = (uint32_t)(base + (uint64_t)r * 64);
    }
    // tile-local digit starts: exclusive scan of this tile's digit counts; thread t owns the DPT
    // consecutive digits from t * DPT
    uint32_t cnt[DPT], ofv[DPT];
    uint32_t tsum = 0;
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
        cnt[j] = hist[(size_t)tile * ND + threadIdx.x * DPT + j];
        ofv[j] = offs[(size_t)tile * ND + threadIdx.x * DPT + j];
        tsum += cnt[j];
    }
    uint32_t inc = tsum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= (uint32_t)o) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    for (int d = threadIdx.x; d < ND; d += RB)
#pragma unroll
        for (int w = 0; w < RB / 64; ++w) wc[w][d] = 0;
    lds_barrier();
    uint32_t wb = 0;
    for (int w = 0; w < wave; ++w) wb += wsum[w];
    uint32_t lstart[DPT];
    {
        uint32_t run = wb + inc - tsum;
#pragma unroll
        for (int j = 0; j < DPT; ++j) {
            const int d = threadIdx.x * DPT + j;
            lstart[j] = run;
            goff[d] = ofv[j] - run;
            run += cnt[j];
        }
    }
    // ranks inside the wave's segment
    uint32_t rank[RI];
#pragma unroll
    for (int r = 0; r < RI; ++r) {
        bool valid = base + (uint64_t)r * 64 < n;
        uint32_t d = (uint32_t)((key[r] >> shift) & (K)(ND - 1));
        uint64_t peers = match_digit<DB>(d, valid);
        uint32_t before = valid ? wc[wave][d] : 0u;
        rank[r] = before + (uint32_t)__popcll(peers & lanemask_lt());
        if (valid && (__ffsll((unsigned long long)peers) - 1) == (int)lane)
            wc[wave][d] = (uint16_t)(before + (uint32_t)__popcll(peers));
    }
    lds_barrier();
#pragma unroll
    for (int j = 0; j < DPT; ++j) {  // per-wave counts -> tile-local offsets of each wave's run
        const int d = threadIdx.x * DPT + j;
        uint32_t acc = lstart[j];
#pragma unroll
        for (int w = 0; w < RB / 64; ++w) {
            uint32_t c = wc[w][d];
            wc[w][d] = (uint16_t)acc;
            acc += c;
        }
    }
    lds_barrier();
#pragma unroll
    for (int r = 0; r < RI; ++r) {
        if (base + (uint64_t)r * 64 < n) {
            uint32_t d = (uint32_t)((key[r] >> shift) & (K)(ND - 1));
            uint32_t pos = wc[wave][d] + rank[r];
            s_key[pos] = key[r];
            s_val[pos] = val[r];
        }
    }
    lds_barrier();
    const uint32_t nvalid = (uint32_t)((n - tbase) < (uint64_t)RT ? (n - tbase) : (uint64_t)RT);
    for (uint32_t i = threadIdx.x; i < nvalid; i += RB) {
        K k = s_key[i];
        uint32_t d = (uint32_t)((k >> shift) & (K)(ND - 1));
        uint32_t g = goff[d] + i;
        kout[g] = k;
        vout[g] = s_val[i];
    }
}

// Global digit offsets of every tile: offs[t][d] = (keys with a smaller digit) + (keys with
// digit d in earlier tiles). hist and offs are tile-major ([tile][256]), so every pass below
// reads or writes whole 1 KB rows. Column sums per group of CG tiles, one block that scans the
// groups and the digits, then each group's running prefix down its tiles.
constexpr int CG = 64;  // tiles per column group

__global__ __launch_bounds__(256) void k_col_sums(const uint32_t* __restrict__ hist, uint32_t tiles,
                                                  uint32_t* __restrict__ gsum, uint32_t nd) {
    const uint32_t t0 = blockIdx.x * CG;
    const uint32_t d = blockIdx.y * 256 + threadIdx.x;  // digit
    uint32_t v[CG];
#pragma unroll
    for (int k = 0; k < CG; ++k) v[k] = t0 + k < tiles ? hist[(size_t)(t0 + k) * nd + d] : 0u;
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < CG; ++k) acc += v[k];
    gsum[(size_t)blockIdx.x * nd + d] = acc;
}

// One block: exclusive prefix of each digit over the groups (32 loads in flight per thread: the
// block walks the groups one latency at a time), and the exclusive scan of the digit totals into
// row `groups`.
__global__ __launch_bounds__(256) void k_col_top(uint32_t* __restrict__ gsum, uint32_t groups, uint32_t nd) {
    // block y takes digits 256 y .. 256 y + 255: each digit's exclusive prefix over the groups (32
    // loads in flight per thread: the block walks the groups one latency at a time), its total into
    // row `groups`
    const uint32_t d = blockIdx.x * 256 + threadIdx.x;
    uint32_t run = 0;
    for (uint32_t g0 = 0; g0 < groups; g0 += 32) {
        uint32_t v[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) v[k] = gsum[(size_t)(g0 + k < groups ? g0 + k : groups - 1) * nd + d];
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            if (g0 + k < groups) {
                gsum[(size_t)(g0 + k) * nd + d] = run;
                run += v[k];
            }
        }
    }
    gsum[(size_t)groups * nd + d] = run;
}

// One block: the exclusive scan of the digit totals in row `groups` (keys with a smaller digit).
__global__ __launch_bounds__(256) void k_col_digits(uint32_t* __restrict__ gsum, uint32_t groups, uint32_t nd) {
    __shared__ uint32_t wsum[4];
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    const uint32_t dpt = nd / 256;  // digits per thread: d = threadIdx.x * dpt + j
    uint32_t* row = gsum + (size_t)groups * nd;
    uint32_t tot[8];
    uint32_t tsum = 0;
    for (uint32_t j = 0; j < dpt; ++j) {
        tot[j] = row[threadIdx.x * dpt + j];
        tsum += tot[j];
    }
    uint32_t inc = tsum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= (uint32_t)o) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < wave; ++w) base += wsum[w];
    uint32_t run = base + inc - tsum;
    for (uint32_t j = 0; j < dpt; ++j) {
        row[threadIdx.x * dpt + j] = run;
        run += tot[j];
    }
}

// 8-bit digits: each digit's exclusive prefix over the groups and its total (row `groups`), in
// place of k_col_top and k_col_digits. Block b takes digits 64 b .. 64 b + 63 with SCAN_T
// threads per digit, each over a contiguous run of groups: one load latency per walk, where the
// one-block scan walks the groups 32 at a time. The digit totals are scanned by k_col_offsets.
constexpr int SCAN_T = 16;
__global__ __launch_bounds__(1024) void k_col_scan(uint32_t* __restrict__ gsum, uint32_t groups) {
    __shared__ uint32_t part[SCAN_T][64];
    const uint32_t dl = threadIdx.x & 63u, q = threadIdx.x >> 6;
    const uint32_t d = blockIdx.x * 64u + dl;
    const uint32_t gpt = (groups + SCAN_T - 1) / SCAN_T;
    const uint32_t g0 = q * gpt;
    const uint32_t g1 = g0 + gpt < groups ? g0 + gpt : groups;
    // loads at clamped rows, masked where summed (a guarded load compiles to a branch and a wait)
    auto row = [&](uint32_t g) -> uint32_t { return g < g1 ? g : g1 - 1; };
    uint32_t sum = 0;
    for (uint32_t g = g0; g < g1; g += 16) {
        uint32_t v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = gsum[(size_t)row(g + k) * 256 + d];
#pragma unroll
        for (int k = 0; k < 16; ++k) sum += g + k < g1 ? v[k] : 0u;
    }
    part[q][dl] = sum;
    __syncthreads();
    uint32_t run = 0;
    for (uint32_t j = 0; j < q; ++j) run += part[j][dl];
    if (q == SCAN_T - 1) gsum[(size_t)groups * 256 + d] = run + sum;
    for (uint32_t g = g0; g < g1; g += 16) {
        uint32_t v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = gsum[(size_t)row(g + k) * 256 + d];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (g + k < g1) {
                gsum[(size_t)(g + k) * 256 + d] = run;
                run += v[k];
            }
        }
    }
}

// Each tile's offsets: its group's prefix plus the running count down the group's tiles, plus the
// digit's base. SCAN: the digit totals in row `groups` are not scanned yet (k_col_scan); every
// block scans them itself (nd = 256, one digit per thread), 1 KB each, instead of a kernel of its
// own doing it once.
template <bool SCAN>
__global__ __launch_bounds__(256) void k_col_offsets(const uint32_t* __restrict__ hist, uint32_t tiles,
                                                     uint32_t groups, const uint32_t* __restrict__ gsum,
                                                     uint32_t* __restrict__ offs, uint32_t nd) {
    const uint32_t t0 = blockIdx.x * CG;
    const uint32_t d = blockIdx.y * 256 + threadIdx.x;  // digit
    uint32_t v[CG];
#pragma unroll
    for (int k = 0; k < CG; ++k) v[k] = t0 + k < tiles ? hist[(size_t)(t0 + k) * nd + d] : 0u;
    uint32_t base;
    if constexpr (SCAN) {
        __shared__ uint32_t wsum[4];
        const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
        const uint32_t tot = gsum[(size_t)groups * nd + d];
        uint32_t inc = tot;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(inc, o, 64);
            if (lane >= (uint32_t)o) inc += y;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        base = inc - tot;
        for (uint32_t w = 0; w < wave; ++w) base += wsum[w];
    } else {
        base = gsum[(size_t)groups * nd + d];
    }
    uint32_t run = gsum[(size_t)blockIdx.x * nd + d] + base;
#pragma unroll
    for (int k = 0; k < CG; ++k) {
        if (t0 + k < tiles) offs[(size_t)(t0 + k) * nd + d] = run;
        run += v[k];
    }
}

// ---------------------------------------------------------------- single-pass radix passes
// The alternative to histogram + column scans + scatter per pass (option sort_onesweep): the
// digit counts of every pass over all keys come from one read (they do not depend on the key
// order), and each pass is then one kernel. Tiles take a ticket in launch order, rank their keys
// as the scatter above does, publish their per-digit counts, and find their global offsets by
// looking back over the earlier tiles' published counts until one that already holds its
// inclusive prefix (decoupled look-back, Merrill & Garland 2016; one pass per digit as in
// Adinets & Merrill 2022, "Onesweep"). A status word is flag (2 bits) | count (30 bits), written
// and read with device-scope atomics, so no fence is needed. Every tile that a tile waits for
// took its ticket earlier and is running; the wait is still bounded, and on exhaustion the
// kernel records a fault and goes on (a wrong order, never a hang).
constexpr uint32_t OS_AGG = 1u << 30, OS_INC = 2u << 30, OS_MASK = (1u << 30) - 1u;
constexpr int OS_SPIN = 1 << 20;

// ghist[p][d]: keys whose digit p is d, for p < passes (ghist zero on entry).
template <typename K>
__global__ __launch_bounds__(256) void k_radix_ghist(const K* __restrict__ keys, uint64_t n, int passes,
                                                     uint32_t* __restrict__ ghist) {
    __shared__ uint32_t h[8][256];
    for (int p = 0; p < 8; ++p) h[p][threadIdx.x] = 0;
    __syncthreads();
    // 16 keys per thread per step, all loads issued first (coalesced: key j of a step is
    // 256 * j keys after key 0 of the workgroup's step)
    constexpr int KB = 16;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x * KB;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x * KB; b < n; b += step) {
        K k[KB];
#pragma unroll
        for (int j = 0; j < KB; ++j) {
            const uint64_t i = b + (uint64_t)j * blockDim.x + threadIdx.x;
            k[j] = keys[i < n ? i : n - 1];
        }
#pragma unroll
        for (int j = 0; j < KB; ++j) {
            const bool valid = b + (uint64_t)j * blockDim.x + threadIdx.x < n;
            if (__ballot(valid) == 0) break;
            if (valid)
                for (int p = 0; p < passes; ++p) hist_add(h[p], (uint32_t)((k[j] >> (8 * p)) & 255u));
        }
    }
    __syncthreads();
    for (int p = 0; p < passes; ++p) {
        const uint32_t v = h[p][threadIdx.x];
        if (v) atomicAdd(&ghist[p * 256 + threadIdx.x], v);
    }
}

// dstart[p][d]: exclusive scan of ghist[p] over the digits (one block, thread = digit).
__global__ __launch_bounds__(256) void k_radix_gscan(const uint32_t* __restrict__ ghist, int passes,
                                                     uint32_t* __restrict__ dstart) {
    __shared__ uint32_t wsum[4];
    const uint32_t d = threadIdx.x, lane = __lane_id(), wave = threadIdx.x >> 6;
    for (int p = 0; p < passes; ++p) {
        const uint32_t c = ghist[p * 256 + d];
        uint32_t inc = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(inc, o, 64);
            if (lane >= (uint32_t)o) inc += y;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        uint32_t base = 0;
        for (uint32_t w = 0; w < wave; ++w) base += wsum[w];
        dstart[p * 256 + d] = base + inc - c;
        __syncthreads();
    }
}

template <typename K, int RI>
__global__ __launch_bounds__(RB) void k_radix_onesweep(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                       K* __restrict__ kout, uint32_t* __restrict__ vout, uint64_t n,
                                                       int shift, const uint32_t* __restrict__ dstart,
                                                       uint32_t* status, uint32_t* ticket, uint32_t* fault) {
    constexpr int RT = RB * RI, RW = RT / (RB / 64);
    __shared__ K s_key[RT];
    __shared__ uint32_t s_val[RT];
    __shared__ uint32_t wc[RB / 64][256];  // per-wave running digit counts, then per-wave offsets
    __shared__ uint32_t goff[256];         // global offset of the digit's run minus its local start
    __shared__ uint32_t wsum[RB / 64];
    __shared__ uint32_t s_tile, s_bad;
    const int wave = threadIdx.x >> 6;
    const uint32_t lane = __lane_id();
    if (threadIdx.x == 0) {
        s_tile = atomicAdd(ticket, 1u);
        s_bad = 0;
    }
#pragma unroll
    for (int w = 0; w < RB / 64; ++w) wc[w][threadIdx.x] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t tbase = (uint64_t)tile * RT;
    const uint64_t base = tbase + (uint64_t)wave * RW + lane;
    K key[RI];
    uint32_t val[RI];
#pragma unroll
    for (int r = 0; r < RI; ++r) {
        uint64_t i = base + (uint64_t)r * 64;
        bool valid = i < n;
        key[r] = valid ? kin[i] : (K)0;
    }
    if (vin) {
#pragma unroll
        for (int r = 0; r < RI; ++r) {
            uint64_t i = base + (uint64_t)r * 64;
            val[r] = i < n ? vin[i] : 0u;
        }
    } else {  // the first pass: the values are the indices
#pragma unroll
        for (int r = 0; r < RI; ++r) val[r] = (uint32_t)(base + (uint64_t)r * 64);
    }
    uint32_t rank[RI];
#pragma unroll
    for (int r = 0; r < RI; ++r) {
        bool valid = base + (uint64_t)r * 64 < n;
        uint32_t d = (uint32_t)((key[r] >> shift) & 255u);
        uint64_t peers = match_digit(d, valid);
        uint32_t before = valid ? wc[wave][d] : 0u;
        rank[r] = before + (uint32_t)__popcll(peers & lanemask_lt());
        if (valid && (__ffsll((unsigned long long)peers) - 1) == (int)lane) wc[wave][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // thread = digit: the tile's count, published at once; tile-local start of the digit; each
    // wave's offset within the digit's run
    const uint32_t dg = threadIdx.x;
    uint32_t cw[RB / 64];
    uint32_t cnt = 0;
#pragma unroll
    for (int w = 0; w < RB / 64; ++w) {
        cw[w] = wc[w][dg];
        cnt += cw[w];
    }
    uint32_t* st = status + (size_t)tile * 256 + dg;
    __hip_atomic_store(st, (tile == 0 ? OS_INC : OS_AGG) | cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t inc = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= (uint32_t)o) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t wb = 0;
    for (int w = 0; w < wave; ++w) wb += wsum[w];
    const uint32_t lstart = wb + inc - cnt;
    {
        uint32_t acc = lstart;
#pragma unroll
        for (int w = 0; w < RB / 64; ++w) {
            wc[w][dg] = acc;
            acc += cw[w];
        }
    }
    // look back: keys with digit dg in all earlier tiles
    uint32_t excl = 0;
    if (tile > 0) {
        for (int64_t t = (int64_t)tile - 1; t >= 0; --t) {
            const uint32_t* pt = status + (size_t)t * 256 + dg;
            uint32_t v = __hip_atomic_load(pt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int spin = 0;
            while ((v >> 30) == 0u) {
                if (++spin > OS_SPIN) {
                    *fault = 1u;
                    s_bad = 1u;
                    v = OS_INC;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                v = __hip_atomic_load(pt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            excl += v & OS_MASK;
            if ((v >> 30) == 2u) break;
        }
        __hip_atomic_store(st, OS_INC | (excl + cnt), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    goff[dg] = dstart[dg] + excl - lstart;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RI; ++r) {
        if (base + (uint64_t)r * 64 < n) {
            uint32_t d = (uint32_t)((key[r] >> shift) & 255u);
            uint32_t pos = wc[wave][d] + rank[r];
            s_key[pos] = key[r];
            s_val[pos] = val[r];
        }
    }
    __syncthreads();
    if (s_bad) return;  // offsets unknown: write nothing (the fault flag reports it)
    const uint32_t nvalid = (uint32_t)((n - tbase) < (uint64_t)RT ? (n - tbase) : (uint64_t)RT);
    for (uint32_t i = threadIdx.x; i < nvalid; i += RB) {
        K k = s_key[i];
        uint32_t d = (uint32_t)((k >> shift) & 255u);
        uint32_t g = goff[d] + i;
        if (g < n) {
            kout[g] = k;
            vout[g] = s_val[i];
        }
    }
}

// ---------------------------------------------------------------- gather / hierarchy + boxes

// Common-prefix length of sorted leaves i and j: of the keys, or key bits + that of the indices
// when the keys are equal (Karras 2012, sec. 4: duplicates ordered by index).
template <typename K>
__device__ __forceinline__ int kdelta(const K* __restrict__ keys, int64_t i, int64_t j) {
    K a = keys[i], b = keys[j];
    if (a == b) return (int)(8 * sizeof(K)) + __clz((uint32_t)i ^ (uint32_t)j);
    if constexpr (sizeof(K) == 8)
        return __clzll((long long)(a ^ b));
    else
        return __clz((int)(a ^ b));
}

// delta(k) of leaves k and k + 1; -1 past either end of [0, n - 2].
template <typename K>
__device__ __forceinline__ int split_delta(const K* __restrict__ keys, int64_t n, int64_t k) {
    if (k < 0 || k >= n - 1) return -1;
    return kdelta(keys, k, k + 1);
}

// The hierarchy and its boxes in one bottom-up pass. A node covering leaves [l, r] is the left
// child of the node split between r and r + 1 when delta(r) > delta(l - 1), otherwise the right
// child of the node split between l - 1 and l (the parent rule of Apetrei 2014, "Fast and simply
// scalable bounding volume hierarchy construction"; the two deltas never tie for distinct
// augmented keys). Nodes carry the numbering of Karras 2012 — the root is 0, a left child is
// numbered by its last leaf and a right child by its first — so the tree and its numbering are
// those of the top-down construction.
//
// Each thread climbs from a leaf. Two children meet at their parent's split position s; the
// first to arrive parks its box, reference, height and far range end there and stops, the
// second merges, writes nothing until the merged node knows its own side (which fixes its
// number), then writes the whole 64-byte node and climbs on. A device-scope fence per climb step
// would write back L2 on every step, so the meeting places are split by where they lie:
//   chunk:  workgroup g owns leaves [g*BC, (g+1)*BC) (BC = 1024, 512 or 256, a launch option);
//           splits between two of its leaves meet in
//           LDS (a workgroup-scope atomic decides the order). A split at a chunk boundary, or an
//           in-chunk split whose other child never arrives (it straddles the boundary), meets in
//           global memory, meet[s], where a device-scope counter flags[s] counts the arrivals;
//           the second arrival appends s to the first worklist.
//   rounds: each worklist split merges its two records, writes its node, delivers to its own
//           parent's meet record and bumps that counter; the second arrival appends the parent's
//           split to the next worklist. Kernel boundaries make the records of earlier rounds
//           visible.
//   finish: once a worklist is small, one workgroup runs the remaining rounds itself with
//           barriers between them (workgroup scope, no cache writeback).
// Straddling nodes are a small fraction: the ancestors of one node per chunk boundary.
// Arrival at a global meeting place: every split receives exactly two arrivals per build, so an
// exchange of the build's generation number gen into flags[s] tells them apart (the second one
// reads gen back) without clearing the 4n-byte array between builds; it starts as all ones. The
// level-synchronous chunk pass ahead of span passes exchanges nothing: no split it delivers to
// globally can get its other arrival in that pass except from the neighbouring chunk at a chunk
// boundary, so it stores gen (per side in flags2 at boundaries) and the later passes read it.
constexpr int GEN_SLOT = BOX_ROUNDS + 3;  // box_cnt slot holding the generation number
constexpr int ROUNDS_SLOT = BOX_ROUNDS + 2;  // box_cnt slot: kernel rounds this build needed

__device__ __forceinline__ void leaf_box(const float4 s, float b[6]) {
    b[0] = s.x - s.w; b[1] = s.y - s.w; b[2] = s.z - s.w;
    b[3] = s.x + s.w; b[4] = s.y + s.w; b[5] = s.z + s.w;
}

// Whole node: the two child boxes, the child references, the pair flags and the height, one
// 64-byte line. A node whose two children are leaves (a leaf pair: leaves k and k + 1, height 1)
// is not stored: its parent references it as leaf k and sets the pair flag of that side
// (bit 0: child 0, bit 1: child 1), and a traversal tests both spheres there. The root is always
// stored. pf(): the flags from the children's references and heights (a pair is the only leaf
// reference of height 1).
__device__ __forceinline__ int pair_flags(int2 k, int h0, int h1) {
    return ((k.x < 0 && h0 == 1) ? 1 : 0) | ((k.y < 0 && h1 == 1) ? 2 : 0);
}
__device__ __forceinline__ void store_node(Node* nd, const float f[12], int2 k, int pf, int h) {
    nd->a = make_float4(f[0], f[1], f[2], f[3]);
    nd->b = make_float4(f[4], f[5], f[6], f[7]);
    nd->c = make_float4(f[8], f[9], f[10], f[11]);
    nd->d = make_int4(k.x, k.y, pf, h);
}

// One child's record at a global meeting place: box in slot `side` of a,b,c; d = (ref0, ref1,
// height0, height1); mf = (far0, far1, dx0, dx1) with far[side] the child's range end away from
// the split and dx[side] the one delta the merged parent needs from this child (delta(l - 1) of a
// left child, delta(r) of a right child), so that merging reads no keys.
__device__ __forceinline__ void meet_put(Node* mt, int4* mf, int side, const float b[6], int32_t ref, int h,
                                         int32_t far, int dx) {
    float* f = reinterpret_cast<float*>(mt) + side * 6;
#pragma unroll
    for (int a = 0; a < 6; ++a) f[a] = b[a];
    int* d = reinterpret_cast<int*>(&mt->d);
    d[side] = ref;
    d[2 + side] = h;
    int* m = reinterpret_cast<int*>(mf);
    m[side] = far;
    m[2 + side] = dx;
}

// Block-aggregated worklist append: up to 256 entries gathered in LDS, one global atomic per
// block; entries past that go straight to the list.
struct SplitList {
    uint32_t* s_list;
    uint32_t* s_n;
    uint32_t* s_base;
    __device__ __forceinline__ void push(uint32_t v, uint32_t* list, uint32_t* cnt) {
        uint32_t q = atomicAdd(s_n, 1u);
        if (q < 256)
            s_list[q] = v;
        else
            list[atomicAdd(cnt, 1u)] = v;
    }
    __device__ __forceinline__ void flush(uint32_t* list, uint32_t* cnt) {
        __syncthreads();
        uint32_t nl = *s_n < 256u ? *s_n : 256u;
        if (threadIdx.x == 0 && nl) *s_base = atomicAdd(cnt, nl);
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < nl; k += blockDim.x) list[*s_base + k] = s_list[k];
        __syncthreads();
        if (threadIdx.x == 0) *s_n = 0;
        __syncthreads();
    }
};

// The chunk pass, one thread per leaf; n >= 2.
template <typename K, int BC>
__global__ __launch_bounds__(BC) void k_bu_chunk(const K* __restrict__ keys, const float4* __restrict__ leaf_sph,
                                                 int64_t n, Node* __restrict__ nodes, Node* meet, int4* meet_far,
                                                 uint32_t* flags, uint32_t* list, uint32_t* cnt, SceneInfo* info) {
    __shared__ float s_box[BC][12];  // the two child boxes parked at each in-chunk split
    __shared__ int32_t s_ref[BC][2];
    __shared__ int32_t s_far[BC][2];
    __shared__ int s_delta[BC + 1];  // delta(cs - 1 + j): every delta a climb inside the chunk reads
    __shared__ uint8_t s_h[BC][2];
    __shared__ uint8_t s_side[BC];   // side of the parked child (meaningful with one arrival)
    __shared__ uint32_t s_cnt[BC];   // arrivals per in-chunk split
    __shared__ uint32_t s_list[256];
    __shared__ uint32_t s_n, s_base;
    SplitList ap{s_list, &s_n, &s_base};
    const uint32_t gen = cnt[GEN_SLOT];
    const int pair_h = (int)info->pair_h;
    const int64_t cs = (int64_t)blockIdx.x * BC;
    const int64_t ce = cs + BC < n ? cs + BC : n;
    const int64_t k = cs + threadIdx.x;
    const float4 sph = k < ce ? leaf_sph[k] : make_float4(0.f, 0.f, 0.f, 0.f);  // in flight over the barrier
    s_cnt[threadIdx.x] = 0;
    for (int j = threadIdx.x; j <= BC; j += BC) s_delta[j] = split_delta(keys, n, cs - 1 + j);
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    if (k < ce) {
        float b[6];
        leaf_box(sph, b);
        float f[12];
        int2 kd = make_int2(0, 0);
        int pf = 0;
        int h = 0;
        int64_t l = k, r = k;
        int32_t ref = ~(int32_t)k;
        bool internal = false;
        while (true) {
            if (l == 0 && r == n - 1) {  // the root
                store_node(nodes, f, kd, pf, h);
                info->height = (uint32_t)h;
                break;
            }
            const int dl = s_delta[l - cs], dr = s_delta[r + 1 - cs];  // l - 1 >= cs - 1, r <= ce - 1
            const bool left = dr > dl;
            const int64_t s = left ? r : l - 1;
            if (internal) {
                if (h == pair_h) {  // a leaf pair: referenced as its first leaf, not stored
                    ref = ~(int32_t)l;
                } else {
                    const int64_t idx = left ? r : l;
                    ref = (int32_t)idx;
                    store_node(nodes + idx, f, kd, pf, h);
                }
            }
            const int side = left ? 0 : 1;
            const int32_t far = (int32_t)(left ? l : r);
            if (s < cs || s + 1 >= ce) {  // the split lies on a chunk boundary
                meet_put(meet + s, meet_far + s, side, b, ref, h, far, left ? dl : dr);
                if (atomicExch(flags + s, gen) == gen) ap.push((uint32_t)s, list, cnt + 1);
                break;
            }
            const int slot = (int)(s - cs);
#pragma unroll
            for (int a = 0; a < 6; ++a) s_box[slot][side * 6 + a] = b[a];
            s_ref[slot][side] = ref;
            s_far[slot][side] = far;
            s_h[slot][side] = (uint8_t)h;
            s_side[slot] = (uint8_t)side;
            uint32_t old = __hip_atomic_fetch_add(&s_cnt[slot], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old == 0u) break;  // the sibling's thread merges
#pragma unroll
            for (int a = 0; a < 12; ++a) f[a] = s_box[slot][a];
#pragma unroll
            for (int a = 0; a < 6; ++a) f[side * 6 + a] = b[a];
            kd = side == 0 ? make_int2(ref, s_ref[slot][1]) : make_int2(s_ref[slot][0], ref);
            const int hs = s_h[slot][1 - side];
            pf = side == 0 ? pair_flags(kd, h, hs) : pair_flags(kd, hs, h);
            const int64_t fo = s_far[slot][1 - side];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                b[a] = fminf(f[a], f[6 + a]);
                b[3 + a] = fmaxf(f[3 + a], f[9 + a]);
            }
            h = 1 + (h > hs ? h : hs);
            if (side == 0)
                r = fo;
            else
                l = fo;
            internal = true;
        }
    }
    __syncthreads();
    // in-chunk splits with one arrival: the other child straddles the chunk, so the parked
    // child moves to the global meeting place
    const int nsplit = (int)(ce - cs) - 1;
    for (int slot = threadIdx.x; slot < nsplit; slot += BC) {
        if (s_cnt[slot] != 1u) continue;
        const int side = s_side[slot];
        const int64_t s = cs + slot;
        float b[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) b[a] = s_box[slot][side * 6 + a];
        const int64_t fr = s_far[slot][side];  // delta(l - 1) of a left child, delta(r) of a right one
        meet_put(meet + s, meet_far + s, side, b, s_ref[slot][side], s_h[slot][side], (int32_t)fr,
                 s_delta[side == 0 ? fr - cs : fr + 1 - cs]);
        if (atomicExch(flags + s, gen) == gen) ap.push((uint32_t)s, list, cnt + 1);
    }
    ap.flush(list, cnt + 1);
}

// The chunk pass, level-synchronous and persistent: the same tree, numbering and meeting rules as
// k_bu_chunk, but the climb runs in rounds over compact LDS lists of the splits whose two
// children have arrived, so a wavefront's lanes stay busy instead of following the few long
// climbs. Workgroup g takes chunks g, g + G, g + 2G, ... (G = the grid) and also gathers the leaf
// spheres into Morton order (leaf_sph[k] = sph[ids[k]]). Round 0 delivers the leaves; round r
// merges the splits completed in round r - 1, writes their nodes and delivers them to their
// parents, and the active items stay packed in the first waves. LDS records are chunk-relative
// 16-bit (a reference is the leaf or node index minus the chunk start, bit 15 marking a leaf).
// Rounds touch LDS only: deliveries to the two boundary splits of a chunk write their global
// meeting record at once but take their arrival count after the rounds, together with the
// in-chunk splits left with one arrival. The loads of the next chunk's spheres and keys (and of
// the ids one chunk further) are issued just before those arrival counts, so that the one wait
// for global memory per chunk covers both. n >= 2.
// Two 1024-thread workgroups per CU (their LDS allows two) are 8 waves per SIMD: at most 64 VGPRs.
template <typename K, int BC>
__global__ __launch_bounds__(BC) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_bu_chunk_ls(
    const K* __restrict__ keys, const uint32_t* __restrict__ ids, const float4* __restrict__ sph,
    float4* __restrict__ leaf_sph, int64_t n64, Node* __restrict__ nodes, Node* meet, int4* meet_far, uint32_t* flags,
    uint32_t* flags2, uint32_t* list, uint32_t* cnt, SceneInfo* info, int wave_items) {
    __shared__ float s_box[BC][12];     // the two child boxes parked at each in-chunk split
    __shared__ uint16_t s_ref[BC][2];   // chunk-relative child references
    __shared__ uint16_t s_far[BC][2];   // chunk-relative far range ends
    __shared__ int8_t s_delta[BC + 1];  // delta(cs - 1 + j)
    // per side: the child's height (low byte) and the one delta its parent needs (high byte: delta(l - 1)
    // of a left child, delta(r) of a right child); both sides are read with one 4-byte load
    __shared__ uint16_t s_hd[BC][2];
    __shared__ uint32_t s_cnt[BC];      // arrivals: + 1 for side 0, + 0x10000 for side 1
    __shared__ uint16_t s_lst[2][BC];   // splits completed in the previous / current round
    __shared__ uint32_t s_nl[3];        // list sizes: read in round r, appended in r + 1, cleared in r + 2
    __shared__ int32_t s_pend[2];       // boundary splits delivered to in this chunk (left, right)
    __shared__ uint32_t s_list[256];
    __shared__ uint32_t s_n, s_base;
    __shared__ float4 s_nb[BC / 64];    // the sphere of each wavefront's first leaf (its left neighbour's partner)
    SplitList ap{s_list, &s_n, &s_base};
    const uint32_t gen = cnt[GEN_SLOT];
    const int pair_h = (int)info->pair_h;
    const int32_t n = (int32_t)n64;
    const int32_t nchunks = (n + BC - 1) / BC;
    const int32_t tid = (int32_t)threadIdx.x;
    const int32_t G = (int32_t)gridDim.x;
    if (tid == 0) s_n = 0;
    // prefetch pipeline: ids of chunk + 2G, spheres and keys of chunk + G. Every load is
    // unconditional, at an index clamped into [0, n - 1]; validity is applied where the values
    // are used, so that no load result merges with a constant at a branch join.
    auto clampi = [&](int64_t i) -> int64_t { return i < 0 ? 0 : (i > n64 - 1 ? n64 - 1 : i); };
    auto load_id = [&](int32_t ch) -> uint32_t { return ids[clampi((int64_t)ch * BC + tid)]; };
    // keys i, i + 1 and i + 2 for i = cs - 1 + tid: delta(i), and delta(i + 1) for the last
    // thread (lane-dependent addresses keep these vector loads)
    auto load_keys = [&](int32_t ch, K& ka, K& kb, K& kc) {
        const int64_t i = (int64_t)ch * BC - 1 + tid;
        ka = keys[clampi(i)];
        kb = keys[clampi(i + 1)];
        kc = keys[clampi(i + 2)];
    };
    int32_t ch = (int32_t)blockIdx.x;
    uint32_t id_cur = load_id(ch);
    float4 sp_cur = sph[id_cur];
    K ka_cur, kb_cur, kc_cur;
    load_keys(ch, ka_cur, kb_cur, kc_cur);
    uint32_t id_next = load_id(ch + G);
    for (; ch < nchunks; ch += G) {
        const int32_t cs = ch * BC;
        const int32_t ce = cs + BC < n ? cs + BC : n;
        const int32_t k = cs + tid;
        if (k < ce) {
            leaf_sph[k] = sp_cur;
        }
        // chunk set-up: delta(cs - 1 + j) for j in [0, BC] from the keys in registers
        {
            // delta(i) of keys a = key i, b = key i + 1 (-1 past either end), as split_delta
            auto dl = [&](int32_t i, K a, K b) -> int {
                if (i < 0 || i + 1 >= n) return -1;
                if (a == b) return (int)(8 * sizeof(K)) + __clz((uint32_t)i ^ (uint32_t)(i + 1));
                if constexpr (sizeof(K) == 8)
                    return __clzll((long long)(a ^ b));
                else
                    return __clz((int)(a ^ b));
            };
            const int32_t i = cs - 1 + tid;
            s_delta[tid] = (int8_t)dl(i, ka_cur, kb_cur);
            if (tid == BC - 1) s_delta[BC] = (int8_t)dl(i + 1, kb_cur, kc_cur);
            if (tid == 0) {
                s_nl[0] = 0;
                s_nl[1] = 0;
                s_nl[2] = 0;
                s_pend[0] = -1;
                s_pend[1] = -1;
            }
        }
        s_cnt[tid] = 0;
        if ((tid & 63) == 0) s_nb[tid >> 6] = sp_cur;
        __syncthreads();
        // deliver node X = [l, r] (box b, height h, dl = delta(l - 1), dr = delta(r); its
        // children's boxes f and references kd when internal) to its parent. When this completes an
        // in-chunk split, the split goes to lst_next, or with lst_next null (a climbing lane) its
        // slot is returned for this lane to merge next; otherwise -1.
        auto deliver = [&](int32_t l, int32_t r, int dl, int dr, const float b[6], int h, bool internal,
                           const float f[12], int2 kd, int pf, int32_t leaf, uint32_t* nl_next,
                           uint16_t* lst_next, bool xw = false) -> int {
            if (l == 0 && r == n - 1) {  // the root
                store_node(nodes, f, kd, pf, h);
                info->height = (uint32_t)h;
                return -1;
            }
            if (internal && h == pair_h) {  // a leaf pair: delivered as its first leaf (height 1), not stored
                internal = false;
                leaf = l;
            }
            const bool left = dr > dl;
            const int32_t sp = left ? r : l - 1;
            int32_t ref;
            uint16_t rrel;
            if (internal) {
                const int32_t idx = left ? r : l;
                store_node(nodes + idx, f, kd, pf, h);
                ref = idx;
                rrel = (uint16_t)(idx - cs);
            } else {
                ref = ~leaf;
                rrel = (uint16_t)(0x8000u | (uint32_t)(leaf - cs));
            }
            const int side = left ? 0 : 1;
            const int32_t far = left ? l : r;
            if (sp < cs || sp + 1 >= ce) {  // a boundary split: sp = cs - 1 or ce - 1
                meet_put(meet + sp, meet_far + sp, side, b, ref, h, far, left ? dl : dr);
                s_pend[sp < cs ? 0 : 1] = sp;
                return -1;
            }
            const int slot = sp - cs;
#pragma unroll
            for (int a = 0; a < 6; ++a) s_box[slot][side * 6 + a] = b[a];
            s_ref[slot][side] = rrel;
            s_far[slot][side] = (uint16_t)(far - cs);
            // the parent's own dl (left) or dr (right) beside the height
            s_hd[slot][side] = (uint16_t)((uint32_t)h | ((uint32_t)(uint8_t)(int8_t)(left ? dl : dr) << 8));
            // a round's merge happens in the next round, after a barrier, and a climbing lane reads
            // the record right after the count. Within one wavefront LDS executes its accesses in
            // issue order: a relaxed count suffices, with compiler fences that keep the record's
            // writes before it and the climb's reads after it. Lanes of several wavefronts climbing
            // together (xw) take an acquire-release count. Each side adds its own unit, so the count
            // also tells which side a lone arrival came from.
            uint32_t old;
            if (xw) {
                old = __hip_atomic_fetch_add(&s_cnt[slot], side ? 0x10000u : 1u, __ATOMIC_ACQ_REL,
                                             __HIP_MEMORY_SCOPE_WORKGROUP);
            } else {
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                old = __hip_atomic_fetch_add(&s_cnt[slot], side ? 0x10000u : 1u, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
            }
            if (old == 0u) return -1;
            if (lst_next == nullptr) return slot;
            // (a wavefront-aggregated append, one counter update per wavefront, measured 2% slower)
            lst_next[atomicAdd(nl_next, 1u)] = (uint16_t)slot;
            return -1;
        };
        auto gref = [&](uint16_t rr) -> int32_t {
            return (rr & 0x8000u) ? ~(cs + (int32_t)(rr & 0x7fffu)) : cs + (int32_t)rr;
        };
        // the leaves. Leaves k and k + 1 whose parent is split k (delta(k) above delta(k - 1) and
        // delta(k + 1)) are merged here by the left leaf's thread, exactly as a round would merge
        // them, so that the most numerous merges (the nodes over two leaves) take no round and no
        // LDS records; the right leaf's thread delivers nothing. The right sphere comes from the next
        // lane, or for a wavefront's last lane from the next wavefront's first (s_nb). A pair whose
        // split is a chunk boundary is left to the global meeting place like any boundary split.
        if (k < ce) {
            const int dkm1 = s_delta[tid], dk = s_delta[tid + 1];
            const int dkp1 = s_delta[tid + 2 <= BC ? tid + 2 : BC];  // read only when k + 1 < ce
            const int dkm2 = s_delta[tid >= 1 ? tid - 1 : 0];         // read only when k - 1 >= cs
            const bool pair_l = k + 1 < ce && dk > dkm1 && dk > dkp1;
            const bool pair_r = k - 1 >= cs && dkm1 > dkm2 && dkm1 > dk;
            float4 nb;
            nb.x = __shfl_down(sp_cur.x, 1, 64);
            nb.y = __shfl_down(sp_cur.y, 1, 64);
            nb.z = __shfl_down(sp_cur.z, 1, 64);
            nb.w = __shfl_down(sp_cur.w, 1, 64);
            if ((tid & 63) == 63 && pair_l) nb = s_nb[(tid >> 6) + 1];  // k + 1 < ce: a later wavefront's leaf
            float b[6];
            leaf_box(sp_cur, b);
            if (pair_l) {
                float f[12], u[6];
                leaf_box(nb, f + 6);
#pragma unroll
                for (int a = 0; a < 6; ++a) f[a] = b[a];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    u[a] = fminf(f[a], f[6 + a]);
                    u[3 + a] = fmaxf(f[3 + a], f[9 + a]);
                }
                const int2 kd = make_int2(~k, ~(k + 1));
                deliver(k, k + 1, dkm1, dkp1, u, 1, true, f, kd, pair_flags(kd, 0, 0), 0, &s_nl[0], s_lst[0]);
            } else if (!pair_r) {
                float f0[12];
                deliver(k, k, dkm1, dk, b, 0, false, f0, make_int2(0, 0), 0, k, &s_nl[0], s_lst[0]);
            }
        }
        // merge the two children parked at a completed in-chunk split and deliver the node
        auto merge = [&](int slot, uint32_t* nl_next, uint16_t* lst_next, bool xw = false) -> int {
            float f[12];
#pragma unroll
            for (int a = 0; a < 12; ++a) f[a] = s_box[slot][a];
            const int2 kd = make_int2(gref(s_ref[slot][0]), gref(s_ref[slot][1]));
            const uint32_t hd = *reinterpret_cast<const uint32_t*>(s_hd[slot]);
            const int h0 = (int)(hd & 0xffu), h1 = (int)((hd >> 16) & 0xffu);
            const int h = 1 + (h0 > h1 ? h0 : h1);
            const int32_t l = cs + (int32_t)s_far[slot][0], r = cs + (int32_t)s_far[slot][1];
            const int dl = (int)(int8_t)(hd >> 8), dr = (int)(int8_t)(hd >> 24);
            float b[6];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                b[a] = fminf(f[a], f[6 + a]);
                b[3 + a] = fmaxf(f[3 + a], f[9 + a]);
            }
            return deliver(l, r, dl, dr, b, h, true, f, kd, pair_flags(kd, h0, h1), 0, nl_next, lst_next, xw);
        };
        __syncthreads();
        for (int rd = 0;; ++rd) {
            const int c0 = rd % 3;
            const uint32_t count = s_nl[c0];
            if (count == 0) break;
            if (count <= (uint32_t)wave_items) {
                // the last levels without rounds: each lane of the first waves takes one of the last
                // splits and climbs while its delivery completes the parent (within one wavefront LDS
                // executes accesses in issue order, so the other half of a parent, written before its
                // count by a lane of the same wavefront, is there when the count says so; lanes of
                // several wavefronts use acquire-release counts). No list, counter or fence per
                // level, and no workgroup barrier until the chunk's top.
                const bool xw = count > 64u;
                if ((uint32_t)tid < ((count + 63u) & ~63u)) {
                    int slot = (uint32_t)tid < count ? (int)s_lst[rd & 1][tid] : -1;
                    while (slot >= 0) slot = merge(slot, nullptr, nullptr, xw);
                }
                __syncthreads();
                break;
            }
            if (tid == 0) s_nl[(rd + 2) % 3] = 0;
            uint32_t* nl_next = &s_nl[(rd + 1) % 3];
            const uint16_t* lst = s_lst[rd & 1];
            uint16_t* lst_next = s_lst[(rd + 1) & 1];
            for (uint32_t i = (uint32_t)tid; i < count; i += BC) merge(lst[i], nl_next, lst_next);
            __syncthreads();
        }
        // next chunk's loads (its ids are already here; the ids after it start now), issued before
        // the arrival counts below so that one wait covers both
        const float4 sp_next = sph[id_next];
        K ka_next, kb_next, kc_next;
        load_keys(ch + G, ka_next, kb_next, kc_next);
        const uint32_t id_next2 = load_id(ch + 2 * G);
        // arrival counts at the global meeting places: the boundary splits delivered to, and the
        // in-chunk splits with one arrival (their other child straddles the chunk), whose parked
        // child moves to the global record first. With flags2 no count is exchanged: no other
        // child of these splits can arrive in this pass (a straddler arrives in a later pass, a
        // neighbour chunk's node at a boundary split only here), so the arrivals are recorded
        // with plain stores, a boundary split's per side (flags2[2k + side] for split k * BC - 1),
        // and the later passes tell which splits the chunks completed.
        if (tid < 2 && s_pend[tid] >= 0) {
            const int32_t sp = s_pend[tid];
            if (flags2)  // left boundary (tid 0): its right side; right boundary: its left side
                flags2[2 * (ch + tid) + 1 - tid] = gen;
            else if (atomicExch(flags + sp, gen) == gen)
                ap.push((uint32_t)sp, list, cnt + 1);
        }
        const int nsplit = ce - cs - 1;
        for (int slot = tid; slot < nsplit; slot += BC) {
            const uint32_t ca = s_cnt[slot];
            if (ca != 1u && ca != 0x10000u) continue;
            const int side = ca == 1u ? 0 : 1;
            const int32_t sp = cs + slot;
            float b[6];
#pragma unroll
            for (int a = 0; a < 6; ++a) b[a] = s_box[slot][side * 6 + a];
            const uint32_t hd = s_hd[slot][side];
            meet_put(meet + sp, meet_far + sp, side, b, gref(s_ref[slot][side]), (int)(hd & 0xffu),
                     cs + (int32_t)s_far[slot][side], (int)(int8_t)(hd >> 8));
            if (flags2)
                flags[sp] = gen;
            else if (atomicExch(flags + sp, gen) == gen)
                ap.push((uint32_t)sp, list, cnt + 1);
        }
        if (flags2) {
            // only LDS needs ordering here: the stores above are for later passes, and the next
            // chunk's loads are waited for where they are used
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_s_barrier();
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
        } else {
            __syncthreads();
            if (s_n > 192u) ap.flush(list, cnt + 1);  // s_n is read by every thread after the barrier
        }
        // rotate the pipeline
        id_cur = id_next;
        id_next = id_next2;
        sp_cur = sp_next;
        ka_cur = ka_next;
        kb_cur = kb_next;
        kc_cur = kc_next;
    }
    ap.flush(list, cnt + 1);
}

// Merge the two records at split s, write the node, deliver it to its parent's meeting place.
// Returns the parent's split when this delivery completes it, else -1.
// flags2 (with chunk, the chunk pass's leaves per chunk): that pass's arrivals at the chunk
// boundary splits k * chunk - 1, stored per side as flags2[2k + side] = gen rather than
// exchanged at flags, so a delivery there also completes the split when the other side's entry
// holds gen. Null: every earlier arrival was exchanged at flags.
template <typename K, bool VOL>
__device__ __forceinline__ int64_t bu_step(const K* __restrict__ keys, int64_t n, Node* nodes, Node* meet,
                                           int4* meet_far, uint32_t* flags, int64_t s, SceneInfo* info,
                                           uint32_t gen, const uint32_t* flags2 = nullptr, int64_t chunk = 1) {
    // The two records at s were written by an earlier kernel (visible since this launch began) or,
    // within one workgroup's levels (VOL), by this workgroup before the barrier that ended the
    // previous level: waves of a workgroup run on one CU, whose vector L1 they share, so plain
    // loads see those stores. (System-scope loads here bypassed both cache levels one dependent
    // access at a time: 20 round trips to memory per merge.)
    (void)VOL;
    float f[12];
    int dv[4];
    int fr[4];  // far0, far1, dx0, dx1
    {
        const float4* src = reinterpret_cast<const float4*>(meet + s);
        const float4 a = src[0], b = src[1], c = src[2];
        const int4 d = meet[s].d;
        const int4 ff = meet_far[s];
        f[0] = a.x; f[1] = a.y; f[2] = a.z; f[3] = a.w;
        f[4] = b.x; f[5] = b.y; f[6] = b.z; f[7] = b.w;
        f[8] = c.x; f[9] = c.y; f[10] = c.z; f[11] = c.w;
        dv[0] = d.x; dv[1] = d.y; dv[2] = d.z; dv[3] = d.w;
        fr[0] = ff.x;
        fr[1] = ff.y;
        fr[2] = ff.z;
        fr[3] = ff.w;
    }
    const int2 kd = make_int2(dv[0], dv[1]);
    const int h = 1 + (dv[2] > dv[3] ? dv[2] : dv[3]);
    const int pf = pair_flags(kd, dv[2], dv[3]);
    const int64_t l = fr[0], r = fr[1];
    if (l == 0 && r == n - 1) {
        store_node(nodes, f, kd, pf, h);
        info->height = (uint32_t)h;
        return -1;
    }
    float b[6];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        b[a] = fminf(f[a], f[6 + a]);
        b[3 + a] = fmaxf(f[3 + a], f[9 + a]);
    }
    const int dl = fr[2], dr = fr[3];  // delta(l - 1), delta(r), carried by the children
    const bool left = dr > dl;
    int32_t ref;
    if (h == (int)info->pair_h) {  // a leaf pair: referenced as its first leaf, not stored
        ref = ~(int32_t)l;
    } else {
        const int64_t idx = left ? r : l;
        store_node(nodes + idx, f, kd, pf, h);
        ref = (int32_t)idx;
    }
    const int64_t s2 = left ? r : l - 1;
    meet_put(meet + s2, meet_far + s2, left ? 0 : 1, b, ref, h, (int32_t)(left ? l : r), left ? dl : dr);
    const bool bnd = flags2 != nullptr && (s2 + 1) % chunk == 0;
    const uint32_t other = bnd ? flags2[((s2 + 1) / chunk) * 2 + (left ? 1 : 0)] : 0u;
    const uint32_t x = atomicExch(flags + s2, gen);
    return (x == gen || (bnd && other == gen)) ? s2 : -1;
}

template <typename K>
__device__ void bu_finish(const K* __restrict__ keys, int64_t n, Node* nodes, Node* meet, int4* meet_far,
                          uint32_t* flags, uint32_t* la, uint32_t* lb, uint32_t n_cur, uint32_t* cnt, SceneInfo* info) {
    __shared__ uint32_t s_out;
    const uint32_t gen = cnt[GEN_SLOT];
    uint32_t* lin = la;
    uint32_t* lout = lb;
    while (n_cur > 0) {
        if (threadIdx.x == 0) s_out = 0;
        __syncthreads();
        for (uint32_t base = 0; base < n_cur; base += blockDim.x) {
            uint32_t i = base + threadIdx.x;
            if (i < n_cur) {
                int64_t s = (int64_t)lin[i];
                int64_t s2 = bu_step<K, true>(keys, n, nodes, meet, meet_far, flags, s, info, gen);
                if (s2 >= 0) lout[atomicAdd(&s_out, 1u)] = (uint32_t)s2;
            }
        }
        __threadfence_block();
        __syncthreads();
        n_cur = s_out;
        uint32_t* t = lin;
        lin = lout;
        lout = t;
        __syncthreads();
    }
    if (threadIdx.x == 0) cnt[0] = 1u;
}

// Round r: worklist r lives in list[r & 1], its size in cnt[1 + r]. cnt[0]: 0 while rounds run,
// 2 once a round found its list no longer than tail_items (that round's number in
// cnt[ROUNDS_SLOT]) and left it to the tail, 1 when the tree is complete.
template <typename K>
__global__ __launch_bounds__(256) void k_bu_round(const K* __restrict__ keys, int64_t n, Node* nodes, Node* meet,
                                                  int4* meet_far, uint32_t* flags, uint32_t* l0, uint32_t* l1,
                                                  uint32_t* cnt, int r, uint32_t tail_items, SceneInfo* info) {
    __shared__ uint32_t s_list[256];
    __shared__ uint32_t s_n, s_base;
    if (cnt[0]) return;
    const uint32_t gen = cnt[GEN_SLOT];
    uint32_t n_in = cnt[1 + r];
    uint32_t* lin = (r & 1) ? l1 : l0;
    uint32_t* lout = (r & 1) ? l0 : l1;
    if (n_in <= tail_items) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            cnt[ROUNDS_SLOT] = (uint32_t)r;
            cnt[0] = 2u;
        }
        return;
    }
    SplitList ap{s_list, &s_n, &s_base};
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    // one item per thread while the list is shorter than the grid: every step is a chain of
    // dependent memory accesses, so the items are spread as widely as possible
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n_in; base += stride) {
        const uint32_t i = base + threadIdx.x;
        if (i < n_in) {
            int64_t s2 = bu_step<K, false>(keys, n, nodes, meet, meet_far, flags, (int64_t)lin[i], info, gen);
            if (s2 >= 0) ap.push((uint32_t)s2, lout, cnt + 2 + r);
        }
        ap.flush(lout, cnt + 2 + r);
    }
}

// The rest in one 1024-thread workgroup: the list a round left to it, or, when all r launched
// rounds ran, list r.
template <typename K>
__global__ __launch_bounds__(1024) void k_bu_tail(const K* __restrict__ keys, int64_t n, Node* nodes, Node* meet,
                                                  int4* meet_far, uint32_t* flags, uint32_t* l0, uint32_t* l1,
                                                  uint32_t* cnt, int r, SceneInfo* info) {
    const uint32_t st = cnt[0];
    if (st == 1u) return;
    const int rr = st == 2u ? (int)cnt[ROUNDS_SLOT] : r;
    __syncthreads();  // every thread has read cnt before thread 0 writes it
    if (threadIdx.x == 0) cnt[ROUNDS_SLOT] = (uint32_t)rr;
    uint32_t* lin = (rr & 1) ? l1 : l0;
    uint32_t* lout = (rr & 1) ? l0 : l1;
    bu_finish<K>(keys, n, nodes, meet, meet_far, flags, lin, lout, cnt[1 + rr], cnt, info);
}

constexpr uint32_t SPAN_LDS = 2048;  // level-list entries a span workgroup keeps in LDS

// The levels of one workgroup's merges over the splits completed before it: each level merges
// the splits the level before completed, with one barrier between levels (three counters rotate:
// read, appended, cleared). A completed split inside [lo, hi - 2] goes to the next level, any
// other (a boundary split lo - 1 or hi - 1) to the export list xl (count *xn). The level lists
// are LDS (s_l) when level 0 fits, else the global lists gl0 and gl1. Records and lists are read
// with plain loads: the previous level's stores, completed before the barrier, are visible to
// every wave of the workgroup.
template <typename K>
__device__ void span_levels(const K* __restrict__ keys, int64_t n, Node* nodes, Node* meet, int4* meet_far,
                            uint32_t* flags, SceneInfo* info, uint32_t gen, int64_t lo, int64_t hi, uint32_t n0,
                            bool in_lds, uint32_t* gl0, uint32_t* gl1, uint32_t (*s_l)[SPAN_LDS], uint32_t* s_nc,
                            uint32_t* xl, uint32_t* xn, const uint32_t* flags2, int64_t chunk) {
    uint32_t n_cur = n0;
    for (int lv = 0; n_cur > 0; ++lv) {
        uint32_t* lin = in_lds ? s_l[lv & 1] : ((lv & 1) ? gl1 : gl0);
        uint32_t* lout = in_lds ? s_l[(lv + 1) & 1] : ((lv & 1) ? gl0 : gl1);
        uint32_t* nc_next = &s_nc[(lv + 1) % 3];
        if (threadIdx.x == 0) s_nc[(lv + 2) % 3] = 0;
        for (uint32_t i = threadIdx.x; i < n_cur; i += blockDim.x) {
            const int64_t sp = (int64_t)lin[i];
            const int64_t s2 = bu_step<K, true>(keys, n, nodes, meet, meet_far, flags, sp, info, gen, flags2, chunk);
            if (s2 < 0) continue;
            if (s2 >= lo && s2 <= hi - 2)
                lout[atomicAdd(nc_next, 1u)] = (uint32_t)s2;
            else
                xl[atomicAdd(xn, 1u)] = (uint32_t)s2;
        }
        __threadfence_block();
        __syncthreads();
        n_cur = s_nc[(lv + 1) % 3];
    }
}

// The straddlers of one span of leaves [lo, hi), one 1024-thread workgroup per span. A node whose
// split lies inside the span (lo <= s <= hi - 2) has both children inside it, so the span's
// workgroup merges it once both have arrived; a delivery that completes one of the span's two
// boundary splits is exported to xl (count cnt[2]) instead, and k_bu_top finishes the exported
// splits. Level 0 is what the chunk pass completed: with flags2 (its per-side arrival records at
// the boundaries of its chunks, chunk leaves each; span is a multiple of chunk) the chunk
// boundaries inside the span with both sides recorded, else its list, which every span scans for
// its own splits. Those at hi - 1, completed on the span's right boundary, are exported at once.
// la and lb (n entries each) hold the span's level lists at offset lo when LDS cannot.
template <typename K>
__global__ __launch_bounds__(1024) void k_bu_span(const K* __restrict__ keys, int64_t n, Node* nodes, Node* meet,
                                                  int4* meet_far, uint32_t* flags, const uint32_t* __restrict__ flags2,
                                                  int64_t chunk, const uint32_t* __restrict__ list0, uint32_t* la,
                                                  uint32_t* lb, uint32_t* xl, uint32_t* cnt, int64_t span,
                                                  SceneInfo* info) {
    __shared__ uint32_t s_l[2][SPAN_LDS];
    __shared__ uint32_t s_nc[3];
    const uint32_t gen = cnt[GEN_SLOT];
    const int64_t lo = (int64_t)blockIdx.x * span;
    const int64_t hi = lo + span < n ? lo + span : n;
    if (threadIdx.x < 3) s_nc[threadIdx.x] = 0;
    __syncthreads();
    auto level0 = [&](int64_t sp) {
        if (sp <= hi - 2) {
            const uint32_t k = atomicAdd(&s_nc[0], 1u);
            if (k < SPAN_LDS) s_l[0][k] = (uint32_t)sp;
            la[lo + k] = (uint32_t)sp;
        } else {
            xl[atomicAdd(cnt + 2, 1u)] = (uint32_t)sp;
        }
    };
    if (flags2) {
        const int64_t last = hi - 1 < n - 2 ? hi - 1 : n - 2;
        for (int64_t sp = lo + chunk - 1 + (int64_t)threadIdx.x * chunk; sp <= last; sp += (int64_t)blockDim.x * chunk) {
            const int64_t k = (sp + 1) / chunk;
            if (flags2[2 * k] == gen && flags2[2 * k + 1] == gen) level0(sp);
        }
    } else {
        const uint32_t n0 = cnt[1];
        for (uint32_t i = threadIdx.x; i < n0; i += blockDim.x) {
            const int64_t sp = (int64_t)list0[i];
            if (sp >= lo && sp <= hi - 1) level0(sp);
        }
    }
    __threadfence_block();
    __syncthreads();
    const uint32_t m = s_nc[0];
    span_levels<K>(keys, n, nodes, meet, meet_far, flags, info, gen, lo, hi, m, m <= SPAN_LDS, la + lo, lb + lo, s_l,
                   s_nc, xl, cnt + 2, flags2, chunk);
}

// The top of the tree: one workgroup merges the spans' exported splits (xl, count cnt[2]) up to
// the root; xl2 is the second global list when LDS cannot hold them. flags2 and chunk as for
// k_bu_span.
template <typename K>
__global__ __launch_bounds__(1024) void k_bu_top(const K* __restrict__ keys, int64_t n, Node* nodes, Node* meet,
                                                 int4* meet_far, uint32_t* flags, const uint32_t* __restrict__ flags2,
                                                 int64_t chunk, uint32_t* xl, uint32_t* xl2, uint32_t* cnt,
                                                 SceneInfo* info) {
    __shared__ uint32_t s_l[2][SPAN_LDS];
    __shared__ uint32_t s_nc[3];
    const uint32_t gen = cnt[GEN_SLOT];
    const uint32_t n0 = cnt[2];
    if (threadIdx.x < 3) s_nc[threadIdx.x] = 0;
    for (uint32_t i = threadIdx.x; i < n0 && i < SPAN_LDS; i += blockDim.x) s_l[0][i] = xl[i];
    __syncthreads();
    // every split lies inside [0, n - 2]: nothing is exported (cnt[3] is an unused counter here)
    span_levels<K>(keys, n, nodes, meet, meet_far, flags, info, gen, 0, n, n0, n0 <= SPAN_LDS, xl, xl2, s_l, s_nc,
                   xl2, cnt + 3, flags2, chunk);
}

// Leaf pairs of the tree over sorted keys: leaves l and l + 1 both attach to split l exactly
// when delta(l) is above delta(l - 1) and delta(l + 1). (Instrumentation: the build does not
// run this.)
template <typename K>
__global__ __launch_bounds__(256) void k_count_pairs(const K* __restrict__ keys, int64_t n,
                                                     unsigned long long* out) {
    __shared__ unsigned long long s_sum;
    if (threadIdx.x == 0) s_sum = 0;
    __syncthreads();
    unsigned long long own = 0;
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < n - 1; l += (int64_t)gridDim.x * blockDim.x) {
        const int d = split_delta(keys, n, l);
        if (d > split_delta(keys, n, l - 1) && d > split_delta(keys, n, l + 1)) ++own;
    }
    if (own) atomicAdd(&s_sum, own);
    __syncthreads();
    if (threadIdx.x == 0 && s_sum) atomicAdd(out, s_sum);
}

// Single-sphere scene: one node whose two children are the same leaf.
__global__ void k_single(const float4* __restrict__ leaf_sph, Node* nodes, SceneInfo* info) {
    float4 s = leaf_sph[0];
    float f[12] = {s.x - s.w, s.y - s.w, s.z - s.w, s.x + s.w, s.y + s.w, s.z + s.w,
                   s.x - s.w, s.y - s.w, s.z - s.w, s.x + s.w, s.y + s.w, s.z + s.w};
    store_node(nodes, f, make_int2(~0, ~0), 0, 1);
    info->height = 1;
}

__global__ void k_regather(const float4* __restrict__ sph, const uint32_t* __restrict__ leaf_id,
                           float4* __restrict__ leaf_sph, uint64_t n) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) leaf_sph[k] = sph[leaf_id[k]];
}

__global__ void k_fill(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

inline unsigned blocks_for(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// Per hierarchy pass: list counters cleared, the pass's generation number, the leaf-pair height
// (one launch instead of three fills).
__global__ void k_build_init(uint32_t* cnt, uint32_t gen, uint32_t* pair_h, uint32_t pv) {
    const uint32_t i = threadIdx.x;
    if (i < (uint32_t)(BOX_ROUNDS + 3)) cnt[i] = 0u;
    if (i == 0) {
        cnt[GEN_SLOT] = gen;
        *pair_h = pv;
    }
}

// Morton codes of the current spheres into ka, then the LSD radix sort of (key, particle index)
// over `passes` 8-bit digits; the sorted keys end in ka, the sorted values in vfinal (the last
// pass writes them there). The Morton pass also produces the first digit's tile histograms, and
// the first pass takes the particle indices as its values.
template <typename K, int RI, int DB>
int morton_sort_t(mmhrt_ctx_s* c, K* ka, K* kb, uint32_t* va, uint32_t* vb, uint32_t* vfinal, uint64_t n) {
    constexpr int RT = RB * RI, ND = 1 << DB;
    const int passes = ((sizeof(K) == 8 ? 63 : 30) + DB - 1) / DB;
    uint32_t tiles = (uint32_t)((n + RT - 1) / RT);
    uint64_t hlen = (uint64_t)tiles * ND;
    uint32_t groups = (tiles + CG - 1) / CG;
    const bool os = DB == 8 && c->opt_sort_onesweep != 0 && n < (1ull << 30);
    if (int e = ensure(c, c->hist, hlen * (os ? std::max(2, passes) : 2), "radix histogram")) return e;
    if (int e = ensure(c, c->scan_tmp, std::max((size_t)(groups + 1) * ND, (size_t)passes * 512 + 16),
                       "radix column sums"))
        return e;
    uint32_t* hist = c->hist.p;
    uint32_t* offs = c->hist.p + hlen;
    // an odd number of passes starts from kb, so that the sorted keys end in ka
    K* src = (passes & 1) ? kb : ka;
    K* dst = (passes & 1) ? ka : kb;
    rec(c, 1);
    hipLaunchKernelGGL((k_morton_tile<K, RI, DB>), dim3(tiles), dim3(RB), 0, c->stream, cur_spheres(c), n, c->info.p,
                       src, hist);
    rec(c, 2);
    uint32_t* vs = nullptr;  // first pass: values are the indices
    uint32_t* vd = vb;
    if (os) {
        // single-pass passes: status words per pass and tile, digit counts of all passes,
        // tickets and the fault flag
        if (!c->sort_fault.p) {
            if (int e = ensure(c, c->sort_fault, 1, "radix fault flag")) return e;
            launch_fill_u32(c->stream, c->sort_fault.p, 0u, 1);
        }
        uint32_t* ghist = c->scan_tmp.p;
        uint32_t* dstart = ghist + passes * 256;
        uint32_t* tickets = dstart + passes * 256;
        launch_fill_u32(c->stream, c->hist.p, 0u, hlen * (uint64_t)passes);
        launch_fill_u32(c->stream, ghist, 0u, (uint64_t)passes * 256);
        launch_fill_u32(c->stream, tickets, 0u, (uint64_t)passes);
        unsigned gb = (unsigned)std::min<uint64_t>(1024, (n + 255) / 256);
        hipLaunchKernelGGL(k_radix_ghist<K>, dim3(gb), dim3(256), 0, c->stream, src, n, passes, ghist);
        hipLaunchKernelGGL(k_radix_gscan, dim3(1), dim3(256), 0, c->stream, ghist, passes, dstart);
        for (int p = 0; p < passes; ++p) {
            hipLaunchKernelGGL((k_radix_onesweep<K, RI>), dim3(tiles), dim3(RB), 0, c->stream, src, vs, dst,
                               p == passes - 1 ? vfinal : vd, n, 8 * p, dstart + p * 256, c->hist.p + hlen * p,
                               tickets + p, c->sort_fault.p);
            std::swap(src, dst);
            vs = vd;
            vd = vd == vb ? va : vb;
        }
        return 0;
    }
    for (int p = 0; p < passes; ++p) {
        int shift = DB * p;
        if (p > 0)
            hipLaunchKernelGGL((k_radix_hist<K, RI, DB>), dim3(tiles), dim3(RB), 0, c->stream, src, n, shift, hist);
        hipLaunchKernelGGL(k_col_sums, dim3(groups, ND / 256), dim3(256), 0, c->stream, hist, tiles, c->scan_tmp.p,
                           (uint32_t)ND);
        if (ND == 256) {
            hipLaunchKernelGGL(k_col_scan, dim3(4), dim3(1024), 0, c->stream, c->scan_tmp.p, groups);
            hipLaunchKernelGGL(k_col_offsets<true>, dim3(groups), dim3(256), 0, c->stream, hist, tiles, groups,
                               c->scan_tmp.p, offs, (uint32_t)ND);
        } else {
            hipLaunchKernelGGL(k_col_top, dim3(ND / 256), dim3(256), 0, c->stream, c->scan_tmp.p, groups,
                               (uint32_t)ND);
            hipLaunchKernelGGL(k_col_digits, dim3(1), dim3(256), 0, c->stream, c->scan_tmp.p, groups, (uint32_t)ND);
            hipLaunchKernelGGL(k_col_offsets<false>, dim3(groups, ND / 256), dim3(256), 0, c->stream, hist, tiles,
                               groups, c->scan_tmp.p, offs, (uint32_t)ND);
        }
        const bool xcd = c->opt_sort_xcd != 0;
        hipLaunchKernelGGL((k_radix_scatter<K, RI, DB>), dim3(xcd ? 8 * ((tiles + 7) / 8) : tiles), dim3(RB), 0,
                           c->stream, src, vs, dst, p == passes - 1 ? vfinal : vd, n, shift, hist, offs, (int)xcd);
        std::swap(src, dst);
        vs = vd;
        vd = vd == vb ? va : vb;
    }
    return 0;  // the keys are in ka
}

// Digits of 8 bits (the default: 4 passes for 30-bit keys, 8 for 63-bit), or wide digits (option
// sort_digit 0: 3 passes of 10 bits, 6 passes of 11 bits). On the 50M FCC scene the three 10-bit
// passes took 1.20 ms against 0.97 ms for four 8-bit passes: a 10-bit scatter pass costs 265 us
// against 183 us (four keys per digit run, against sixteen, write shorter runs).
template <typename K, int RI>
int morton_sort_d(mmhrt_ctx_s* c, K* ka, K* kb, uint32_t* va, uint32_t* vb, uint32_t* vfinal, uint64_t n) {
    if (c->opt_sort_digit != 0 || c->opt_sort_onesweep != 0)
        return morton_sort_t<K, RI, 8>(c, ka, kb, va, vb, vfinal, n);
    return morton_sort_t<K, RI, sizeof(K) == 8 ? 11 : 10>(c, ka, kb, va, vb, vfinal, n);
}

template <typename K>
int morton_sort(mmhrt_ctx_s* c, K* ka, K* kb, uint32_t* va, uint32_t* vb, uint32_t* vfinal, uint64_t n) {
    if (c->opt_sort_ri == 8) return morton_sort_d<K, 8>(c, ka, kb, va, vb, vfinal, n);
    if (c->opt_sort_ri == 32 && sizeof(K) == 4) return morton_sort_d<K, 32>(c, ka, kb, va, vb, vfinal, n);
    return morton_sort_d<K, 16>(c, ka, kb, va, vb, vfinal, n);
}

// Persistent grid of the chunk pass: as many workgroups as fit on the device at once.
template <typename K, int BC>
unsigned chunk_grid(mmhrt_ctx_s* c, uint64_t n) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_bu_chunk_ls<K, BC>, BC, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    const uint64_t chunks = (n + BC - 1) / BC;
    uint64_t g = (uint64_t)per_cu * (uint64_t)(c->cu_count > 0 ? c->cu_count : 256);
    if (c->opt_build_grid > 0 && (uint64_t)c->opt_build_grid < g) g = (uint64_t)c->opt_build_grid;
    return (unsigned)(chunks < g ? chunks : g);
}

template <typename K, int BC>
void launch_chunk_ls(mmhrt_ctx_s* c, const K* keys, const uint32_t* ids, uint32_t* flags2, uint32_t* list) {
    const uint64_t n = c->n;
    hipLaunchKernelGGL((k_bu_chunk_ls<K, BC>), dim3(chunk_grid<K, BC>(c, n)), dim3(BC), 0, c->stream, keys, ids,
                       cur_spheres(c), c->leaf_sph.p, (int64_t)n, c->nodes.p,
                       c->meet.p, c->meet_far.p, c->flags.p, flags2, list, c->box_cnt.p, c->info.p,
                       (int)std::max<int64_t>(0, std::min<int64_t>(c->opt_build_wave, BC)));
}

// Hierarchy and boxes over the sorted keys, gathering the leaf spheres on the way: ids[k] is
// the particle of leaf k (the sorted values after a build, leaf_id on a refit, which keeps the
// tree of the same keys and takes the boxes of the current spheres).
template <typename K>
int hierarchy(mmhrt_ctx_s* c, const K* keys, const uint32_t* ids) {
    uint64_t n = c->n;
    hipStream_t s = c->stream;
    if (n == 1 || c->opt_build_ls == 0)
        hipLaunchKernelGGL(k_regather, dim3(blocks_for(n, 256)), dim3(256), 0, s, cur_spheres(c), ids, c->leaf_sph.p,
                           n);
    if (n == 1) {
        rec(c, 8);
        hipLaunchKernelGGL(k_single, dim3(1), dim3(1), 0, s, c->leaf_sph.p, c->nodes.p, c->info.p);
        rec(c, 9);
        return 0;
    }
    uint64_t m = n - 1;
    if (int e = ensure(c, c->box_cnt, BOX_SLOTS, "box round counters")) return e;
    if (int e = ensure(c, c->meet, m, "meeting records")) return e;
    if (int e = ensure(c, c->meet_far, m, "meeting ranges")) return e;
    const bool fresh = !c->flags.p || c->flags.cap < m;
    if (int e = ensure(c, c->flags, m, "meeting counters")) return e;
    if (fresh) launch_fill_u32(s, c->flags.p, 0xffffffffu, c->flags.cap);  // no generation yet
    // each pass takes a new generation; builds captured into a graph keep theirs, so such a
    // graph is launched once (the batched time series), never replayed
    if (++c->build_gen == 0xffffffffu) c->build_gen = 1;
    hipLaunchKernelGGL(k_build_init, dim3(1), dim3(64), 0, s, c->box_cnt.p, c->build_gen, &c->info.p->pair_h,
                       c->opt_leaf_pairs ? 1u : 0u);
    c->bvh_pairs = c->opt_leaf_pairs != 0;
    // worklists of the straddler rounds: the spare key buffer (free once the sort is done) and
    // the spare value buffer
    uint32_t* l0 = sizeof(K) == 8 ? reinterpret_cast<uint32_t*>(c->klb.p) : reinterpret_cast<uint32_t*>(c->kb.p);
    uint32_t* l1 = c->vb.p;
    // leaves per chunk workgroup: 1024 (two per CU), 512 or 256; level-synchronous rounds unless
    // the option build_ls is 0
    const int bc = c->opt_build_chunk == 256 ? 256 : (c->opt_build_chunk == 512 ? 512 : 1024);
    const int64_t nchunks = ((int64_t)n + bc - 1) / bc;
    // per-side arrival records at the chunk boundaries: the level-synchronous chunk pass followed
    // by span passes stores its boundary arrivals there instead of exchanging them
    uint32_t* flags2 = nullptr;
    if (c->opt_build_ls != 0 && c->opt_build_span > 0) {
        const uint64_t m2 = 2 * (uint64_t)(nchunks + 1);
        const bool fresh2 = !c->flags2.p || c->flags2.cap < m2;
        if (int e = ensure(c, c->flags2, m2, "chunk boundary arrivals")) return e;
        if (fresh2) launch_fill_u32(s, c->flags2.p, 0xffffffffu, c->flags2.cap);
        flags2 = c->flags2.p;
    }
    rec(c, 8);
    if (c->opt_build_ls != 0) {
        if (bc == 256)
            launch_chunk_ls<K, 256>(c, keys, ids, flags2, l0);
        else if (bc == 512)
            launch_chunk_ls<K, 512>(c, keys, ids, flags2, l0);
        else
            launch_chunk_ls<K, 1024>(c, keys, ids, flags2, l0);
    } else if (bc == 256)
        hipLaunchKernelGGL((k_bu_chunk<K, 256>), dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, c->leaf_sph.p,
                           (int64_t)n, c->nodes.p, c->meet.p, c->meet_far.p, c->flags.p, l0, c->box_cnt.p, c->info.p);
    else if (bc == 512)
        hipLaunchKernelGGL((k_bu_chunk<K, 512>), dim3(blocks_for(n, 512)), dim3(512), 0, s, keys, c->leaf_sph.p,
                           (int64_t)n, c->nodes.p, c->meet.p, c->meet_far.p, c->flags.p, l0, c->box_cnt.p, c->info.p);
    else
        hipLaunchKernelGGL((k_bu_chunk<K, 1024>), dim3(blocks_for(n, 1024)), dim3(1024), 0, s, keys, c->leaf_sph.p,
                           (int64_t)n, c->nodes.p, c->meet.p, c->meet_far.p, c->flags.p, l0, c->box_cnt.p, c->info.p);
    rec(c, 9);
    // kernel rounds: as many as the last build ran before its tail, each with a grid for the list
    // that round had (the counts come back through pinned host memory one build late; they are
    // hints only: the rounds loop over any list size and the tail finishes whatever is left)
    if (c->opt_build_span > 0) {
        // spans of build_span chunks, one workgroup each, then the tail over their exports (at
        // most three per span: its right boundary from the chunk pass, its two boundary splits).
        // Without flags2 at most 256 spans: every span's workgroup then scans the chunk pass's
        // whole list, so larger scenes take longer spans rather than more of them
        const int64_t span_chunks =
            flags2 ? c->opt_build_span : std::max<int64_t>(c->opt_build_span, (nchunks + 255) / 256);
        const int64_t span = span_chunks * (int64_t)bc;
        const uint64_t spans = (n + span - 1) / span;
        const size_t xcap = 4 * spans + 64;
        if (int e = ensure(c, c->xlist, 2 * xcap, "span export lists")) return e;
        hipLaunchKernelGGL(k_bu_span<K>, dim3((unsigned)spans), dim3(1024), 0, s, keys, (int64_t)n, c->nodes.p,
                           c->meet.p, c->meet_far.p, c->flags.p, flags2, (int64_t)bc, l0, c->va.p, c->vb.p,
                           c->xlist.p, c->box_cnt.p, span, c->info.p);
        hipLaunchKernelGGL(k_bu_top<K>, dim3(1), dim3(1024), 0, s, keys, (int64_t)n, c->nodes.p, c->meet.p,
                           c->meet_far.p, c->flags.p, flags2, (int64_t)bc, c->xlist.p, c->xlist.p + xcap,
                           c->box_cnt.p, c->info.p);
        return 0;
    }
    const uint32_t tail_items = (uint32_t)(c->opt_build_tail > 0 ? c->opt_build_tail : 1);
    // (two rounds more than the last build needed, all of them after a build that needed none or
    // when no build has run: the tail is fastest on short lists)
    int rounds = BOX_ROUNDS;
    const volatile uint32_t* last = c->rounds_host;
    const bool hint = last && last[ROUNDS_SLOT] > 0 && last[ROUNDS_SLOT] + 2 < (uint32_t)BOX_ROUNDS;
    if (hint) rounds = (int)last[ROUNDS_SLOT] + 2;
    for (int r = 0; r < rounds; ++r) {
        unsigned grid = 1024;
        if (hint && r < (int)last[ROUNDS_SLOT]) {
            const uint64_t items = last[1 + r] + last[1 + r] / 4 + 256;
            grid = (unsigned)std::min<uint64_t>(1024, (items + 255) / 256);
        }
        hipLaunchKernelGGL(k_bu_round<K>, dim3(grid), dim3(256), 0, s, keys, (int64_t)n, c->nodes.p, c->meet.p,
                           c->meet_far.p, c->flags.p, l0, l1, c->box_cnt.p, r, tail_items, c->info.p);
    }
    hipLaunchKernelGGL(k_bu_tail<K>, dim3(1), dim3(1024), 0, s, keys, (int64_t)n, c->nodes.p, c->meet.p,
                       c->meet_far.p, c->flags.p, l0, l1, c->box_cnt.p, rounds, c->info.p);
    if (c->rounds_host)
        (void)hipMemcpyAsync(c->rounds_host, c->box_cnt.p, BOX_SLOTS * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    return 0;
}

template <typename K>
int build_typed(mmhrt_ctx_s* c, K* ka, K* kb) {
    uint64_t n = c->n;
    hipStream_t s = c->stream;
    if (int e = morton_sort<K>(c, ka, kb, c->va.p, c->vb.p, c->leaf_id.p, n)) return e;
    rec(c, 3);
    rec(c, 4);  // the leaf gather runs inside the hierarchy pass
    // the sort's last pass wrote its values (the leaf -> particle map) straight into leaf_id
    return hierarchy<K>(c, ka, c->leaf_id.p);
}

}  // namespace

int count_leaf_pairs(mmhrt_ctx_s* c, int64_t* out) {
    *out = 0;
    if (!c->built || !c->bvh_pairs || c->n < 3) return 0;  // two leaves: the root, always stored
    if (int e = ensure(c, c->pair_count, 1, "pair count")) return e;
    launch_fill_u32(c->stream, reinterpret_cast<uint32_t*>(c->pair_count.p), 0u, 2);
    unsigned long long* o = reinterpret_cast<unsigned long long*>(c->pair_count.p);
    const unsigned g = 4096;
    if (c->bits == 63)
        hipLaunchKernelGGL(k_count_pairs<uint64_t>, dim3(g), dim3(256), 0, c->stream, c->kla.p, (int64_t)c->n, o);
    else
        hipLaunchKernelGGL(k_count_pairs<uint32_t>, dim3(g), dim3(256), 0, c->stream, c->ka.p, (int64_t)c->n, o);
    uint64_t h = 0;
    hipError_t e = hipMemcpyAsync(&h, c->pair_count.p, 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "count_leaf_pairs");
    *out = (int64_t)h;
    return 0;
}

void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, uint64_t n) {
    if (n == 0) return;
    unsigned g = blocks_for(n, 256);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_fill, dim3(g), dim3(256), 0, s, p, v, n);
}

int reduce_scene(mmhrt_ctx_s* c) {
    const int nb = 1024;
    if (int e = ensure(c, c->partial, (size_t)nb * NRED, "reduction partials")) return e;
    if (int e = ensure(c, c->info, 1, "scene info")) return e;
    hipLaunchKernelGGL(k_reduce_partial, dim3(nb), dim3(256), 0, c->stream, cur_spheres(c), c->n, c->partial.p);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, c->stream, c->partial.p, nb, c->info.p, c->bits);
    return 0;
}

// The first half of a build, shared by the treelet P-k-d build: scene bounds, Morton codes and
// the stable radix sort; leaf_id receives the particle of each sorted position, ka / kla the keys.
int sort_spheres(mmhrt_ctx_s* c, int bits) {
    uint64_t n = c->n;
    if (n == 0) return fail(c, MMHRT_E_STATE, "build: no spheres set");
    if (n > 0x7fffffffull) return fail(c, MMHRT_E_UNSUPPORTED, "build: more than 2^31-1 spheres");
    c->bits = bits;
    int e = 0;
    if ((e = ensure(c, c->va, n, "sort values"))) return e;
    if ((e = ensure(c, c->vb, n, "sort values"))) return e;
    if ((e = ensure(c, c->leaf_id, n, "leaf ids"))) return e;
    if (bits == 63) {
        if ((e = ensure(c, c->kla, n, "keys"))) return e;
        if ((e = ensure(c, c->klb, n, "keys"))) return e;
    } else {
        if ((e = ensure(c, c->ka, n, "keys"))) return e;
        if ((e = ensure(c, c->kb, n, "keys"))) return e;
    }
    rec(c, 0);
    if ((e = reduce_scene(c))) return e;
    if (bits == 63) return morton_sort<uint64_t>(c, c->kla.p, c->klb.p, c->va.p, c->vb.p, c->leaf_id.p, n);
    return morton_sort<uint32_t>(c, c->ka.p, c->kb.p, c->va.p, c->vb.p, c->leaf_id.p, n);
}

int build_lbvh(mmhrt_ctx_s* c, int bits) {
    uint64_t n = c->n;
    if (n == 0) return fail(c, MMHRT_E_STATE, "build: no spheres set");
    if (n > 0x7fffffffull) return fail(c, MMHRT_E_UNSUPPORTED, "build: more than 2^31-1 spheres");
    c->bits = bits;
    uint64_t m = n > 1 ? n - 1 : 1;
    int e = 0;
    if ((e = ensure(c, c->va, n, "sort values"))) return e;
    if ((e = ensure(c, c->vb, n, "sort values"))) return e;
    if ((e = ensure(c, c->leaf_sph, n, "leaf spheres"))) return e;
    if ((e = ensure(c, c->leaf_id, n, "leaf ids"))) return e;
    if ((e = ensure(c, c->nodes, m, "nodes"))) return e;
    if (bits == 63) {
        if ((e = ensure(c, c->kla, n, "keys"))) return e;
        if ((e = ensure(c, c->klb, n, "keys"))) return e;
    } else {
        if ((e = ensure(c, c->ka, n, "keys"))) return e;
        if ((e = ensure(c, c->kb, n, "keys"))) return e;
    }
    rec(c, 0);
    if ((e = reduce_scene(c))) return e;
    if (bits == 63)
        e = build_typed<uint64_t>(c, c->kla.p, c->klb.p);
    else
        e = build_typed<uint32_t>(c, c->ka.p, c->kb.p);
    if (e) return e;
    rec(c, 5);
    hipError_t he = hipGetLastError();
    if (he != hipSuccess) return hip_fail(c, he, "LBVH build launch");
    c->built = true;
    c->built_n = n;
    c->accel = 0;
    c->info_valid = false;
    return 0;
}

int refit_lbvh(mmhrt_ctx_s* c) {
    if (!c->built || c->built_n != c->n || c->accel != 0)
        return fail(c, MMHRT_E_STATE, "refit: no matching LBVH to refit");
    rec(c, 0);
    if (int e = reduce_scene(c)) return e;
    rec(c, 1);
    rec(c, 2);
    rec(c, 3);
    rec(c, 4);
    int e = c->bits == 63 ? hierarchy<uint64_t>(c, c->kla.p, c->leaf_id.p) : hierarchy<uint32_t>(c, c->ka.p, c->leaf_id.p);
    if (e) return e;
    rec(c, 5);
    hipError_t he = hipGetLastError();
    if (he != hipSuccess) return hip_fail(c, he, "LBVH refit launch");
    c->info_valid = false;
    return 0;
}

}  // namespace mmhrt
