// lbvh.hip — linear BVH built on the GPU, replacing optixAccelBuild
// (plugins/optix_hpg/src/optix/SphereGeometry.cpp:197-214) and the sphere bounds program
// (plugins/optix_hpg/cuda_resources/sphere.cuh:133-147, fused into the Morton pass here).
//
// Steps, each one kernel (or a few) on the context stream:
//   1. scene reduction: centre bounds, sphere bounds, |coord|max, Morton scale
//   2. Morton codes, 30-bit (10 per axis) or 63-bit (21 per axis), over one uniform scale
//   3. LSD radix sort of (key, particle index), 8-bit digits, stable: a tile's keys are ranked
//      with wavefront ballots (8 ballots give each lane the mask of same-digit lanes)
//   4. gather spheres into Morton order (leaf k), keep leaf -> particle index
//   5. Karras (2012) hierarchy: one thread per internal node, index tie-break for equal keys
//   6. bottom-up boxes: one thread per leaf climbs; the second arrival at a node (atomic
//      counter) writes the node's box into its parent and the subtree height
#include <algorithm>

#include "ctx.hpp"

namespace mmhrt {

namespace {

constexpr int RB = 256;      // threads per block
constexpr int RI = 16;       // keys per thread per tile
constexpr int RT = RB * RI;  // keys per tile

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

// Mask of the lanes of this wave whose (valid) digit equals ours.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        uint64_t bm = __ballot((d >> b) & 1u);
        m &= ((d >> b) & 1u) ? bm : ~bm;
    }
    return m;
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
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        float4 s = sph[i];
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

template <typename K>
__global__ __launch_bounds__(256) void k_morton(const float4* __restrict__ sph, uint64_t n,
                                                const SceneInfo* __restrict__ info, K* __restrict__ keys,
                                                uint32_t* __restrict__ vals) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 s = sph[i];
    float sc = info->scale;
    constexpr uint32_t maxq = sizeof(K) == 8 ? 2097151u : 1023u;
    float f[3] = {(s.x - info->clo[0]) * sc, (s.y - info->clo[1]) * sc, (s.z - info->clo[2]) * sc};
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        uint32_t u = (uint32_t)f[a];
        q[a] = u > maxq ? maxq : u;
    }
    if constexpr (sizeof(K) == 8)
        keys[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    else
        keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    vals[i] = (uint32_t)i;
}

// ---------------------------------------------------------------- radix sort
template <typename K>
__global__ __launch_bounds__(RB) void k_radix_hist(const K* __restrict__ keys, uint64_t n, int shift,
                                                   uint32_t* __restrict__ hist, uint32_t tiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * RT;
    for (int r = 0; r < RI; ++r) {
        uint64_t i = base + (uint64_t)r * RB + threadIdx.x;
        bool valid = i < n;
        uint32_t d = valid ? (uint32_t)((keys[i] >> shift) & 255u) : 0u;
        uint64_t peers = match_digit(d, valid);
        if (valid && (__ffsll((unsigned long long)peers) - 1) == (int)__lane_id())
            atomicAdd(&h[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one tile of RT keys for one 8-bit digit. The tile is first ranked into
// digit-sorted order in LDS (tile-local digit starts from the tile's histogram, wave-ballot ranks
// inside each round), then written out in digit-contiguous runs so that the global writes of
// consecutive threads go to consecutive addresses.
template <typename K>
__global__ __launch_bounds__(RB) void k_radix_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                      K* __restrict__ kout, uint32_t* __restrict__ vout, uint64_t n,
                                                      int shift, const uint32_t* __restrict__ hist,
                                                      const uint32_t* __restrict__ offs, uint32_t tiles) {
    __shared__ K s_key[RT];
    __shared__ uint32_t s_val[RT];
    __shared__ uint32_t running[256];        // next tile-local slot per digit
    __shared__ uint32_t goff[256];           // global offset of the digit's run minus its local start
    __shared__ uint32_t wcnt[RB / 64][256];  // per-wave digit counts of the current round
    __shared__ uint32_t wpre[RB / 64][256];  // slot of the first key of (wave, digit) in this round
    __shared__ uint32_t wsum[RB / 64];
    const int wave = threadIdx.x >> 6;
    const uint32_t lane = __lane_id();
    // tile-local digit starts: exclusive scan of this tile's digit counts (thread = digit)
    uint32_t cnt = hist[(size_t)threadIdx.x * tiles + blockIdx.x];
    uint32_t inc = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= (uint32_t)o) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    for (int w = 0; w < RB / 64; ++w) wcnt[w][threadIdx.x] = 0;
    __syncthreads();
    uint32_t wb = 0;
    for (int w = 0; w < wave; ++w) wb += wsum[w];
    const uint32_t lstart = wb + inc - cnt;
    running[threadIdx.x] = lstart;
    goff[threadIdx.x] = offs[(size_t)threadIdx.x * tiles + blockIdx.x] - lstart;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * RT;
    for (int r = 0; r < RI; ++r) {
        uint64_t i = base + (uint64_t)r * RB + threadIdx.x;
        bool valid = i < n;
        K key = valid ? kin[i] : (K)0;
        uint32_t val = valid ? vin[i] : 0u;
        uint32_t d = (uint32_t)((key >> shift) & 255u);
        uint64_t peers = match_digit(d, valid);
        uint32_t rank = (uint32_t)__popcll(peers & lanemask_lt());
        if (valid && (__ffsll((unsigned long long)peers) - 1) == (int)lane) wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t acc = running[threadIdx.x];
#pragma unroll
            for (int w = 0; w < RB / 64; ++w) {
                uint32_t c = wcnt[w][threadIdx.x];
                wpre[w][threadIdx.x] = acc;
                acc += c;
                wcnt[w][threadIdx.x] = 0;
            }
            running[threadIdx.x] = acc;
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = wpre[wave][d] + rank;
            s_key[pos] = key;
            s_val[pos] = val;
        }
        __syncthreads();
    }
    const uint32_t nvalid = (uint32_t)((n - base) < (uint64_t)RT ? (n - base) : (uint64_t)RT);
    for (uint32_t i = threadIdx.x; i < nvalid; i += RB) {
        K key = s_key[i];
        uint32_t d = (uint32_t)((key >> shift) & 255u);
        uint32_t g = goff[d] + i;
        kout[g] = key;
        vout[g] = s_val[i];
    }
}

// Exclusive scan of a u32 array: per-chunk sums, a one-block scan of the sums, then chunks.
constexpr int SB = 256, SI = 16, ST = SB * SI;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t* total) {
    __shared__ uint32_t wsum[SB / 64];
    uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    uint32_t inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= (uint32_t)o) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t wbase = 0, all = 0;
    for (int w = 0; w < SB / 64; ++w) {
        uint32_t s = wsum[w];
        if (w < (int)wave) wbase += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return wbase + inc - x;
}

__global__ __launch_bounds__(SB) void k_scan_sums(const uint32_t* __restrict__ in, uint64_t len,
                                                  uint32_t* __restrict__ sums) {
    uint64_t base = (uint64_t)blockIdx.x * ST;
    uint32_t s = 0;
    for (int r = 0; r < SI; ++r) {
        uint64_t i = base + (uint64_t)r * SB + threadIdx.x;
        if (i < len) s += in[i];
    }
    uint32_t tot;
    (void)block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SB) void k_scan_top(uint32_t* __restrict__ sums, uint32_t nsums) {
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nsums; base += SB) {
        uint32_t i = base + threadIdx.x;
        uint32_t x = i < nsums ? sums[i] : 0u;
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(x, &tot);
        if (i < nsums) sums[i] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(SB) void k_scan_chunks(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                    uint64_t len, const uint32_t* __restrict__ sums) {
    uint64_t base = (uint64_t)blockIdx.x * ST;
    uint32_t carry = sums[blockIdx.x];
    for (int r = 0; r < SI; ++r) {
        uint64_t i = base + (uint64_t)r * SB + threadIdx.x;
        uint32_t x = i < len ? in[i] : 0u;
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(x, &tot);
        if (i < len) out[i] = carry + ex;
        carry += tot;
    }
}

// ---------------------------------------------------------------- gather / hierarchy / boxes
template <typename K>
__global__ __launch_bounds__(256) void k_gather(const K* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                const float4* __restrict__ sph, float4* __restrict__ leaf_sph,
                                                uint32_t* __restrict__ leaf_id, uint64_t* __restrict__ keys64,
                                                uint64_t n) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t v = vals[k];
    leaf_sph[k] = sph[v];
    leaf_id[k] = v;
    keys64[k] = (uint64_t)keys[k];
}

template <typename K>
__device__ __forceinline__ int kdelta(const K* __restrict__ keys, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    K a = keys[i], b = keys[j];
    if (a == b) return (int)(8 * sizeof(K)) + __clz((uint32_t)i ^ (uint32_t)j);
    if constexpr (sizeof(K) == 8)
        return __clzll((long long)(a ^ b));
    else
        return __clz((int)(a ^ b));
}

template <typename K>
__global__ __launch_bounds__(256) void k_karras(const K* __restrict__ keys, int64_t n, Node* __restrict__ nodes,
                                                int32_t* __restrict__ parent) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t m = n - 1;
    if (i >= m) return;
    if (i == 0) parent[0] = -1;  // the root is nobody's child
    int dir = (kdelta(keys, n, i, i + 1) - kdelta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = kdelta(keys, n, i, i - dir);
    int64_t lmax = 2;
    while (kdelta(keys, n, i, i + lmax * dir) > dmin) lmax <<= 1;
    int64_t l = 0;
    for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
        if (kdelta(keys, n, i, i + (l + t) * dir) > dmin) l += t;
    int64_t j = i + l * dir;
    int dnode = kdelta(keys, n, i, j);
    int64_t s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (kdelta(keys, n, i, i + (s + t) * dir) > dnode) s += t;
    } while (t > 1);
    int64_t gamma = i + s * dir + (dir < 0 ? -1 : 0);
    int64_t first = i < j ? i : j, last = i < j ? j : i;
    int32_t left = (first == gamma) ? ~(int32_t)gamma : (int32_t)gamma;
    int32_t right = (last == gamma + 1) ? ~(int32_t)(gamma + 1) : (int32_t)(gamma + 1);
    nodes[i].d = make_int4(left, right, -1, 0);
    parent[left < 0 ? m + (int64_t)(~left) : (int64_t)left] = (int32_t)i;
    parent[right < 0 ? m + (int64_t)(~right) : (int64_t)right] = (int32_t)i;
}

__device__ __forceinline__ void store_child_box(Node* nd, int slot, const float b[6]) {
    float* f = reinterpret_cast<float*>(nd);
    int o = slot * 6;
#pragma unroll
    for (int k = 0; k < 6; ++k) f[o + k] = b[k];
}

// Bottom-up boxes without cross-workgroup fences. A node's box is final once both of its child
// slots are written; a device-scope fence per climb step (the classic single-kernel scheme)
// would write back L2 on every step, so completion is propagated in rounds instead:
//   init:   every internal node writes the boxes of its leaf children into its own slots and
//           counts them in flags[i]; nodes with two leaf children form the first worklist
//   rounds: each worklist node unions its two slots, writes the result into its parent's slot
//           and bumps the parent's counter; the second arrival appends the parent to the next
//           worklist. Kernel boundaries make the slots of earlier rounds visible.
//   finish: once a worklist is small, one workgroup runs the remaining rounds itself with
//           barriers between them (workgroup scope, no cache writeback).
constexpr int BOX_ROUNDS = 48;      // kernel rounds before the single-workgroup finish
constexpr uint32_t BOX_SMALL = 2048;  // worklist size handed to the single workgroup

__device__ __forceinline__ void leaf_box(const float4 s, float b[6]) {
    b[0] = s.x - s.w; b[1] = s.y - s.w; b[2] = s.z - s.w;
    b[3] = s.x + s.w; b[4] = s.y + s.w; b[5] = s.z + s.w;
}

// Wave-aggregated append: one atomic per wavefront. All lanes of the wave must call it.
__device__ __forceinline__ void wave_append(bool want, uint32_t v, uint32_t* list, uint32_t* cnt) {
    uint64_t m = __ballot(want);
    if (m == 0) return;
    int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(cnt, (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    if (want) list[base + (uint32_t)__popcll(m & lanemask_lt())] = v;
}

// Block-aggregated worklist append: a block gathers up to BOX_TILE entries in LDS, reserves
// space with one global atomic and writes them out contiguously.
constexpr int BOX_ITEMS = 16;
constexpr int BOX_TILE = 256 * BOX_ITEMS;

struct BoxAppend {
    uint32_t* s_list;
    uint32_t* s_n;
    uint32_t* s_base;
    __device__ __forceinline__ void begin() {
        if (threadIdx.x == 0) *s_n = 0;
        __syncthreads();
    }
    __device__ __forceinline__ void push(uint32_t v) { s_list[atomicAdd(s_n, 1u)] = v; }
    __device__ __forceinline__ void flush(uint32_t* list, uint32_t* cnt) {
        __syncthreads();
        uint32_t n = *s_n;
        if (threadIdx.x == 0 && n) *s_base = atomicAdd(cnt, n);
        __syncthreads();
        uint32_t base = *s_base;
        for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) list[base + k] = s_list[k];
        __syncthreads();
    }
};

__global__ __launch_bounds__(256) void k_box_init(const float4* __restrict__ leaf_sph, int64_t m, Node* nodes,
                                                  uint32_t* __restrict__ flags, uint32_t* __restrict__ list,
                                                  uint32_t* cnt) {
    __shared__ uint32_t s_list[BOX_TILE];
    __shared__ uint32_t s_n, s_base;
    BoxAppend ap{s_list, &s_n, &s_base};
    for (int64_t tile = (int64_t)blockIdx.x * BOX_TILE; tile < m; tile += (int64_t)gridDim.x * BOX_TILE) {
        ap.begin();
        for (int j = 0; j < BOX_ITEMS; ++j) {
            int64_t i = tile + (int64_t)j * 256 + threadIdx.x;
            if (i >= m) break;
            int4 d = nodes[i].d;
            uint32_t nl = 0;
            float b[6];
            if (d.x < 0) {
                leaf_box(leaf_sph[~d.x], b);
                store_child_box(nodes + i, 0, b);
                ++nl;
            }
            if (d.y < 0) {
                leaf_box(leaf_sph[~d.y], b);
                store_child_box(nodes + i, 1, b);
                ++nl;
            }
            flags[i] = nl;
            if (nl == 2) ap.push((uint32_t)i);
        }
        ap.flush(list, cnt + 1);
    }
}

// Union the two slots of node p, record its height, hand the box to the parent.
// Returns the parent when this arrival completes it, else -1.
template <bool VOL>
__device__ __forceinline__ int32_t box_step(Node* nodes, const int32_t* __restrict__ parent, uint32_t* flags,
                                            int32_t p, SceneInfo* info) {
    float c[12];
    int4 d;
    if constexpr (VOL) {
        const volatile float* f = reinterpret_cast<const volatile float*>(nodes + p);
#pragma unroll
        for (int q = 0; q < 12; ++q) c[q] = f[q];
        const volatile int* di = reinterpret_cast<const volatile int*>(&nodes[p].d);
        d = make_int4(di[0], di[1], di[2], di[3]);
    } else {
        const float* f = reinterpret_cast<const float*>(nodes + p);
#pragma unroll
        for (int q = 0; q < 12; ++q) c[q] = f[q];
        d = nodes[p].d;
    }
    float b[6];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        b[a] = fminf(c[a], c[6 + a]);
        b[3 + a] = fmaxf(c[3 + a], c[9 + a]);
    }
    int h0 = 0, h1 = 0;
    if (d.x >= 0) h0 = VOL ? ((const volatile int*)&nodes[d.x].d)[3] : nodes[d.x].d.w;
    if (d.y >= 0) h1 = VOL ? ((const volatile int*)&nodes[d.y].d)[3] : nodes[d.y].d.w;
    int h = 1 + (h0 > h1 ? h0 : h1);
    reinterpret_cast<int*>(&nodes[p].d)[3] = h;
    int32_t pp = parent[p];
    if (pp < 0) {
        info->height = (uint32_t)h;
        return -1;
    }
    int first = VOL ? ((const volatile int*)&nodes[pp].d)[0] : nodes[pp].d.x;
    store_child_box(nodes + pp, first == p ? 0 : 1, b);
    uint32_t old = atomicAdd(flags + pp, 1u);
    return old == 1u ? pp : -1;
}

__device__ void box_finish(Node* nodes, const int32_t* __restrict__ parent, uint32_t* flags, uint32_t* la,
                           uint32_t* lb, uint32_t n_cur, uint32_t* cnt, SceneInfo* info) {
    __shared__ uint32_t s_out;
    uint32_t* lin = la;
    uint32_t* lout = lb;
    while (n_cur > 0) {
        if (threadIdx.x == 0) s_out = 0;
        __syncthreads();
        for (uint32_t base = 0; base < n_cur; base += blockDim.x) {
            uint32_t i = base + threadIdx.x;
            if (i < n_cur) {
                int32_t p = (int32_t)((const volatile uint32_t*)lin)[i];
                int32_t pp = box_step<true>(nodes, parent, flags, p, info);
                if (pp >= 0) lout[atomicAdd(&s_out, 1u)] = (uint32_t)pp;
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

// Round r: worklist r lives in list[r & 1], its size in cnt[1 + r]; cnt[0] marks completion.
__global__ __launch_bounds__(256) void k_box_round(Node* nodes, const int32_t* __restrict__ parent, uint32_t* flags,
                                                   uint32_t* l0, uint32_t* l1, uint32_t* cnt, int r, SceneInfo* info) {
    __shared__ uint32_t s_list[BOX_TILE];
    __shared__ uint32_t s_n, s_base;
    if (cnt[0]) return;
    uint32_t n_in = cnt[1 + r];
    uint32_t* lin = (r & 1) ? l1 : l0;
    uint32_t* lout = (r & 1) ? l0 : l1;
    if (n_in <= BOX_SMALL) {
        if (blockIdx.x == 0) box_finish(nodes, parent, flags, lin, lout, n_in, cnt, info);
        return;
    }
    BoxAppend ap{s_list, &s_n, &s_base};
    for (uint32_t tile = blockIdx.x * BOX_TILE; tile < n_in; tile += gridDim.x * BOX_TILE) {
        ap.begin();
        for (int j = 0; j < BOX_ITEMS; ++j) {
            uint32_t i = tile + j * 256 + threadIdx.x;
            if (i >= n_in) break;
            int32_t pp = box_step<false>(nodes, parent, flags, (int32_t)lin[i], info);
            if (pp >= 0) ap.push((uint32_t)pp);
        }
        ap.flush(lout, cnt + 2 + r);
    }
}

// Whatever is left after BOX_ROUNDS kernel rounds: one workgroup finishes it.
__global__ __launch_bounds__(256) void k_box_tail(Node* nodes, const int32_t* __restrict__ parent, uint32_t* flags,
                                                  uint32_t* l0, uint32_t* l1, uint32_t* cnt, SceneInfo* info) {
    if (cnt[0]) return;
    int r = BOX_ROUNDS;
    uint32_t* lin = (r & 1) ? l1 : l0;
    uint32_t* lout = (r & 1) ? l0 : l1;
    box_finish(nodes, parent, flags, lin, lout, cnt[1 + r], cnt, info);
}

// Single-sphere scene: one node whose two children are the same leaf.
__global__ void k_single(const float4* __restrict__ leaf_sph, Node* nodes, SceneInfo* info) {
    float4 s = leaf_sph[0];
    float b[6] = {s.x - s.w, s.y - s.w, s.z - s.w, s.x + s.w, s.y + s.w, s.z + s.w};
    store_child_box(nodes, 0, b);
    store_child_box(nodes, 1, b);
    nodes[0].d = make_int4(~0, ~0, -1, 1);
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

template <typename K>
int sort_pairs(mmhrt_ctx_s* c, K* ka, K* kb, uint32_t* va, uint32_t* vb, uint64_t n, int passes) {
    uint32_t tiles = (uint32_t)((n + RT - 1) / RT);
    uint64_t hlen = (uint64_t)tiles * 256;
    uint32_t nchunks = (uint32_t)((hlen + ST - 1) / ST);
    if (int e = ensure(c, c->hist, hlen * 2, "radix histogram")) return e;
    if (int e = ensure(c, c->scan_tmp, nchunks, "scan sums")) return e;
    uint32_t* hist = c->hist.p;
    uint32_t* offs = c->hist.p + hlen;
    K* src = ka;
    K* dst = kb;
    uint32_t* vs = va;
    uint32_t* vd = vb;
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        hipLaunchKernelGGL(k_radix_hist<K>, dim3(tiles), dim3(RB), 0, c->stream, src, n, shift, hist, tiles);
        hipLaunchKernelGGL(k_scan_sums, dim3(nchunks), dim3(SB), 0, c->stream, hist, hlen, c->scan_tmp.p);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SB), 0, c->stream, c->scan_tmp.p, nchunks);
        hipLaunchKernelGGL(k_scan_chunks, dim3(nchunks), dim3(SB), 0, c->stream, hist, offs, hlen, c->scan_tmp.p);
        hipLaunchKernelGGL(k_radix_scatter<K>, dim3(tiles), dim3(RB), 0, c->stream, src, vs, dst, vd, n, shift, hist,
                           offs, tiles);
        K* tk = src;
        src = dst;
        dst = tk;
        uint32_t* tv = vs;
        vs = vd;
        vd = tv;
    }
    return 0;  // passes is even: result is back in (ka, va)
}

template <typename K>
int build_typed(mmhrt_ctx_s* c, K* ka, K* kb) {
    uint64_t n = c->n;
    hipStream_t s = c->stream;
    rec(c, 1);
    hipLaunchKernelGGL(k_morton<K>, dim3(blocks_for(n, 256)), dim3(256), 0, s, cur_spheres(c), n, c->info.p, ka, c->va.p);
    rec(c, 2);
    int passes = sizeof(K) == 8 ? 8 : 4;
    if (int e = sort_pairs<K>(c, ka, kb, c->va.p, c->vb.p, n, passes)) return e;
    rec(c, 3);
    hipLaunchKernelGGL(k_gather<K>, dim3(blocks_for(n, 256)), dim3(256), 0, s, ka, c->va.p, cur_spheres(c), c->leaf_sph.p,
                       c->leaf_id.p, c->keys64.p, n);
    if (n > 1) {
        hipLaunchKernelGGL(k_karras<K>, dim3(blocks_for(n - 1, 256)), dim3(256), 0, s, ka, (int64_t)n, c->nodes.p,
                           c->parent.p);
    }
    rec(c, 4);
    return 0;
}

}  // namespace

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

static int boxes(mmhrt_ctx_s* c) {
    uint64_t n = c->n;
    hipStream_t s = c->stream;
    if (n == 1) {
        hipLaunchKernelGGL(k_single, dim3(1), dim3(1), 0, s, c->leaf_sph.p, c->nodes.p, c->info.p);
        return 0;
    }
    uint64_t m = n - 1;
    if (int e = ensure(c, c->box_cnt, BOX_ROUNDS + 4, "box round counters")) return e;
    launch_fill_u32(s, c->box_cnt.p, 0u, BOX_ROUNDS + 4);
    unsigned ig = (unsigned)std::min<uint64_t>((m + BOX_TILE - 1) / BOX_TILE, 8192);
    hipLaunchKernelGGL(k_box_init, dim3(ig), dim3(256), 0, s, c->leaf_sph.p, (int64_t)m, c->nodes.p,
                       c->flags.p, c->va.p, c->box_cnt.p);
    for (int r = 0; r < BOX_ROUNDS; ++r)
        hipLaunchKernelGGL(k_box_round, dim3(1024), dim3(256), 0, s, c->nodes.p, c->parent.p, c->flags.p, c->va.p,
                           c->vb.p, c->box_cnt.p, r, c->info.p);
    hipLaunchKernelGGL(k_box_tail, dim3(1), dim3(256), 0, s, c->nodes.p, c->parent.p, c->flags.p, c->va.p, c->vb.p,
                       c->box_cnt.p, c->info.p);
    return 0;
}

int build_lbvh(mmhrt_ctx_s* c, int bits) {
    uint64_t n = c->n;
    if (n == 0) return fail(c, MMHRT_E_STATE, "build: no spheres set");
    if (n > 0x7fffffffull) return fail(c, MMHRT_E_UNSUPPORTED, "build: more than 2^31-1 spheres");
    c->bits = bits;
    uint64_t m = n > 1 ? n - 1 : 1;
    int e = 0;
    if ((e = ensure(c, c->keys64, n, "keys"))) return e;
    if ((e = ensure(c, c->va, n, "sort values"))) return e;
    if ((e = ensure(c, c->vb, n, "sort values"))) return e;
    if ((e = ensure(c, c->leaf_sph, n, "leaf spheres"))) return e;
    if ((e = ensure(c, c->leaf_id, n, "leaf ids"))) return e;
    if ((e = ensure(c, c->nodes, m, "nodes"))) return e;
    if ((e = ensure(c, c->parent, 2 * n, "parents"))) return e;
    if ((e = ensure(c, c->flags, m, "flags"))) return e;
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
    if ((e = boxes(c))) return e;
    rec(c, 5);
    hipError_t he = hipGetLastError();
    if (he != hipSuccess) return hip_fail(c, he, "LBVH build launch");
    c->built = true;
    c->built_n = n;
    c->info_valid = false;
    return 0;
}

int refit_lbvh(mmhrt_ctx_s* c) {
    if (!c->built || c->built_n != c->n) return fail(c, MMHRT_E_STATE, "refit: no matching LBVH to refit");
    uint64_t n = c->n;
    rec(c, 0);
    if (int e = reduce_scene(c)) return e;
    rec(c, 1);
    rec(c, 2);
    rec(c, 3);
    hipLaunchKernelGGL(k_regather, dim3(blocks_for(n, 256)), dim3(256), 0, c->stream, cur_spheres(c), c->leaf_id.p,
                       c->leaf_sph.p, n);
    rec(c, 4);
    if (int e = boxes(c)) return e;
    rec(c, 5);
    hipError_t he = hipGetLastError();
    if (he != hipSuccess) return hip_fail(c, he, "LBVH refit launch");
    c->info_valid = false;
    return 0;
}

}  // namespace mmhrt
