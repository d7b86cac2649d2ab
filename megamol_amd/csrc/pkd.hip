// pkd.hip — treelets of P-k-d trees: the compact acceleration structure for very large scenes
// (plugins/optix_owl/src/TreeletsCreate.h:117-166 prePartition_inPlace, PKDCreate.h:86-160 recBuild,
// used by TreeletsRenderer.cpp:95-110).
//
// The spheres are sorted by Morton code (the first half of an LBVH build) and cut into treelets
// of M consecutive spheres (M a power of two; the last treelet is shorter). The reference cuts at
// spatial midpoints until a part holds at most `threshold` particles; a run of the Morton order is
// a cell of the same kind of midpoint subdivision, and a fixed M lets the traversal find a
// treelet by a shift. Each treelet becomes a left-balanced k-d tree over its own spheres in heap
// order (node i of treelet t at position t*M + i, children 2i+1 and 2i+2): node i's split axis
// is the widest extent of its cell (ties to the lower axis), its sphere the one of rank
// |left subtree| among its subtree's spheres ordered by the coordinate on that axis (ties: the
// lower Morton position), and the children's cells are the node's cell cut at that coordinate.
// The root cell is the treelet's sphere bounds. A complete binary tree of boxes over the treelets
// (64-byte nodes like the LBVH's, heap order, treelet t at leaf slot t) sits on top.
//
// Memory beyond the spheres in treelet order: one byte per sphere (the split axis), 28 bytes per
// treelet and 64 bytes per top node, against 64 bytes per stored LBVH node.
//
// One 256-thread workgroup builds one treelet in LDS, level by level: the spheres never move;
// 64-bit keys (segment start, coordinate on the segment's axis, local index) are bitonic-sorted
// once per level, which orders every segment on its own axis in place, and the element of rank
// |left subtree| in each segment is fixed at its in-order position. A key whose segment start is
// its own position marks an element already fixed: the keys stay ordered by position, so it never
// moves.
#include <float.h>

#include <algorithm>

#include "ctx.hpp"

namespace mmhrt {
namespace {

constexpr float PKD_ABSENT = 3.0e38f;  // the box no ray reaches: an absent child of the top tree

__device__ __forceinline__ uint32_t ford(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// nodes of the subtree of heap node i in a left-balanced tree of m nodes
__device__ __forceinline__ uint32_t subtree_size(uint32_t i, uint32_t m) {
    uint32_t first = i, width = 1, cnt = 0;
    while (first < m) {
        uint32_t end = first + width;
        cnt += end <= m ? width : m - first;
        first = 2 * first + 1;
        width *= 2;
    }
    return cnt;
}

__device__ __forceinline__ float coord(const float4& s, int d) { return d == 0 ? s.x : (d == 1 ? s.y : s.z); }

__device__ __forceinline__ int widest_axis(const float* b) {
    float e0 = b[3] - b[0], e1 = b[4] - b[1], e2 = b[5] - b[2];
    int d = 0;
    float e = e0;
    if (e1 > e) {
        d = 1;
        e = e1;
    }
    if (e2 > e) d = 2;
    return d;
}

template <int M>
__global__ __launch_bounds__(256) void k_pkd_treelet(const float4* __restrict__ sph, uint32_t* __restrict__ ids,
                                                     float4* __restrict__ out_sph, uint8_t* __restrict__ out_dim,
                                                     float* __restrict__ tl_box, float* __restrict__ tl_rmax,
                                                     uint64_t n) {
    constexpr int NT = M / 256;
    __shared__ float4 s_sph[M];
    __shared__ uint32_t s_id[M];
    __shared__ uint64_t s_key[M];
    __shared__ float s_box[M * 6];   // cell of heap node i
    __shared__ uint16_t s_start[M];  // in-order start of node i's segment
    __shared__ uint16_t s_pos[M];    // segment (heap node) holding position p; bit 15: p is fixed
    __shared__ uint16_t s_elem[M];   // node i's sphere (local index)
    __shared__ uint8_t s_dim[M];
    __shared__ float s_red[4][7];
    const uint32_t tid = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * M;
    const uint32_t m = (uint32_t)std::min<uint64_t>(M, n - base);
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    float rmax = 0.0f;
    uint32_t idv[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        const uint32_t j = tid + 256u * k;
        idv[k] = j < m ? ids[base + j] : 0u;
    }
#pragma unroll
    for (int k = 0; k < NT; ++k) {
        const uint32_t j = tid + 256u * k;
        if (j < m) {
            const float4 q = sph[idv[k]];
            s_sph[j] = q;
            s_id[j] = idv[k];
            lo[0] = fminf(lo[0], q.x - q.w);
            lo[1] = fminf(lo[1], q.y - q.w);
            lo[2] = fminf(lo[2], q.z - q.w);
            hi[0] = fmaxf(hi[0], q.x + q.w);
            hi[1] = fmaxf(hi[1], q.y + q.w);
            hi[2] = fmaxf(hi[2], q.z + q.w);
            rmax = fmaxf(rmax, q.w);
        }
        if (j < M) {
            s_key[j] = j;
            s_pos[j] = 0;
        }
    }
    // treelet bounds: min / max are exact in any order
    for (int o = 32; o > 0; o >>= 1) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], __shfl_xor(lo[a], o, 64));
            hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], o, 64));
        }
        rmax = fmaxf(rmax, __shfl_xor(rmax, o, 64));
    }
    const uint32_t wv = tid >> 6;
    if ((tid & 63) == 0) {
        for (int a = 0; a < 3; ++a) {
            s_red[wv][a] = lo[a];
            s_red[wv][3 + a] = hi[a];
        }
        s_red[wv][6] = rmax;
    }
    __syncthreads();
    if (tid == 0) {
        float b[6];
        for (int a = 0; a < 3; ++a) {
            b[a] = fminf(fminf(s_red[0][a], s_red[1][a]), fminf(s_red[2][a], s_red[3][a]));
            b[3 + a] = fmaxf(fmaxf(s_red[0][3 + a], s_red[1][3 + a]), fmaxf(s_red[2][3 + a], s_red[3][3 + a]));
        }
        const float r = fmaxf(fmaxf(s_red[0][6], s_red[1][6]), fmaxf(s_red[2][6], s_red[3][6]));
        for (int a = 0; a < 6; ++a) {
            s_box[a] = b[a];
            tl_box[(size_t)blockIdx.x * 6 + a] = b[a];
        }
        tl_rmax[blockIdx.x] = r;
        s_dim[0] = (uint8_t)widest_axis(b);
        s_start[0] = 0;
    }
    uint32_t mp = 1;
    while (mp < m) mp <<= 1;
    __syncthreads();
    for (uint32_t first = 0; first < m; first = 2 * first + 1) {
        const uint32_t last = std::min(2 * first, m - 1);  // the nodes of this level
        if (2 * first + 1 < m) {
            // keys: (segment start, coordinate on the segment's axis, local index); a fixed
            // element's segment is its own position
            for (uint32_t p = tid; p < mp; p += 256) {
                uint64_t key = ~0ull;
                if (p < m) {
                    const uint32_t j = (uint32_t)(s_key[p] & 0x3ffu);
                    const uint32_t code = s_pos[p];
                    if (code & 0x8000u) {
                        key = ((uint64_t)p << 42) | j;
                    } else {
                        const float cv = coord(s_sph[j], s_dim[code]);
                        key = ((uint64_t)s_start[code] << 42) | ((uint64_t)ford(cv) << 10) | j;
                    }
                }
                s_key[p] = key;
            }
            __syncthreads();
            for (uint32_t k = 2; k <= mp; k <<= 1)
                for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
                    for (uint32_t q = tid; q < mp / 2; q += 256) {
                        const uint32_t a = ((q & ~(jj - 1)) << 1) | (q & (jj - 1));
                        const uint32_t b = a | jj;
                        const uint64_t x = s_key[a], y = s_key[b];
                        if ((x > y) == ((a & k) == 0)) {
                            s_key[a] = y;
                            s_key[b] = x;
                        }
                    }
                    __syncthreads();
                }
        }
        // fix each segment's element of rank |left subtree|; the rest go to the children
        for (uint32_t p = tid; p < m; p += 256) {
            const uint32_t code = s_pos[p];
            if (code & 0x8000u) continue;
            const uint32_t i = code;
            const uint32_t L = 2 * i + 1 < m ? subtree_size(2 * i + 1, m) : 0u;
            const uint32_t med = s_start[i] + L;
            if (p == med) {
                s_pos[p] = (uint16_t)(i | 0x8000u);
                s_elem[i] = (uint16_t)(s_key[p] & 0x3ffu);
            } else {
                s_pos[p] = (uint16_t)(p < med ? 2 * i + 1 : 2 * i + 2);
            }
        }
        __syncthreads();
        // the children's cells, axes and segment starts
        for (uint32_t i = first + tid; i <= last; i += 256) {
            const int d = s_dim[i];
            const float cv = coord(s_sph[s_elem[i]], d);
            const uint32_t L = 2 * i + 1 < m ? subtree_size(2 * i + 1, m) : 0u;
            float b[6];
            for (int a = 0; a < 6; ++a) b[a] = s_box[6 * i + a];
            if (2 * i + 1 < m) {
                float c[6];
                for (int a = 0; a < 6; ++a) c[a] = b[a];
                c[3 + d] = cv;
                for (int a = 0; a < 6; ++a) s_box[6 * (2 * i + 1) + a] = c[a];
                s_dim[2 * i + 1] = (uint8_t)widest_axis(c);
                s_start[2 * i + 1] = s_start[i];
            }
            if (2 * i + 2 < m) {
                float c[6];
                for (int a = 0; a < 6; ++a) c[a] = b[a];
                c[d] = cv;
                for (int a = 0; a < 6; ++a) s_box[6 * (2 * i + 2) + a] = c[a];
                s_dim[2 * i + 2] = (uint8_t)widest_axis(c);
                s_start[2 * i + 2] = (uint16_t)(s_start[i] + L + 1);
            }
        }
        __syncthreads();
    }
    for (uint32_t i = tid; i < m; i += 256) {
        const uint32_t j = s_elem[i];
        out_sph[base + i] = s_sph[j];
        ids[base + i] = s_id[j];
        out_dim[base + i] = s_dim[i];
    }
}

// One level of the top tree (heap nodes first .. first + count - 1), bottom up: a child slot at or
// past L - 1 is treelet slot - (L - 1) (absent from T on); an internal child's box is the union of
// its present children.
__global__ __launch_bounds__(256) void k_pkd_top_level(Node* __restrict__ nodes, const float* __restrict__ tl_box,
                                                       uint32_t T, uint32_t L, uint32_t first, uint32_t count) {
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    if (q >= count) return;
    const uint32_t k = first + q;
    float bx[2][6];
    int32_t ref[2];
    for (int c = 0; c < 2; ++c) {
        const uint32_t ch = 2 * k + 1 + (uint32_t)c;
        bool present = false;
        if (ch >= L - 1) {
            const uint32_t t = ch - (L - 1);
            if (t < T) {
                present = true;
                for (int a = 0; a < 6; ++a) bx[c][a] = tl_box[(size_t)t * 6 + a];
                ref[c] = ~(int32_t)t;
            }
        } else {
            const Node cn = nodes[ch];
            const float* f = reinterpret_cast<const float*>(&cn);
            const bool p0 = cn.d.x != INT32_MIN, p1 = cn.d.y != INT32_MIN;
            if (p0 || p1) {
                present = true;
                for (int a = 0; a < 3; ++a) {
                    const float l0 = f[a], h0 = f[3 + a], l1 = f[6 + a], h1 = f[9 + a];
                    bx[c][a] = p0 && p1 ? fminf(l0, l1) : (p0 ? l0 : l1);
                    bx[c][3 + a] = p0 && p1 ? fmaxf(h0, h1) : (p0 ? h0 : h1);
                }
                ref[c] = (int32_t)ch;
            }
        }
        if (!present) {
            for (int a = 0; a < 6; ++a) bx[c][a] = PKD_ABSENT;
            ref[c] = INT32_MIN;
        }
    }
    Node o;
    o.a = make_float4(bx[0][0], bx[0][1], bx[0][2], bx[0][3]);
    o.b = make_float4(bx[0][4], bx[0][5], bx[1][0], bx[1][1]);
    o.c = make_float4(bx[1][2], bx[1][3], bx[1][4], bx[1][5]);
    o.d = make_int4(ref[0], ref[1], 0, 0);
    nodes[k] = o;
}

}  // namespace

int build_pkd(mmhrt_ctx_s* c, int bits) {
    const int64_t M = c->opt_treelet;
    if (M != 256 && M != 512 && M != 1024)
        return fail(c, MMHRT_E_INVALID, "build: option treelet must be 256, 512 or 1024");
    if (int e = sort_spheres(c, bits)) return e;
    const uint64_t n = c->n;
    const uint32_t T = (uint32_t)((n + M - 1) / M);
    uint32_t L = 2;
    while (L < T) L *= 2;
    int e = 0;
    if ((e = ensure(c, c->leaf_sph, n, "leaf spheres"))) return e;
    if ((e = ensure(c, c->pkd_dim, n, "split axes"))) return e;
    if ((e = ensure(c, c->tl_box, (size_t)T * 6, "treelet boxes"))) return e;
    if ((e = ensure(c, c->tl_rmax, T, "treelet radii"))) return e;
    if ((e = ensure(c, c->nodes, L - 1, "top nodes"))) return e;
    rec(c, 3);
    rec(c, 4);
    const float4* sph = cur_spheres(c);
    if (M == 256)
        hipLaunchKernelGGL(k_pkd_treelet<256>, dim3(T), dim3(256), 0, c->stream, sph, c->leaf_id.p, c->leaf_sph.p,
                           c->pkd_dim.p, c->tl_box.p, c->tl_rmax.p, n);
    else if (M == 512)
        hipLaunchKernelGGL(k_pkd_treelet<512>, dim3(T), dim3(256), 0, c->stream, sph, c->leaf_id.p, c->leaf_sph.p,
                           c->pkd_dim.p, c->tl_box.p, c->tl_rmax.p, n);
    else
        hipLaunchKernelGGL(k_pkd_treelet<1024>, dim3(T), dim3(256), 0, c->stream, sph, c->leaf_id.p, c->leaf_sph.p,
                           c->pkd_dim.p, c->tl_box.p, c->tl_rmax.p, n);
    for (uint32_t first = L / 2 - 1, count = L / 2;; first = (first - 1) / 2, count /= 2) {
        hipLaunchKernelGGL(k_pkd_top_level, dim3((count + 255) / 256), dim3(256), 0, c->stream, c->nodes.p,
                           c->tl_box.p, T, L, first, count);
        if (first == 0) break;
    }
    rec(c, 5);
    hipError_t he = hipGetLastError();
    if (he != hipSuccess) return hip_fail(c, he, "P-k-d build launch");
    c->pkd_M = (uint32_t)M;
    c->pkd_T = T;
    c->pkd_L = L;
    c->built = true;
    c->built_n = n;
    c->accel = 1;
    c->bvh_pairs = false;
    c->info_valid = false;
    return 0;
}

}  // namespace mmhrt
