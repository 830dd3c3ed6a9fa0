// Face topology of a prism / tet mesh on the GPU: the deduplicated triangle records with owner,
// neighbour, cube-face codes and FP64 geometry.
//
// Semantics of FaceExtractor.extract_faces_mixed
// (src/autoflowcfd/grid/mesh_gen/extraction/face_extractor.py:296-376) and the face connectivity
// (src/autoflowcfd/grid/connectivity/face_connectivity.py:53-231):
//   * every prism contributes 8 triangles (bottom, top, and its three quads split on the
//     global-node-id diagonal rule of face_extraction_kernels.py:247-322), every tet its 4 faces;
//   * records are the distinct sorted node triples, in ascending lexicographic order of the triple;
//     the first occurrence (in cell order, then local triangle order) is the owner, the second the
//     neighbour; a triple shared by more than two cells is an error;
//   * centre = (p0 + p1 + p2) / 3 and area vector = (p1 - p0) x (p2 - p0) / 2 over the sorted triple,
//     flipped to point away from the owner's vertex centroid (face_geometry_finalize.py:92-160);
//   * the cube-face code of each side is the cell face whose corners hold the triangle's nodes
//     (exactly one must match).
//
// GPU form: one thread per triangle builds its sorted triple and a radix key; a stable device
// radix sort (hipCUB over rocPRIM) orders the triangles by key with their index as value, so
// equal triples keep cell order; a run-start flag and an inclusive scan number the records; one
// thread per run writes the record.  Node ids below 2^21 pack the triple into one 63-bit key
// (one sort); larger meshes sort twice (the low pair, then stably the first id).  The FP64
// geometry is evaluated without contraction into fused multiply-adds, in the operation order of
// the NumPy formulation, so every value is the correctly rounded result of the same operations.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <string>

#include "../../include/afr.h"

void afr_set_error(const std::string& msg);

namespace {

constexpr int kThreads = 256;

// local vertex indices of the real cube faces (codes 0..5 = a-, a+, b-, b+, c-, c+; -1 unused),
// the tables of src/autoflowcfd/grid/curved_mapping/curved_mapping.py:254-275
__constant__ int kPrismFace[6][4] = {{0, 2, 5, 3}, {1, 2, 5, 4}, {0, 1, 4, 3}, {-1, -1, -1, -1},
                                     {0, 1, 2, -1}, {3, 4, 5, -1}};
__constant__ int kTetFace[6][4] = {{0, 2, 3, -1}, {1, 2, 3, -1}, {0, 1, 3, -1}, {-1, -1, -1, -1},
                                   {0, 1, 2, -1}, {-1, -1, -1, -1}};

struct TopoArgs {
    int64_t n_prism, n_tet, n_tri;
    const int32_t* prism_conn;
    const int32_t* tet_conn;
};

__device__ __forceinline__ int64_t cell_of(const TopoArgs& a, int64_t t, int* j) {
    if (t < 8 * a.n_prism) {
        *j = (int)(t & 7);
        return t >> 3;
    }
    const int64_t u = t - 8 * a.n_prism;
    *j = (int)(u & 3);
    return a.n_prism + (u >> 2);
}

// triangle t's three node ids in the orientation of its cell's triangulation
__device__ void tri_nodes(const TopoArgs& a, int64_t t, int32_t v[3]) {
    int j;
    const int64_t c = cell_of(a, t, &j);
    if (c < a.n_prism) {
        const int32_t* cn = a.prism_conn + c * 6;
        int32_t b[3] = {cn[0], cn[1], cn[2]}, w[3] = {cn[3], cn[4], cn[5]};
        // order the bottom triangle by global id, the top with it (ids are distinct)
        int o[3] = {0, 1, 2};
        if (b[o[1]] < b[o[0]]) { int x = o[0]; o[0] = o[1]; o[1] = x; }
        if (b[o[2]] < b[o[1]]) { int x = o[1]; o[1] = o[2]; o[2] = x; }
        if (b[o[1]] < b[o[0]]) { int x = o[0]; o[0] = o[1]; o[1] = x; }
        const int32_t v0 = b[o[0]], v1 = b[o[1]], v2 = b[o[2]];
        const int32_t w0 = w[o[0]], w1 = w[o[1]], w2 = w[o[2]];
        switch (j) {
            case 0: v[0] = v0; v[1] = v1; v[2] = v2; break;
            case 1: v[0] = w0; v[1] = w1; v[2] = w2; break;
            case 2: v[0] = v0; v[1] = v1; v[2] = w1; break;
            case 3: v[0] = v0; v[1] = w0; v[2] = w1; break;
            case 4: v[0] = v1; v[1] = v2; v[2] = w2; break;
            case 5: v[0] = v1; v[1] = w1; v[2] = w2; break;
            case 6: v[0] = v0; v[1] = v2; v[2] = w2; break;
            default: v[0] = v0; v[1] = w0; v[2] = w2; break;
        }
    } else {
        const int32_t* cn = a.tet_conn + (c - a.n_prism) * 4;
        switch (j) {
            case 0: v[0] = cn[0]; v[1] = cn[1]; v[2] = cn[2]; break;
            case 1: v[0] = cn[0]; v[1] = cn[1]; v[2] = cn[3]; break;
            case 2: v[0] = cn[0]; v[1] = cn[2]; v[2] = cn[3]; break;
            default: v[0] = cn[1]; v[1] = cn[2]; v[2] = cn[3]; break;
        }
    }
}

__device__ __forceinline__ void sort3(int32_t v[3]) {
    if (v[1] < v[0]) { int32_t x = v[0]; v[0] = v[1]; v[1] = x; }
    if (v[2] < v[1]) { int32_t x = v[1]; v[1] = v[2]; v[2] = x; }
    if (v[1] < v[0]) { int32_t x = v[0]; v[0] = v[1]; v[1] = x; }
}

// a zero-area triangle of a collapsed-edge wedge (two equal node ids) is stored as this triple: it
// sorts after every real triangle in both sort forms and is never a record
constexpr int32_t kNoTri = 0x7fffffff;

__device__ __forceinline__ bool no_tri(const int32_t* tri3, int32_t t) { return tri3[(int64_t)t * 3] == kNoTri; }

// pass 0: sorted triples, keys (one-pass packed key, or the low pair of a two-pass sort)
__global__ void __launch_bounds__(kThreads) k_tri_keys(TopoArgs a, int32_t* __restrict__ tri3,
                                                       uint64_t* __restrict__ key, int32_t* __restrict__ val,
                                                       int packed) {
    for (int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x; t < a.n_tri; t += (int64_t)gridDim.x * kThreads) {
        int32_t v[3];
        tri_nodes(a, t, v);
        sort3(v);
        if (v[0] == v[1] || v[1] == v[2]) {
            // face_extraction_kernels.py:310-318: dropped
            tri3[t * 3 + 0] = tri3[t * 3 + 1] = tri3[t * 3 + 2] = kNoTri;
            key[t] = ~(uint64_t)0;
            val[t] = (int32_t)t;
            continue;
        }
        tri3[t * 3 + 0] = v[0];
        tri3[t * 3 + 1] = v[1];
        tri3[t * 3 + 2] = v[2];
        key[t] = packed ? (((uint64_t)v[0] << 42) | ((uint64_t)v[1] << 21) | (uint64_t)v[2])
                        : (((uint64_t)(uint32_t)v[1] << 32) | (uint64_t)(uint32_t)v[2]);
        val[t] = (int32_t)t;
    }
}

// second pass of the two-pass sort: key = first id of the triple, in the current order
__global__ void __launch_bounds__(kThreads) k_first_keys(int64_t n, const int32_t* __restrict__ tri3,
                                                         const int32_t* __restrict__ val, uint64_t* __restrict__ key) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        key[i] = (uint64_t)(uint32_t)tri3[(int64_t)val[i] * 3];
}

__device__ __forceinline__ bool same3(const int32_t* tri3, int32_t t, int32_t u) {
    const int32_t* p = tri3 + (int64_t)t * 3;
    const int32_t* q = tri3 + (int64_t)u * 3;
    return p[0] == q[0] && p[1] == q[1] && p[2] == q[2];
}

__global__ void __launch_bounds__(kThreads) k_run_flags(int64_t n, const int32_t* __restrict__ tri3,
                                                        const int32_t* __restrict__ sorted_t,
                                                        int32_t* __restrict__ flag) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        flag[i] = (!no_tri(tri3, sorted_t[i]) && (i == 0 || !same3(tri3, sorted_t[i], sorted_t[i - 1]))) ? 1 : 0;
}

// cube-face code of the triangle t's node set on cell c; -1 unless exactly one face matches
__device__ int face_code(const TopoArgs& a, int64_t c, const int32_t v[3]) {
    const bool prism = c < a.n_prism;
    const int32_t* cn = prism ? a.prism_conn + c * 6 : a.tet_conn + (c - a.n_prism) * 4;
    int found = -1, n_match = 0;
    for (int code = 0; code < 6; ++code) {
        const int* loc = prism ? kPrismFace[code] : kTetFace[code];
        if (loc[0] < 0) continue;
        const int nc = loc[3] < 0 ? 3 : 4;
        bool all = true;
        for (int k = 0; k < 3; ++k) {
            bool in = false;
            for (int m = 0; m < nc; ++m) in |= cn[loc[m]] == v[k];
            all &= in;
        }
        if (all) {
            found = code;
            ++n_match;
        }
    }
    return n_match == 1 ? found : -1;
}

__global__ void __launch_bounds__(kThreads) k_records(TopoArgs a, const double* __restrict__ nodes,
                                                      const int32_t* __restrict__ tri3,
                                                      const int32_t* __restrict__ sorted_t,
                                                      const int32_t* __restrict__ flag,
                                                      const int32_t* __restrict__ uid, AfrTopoDesc d) {
#pragma clang fp contract(off)
    const int64_t n = a.n_tri;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        if (i == n - 1) *d.n_faces = uid[i];
        if (!flag[i]) continue;
        const int64_t f = (int64_t)uid[i] - 1;
        const int32_t t0 = sorted_t[i];
        int j;
        const int64_t owner = cell_of(a, t0, &j);
        int64_t neighbor = -1;
        int32_t t1 = -1;
        if (i + 1 < n && !flag[i + 1] && !no_tri(tri3, sorted_t[i + 1])) {
            t1 = sorted_t[i + 1];
            neighbor = cell_of(a, t1, &j);
            if (i + 2 < n && !flag[i + 2] && !no_tri(tri3, sorted_t[i + 2])) atomicOr(d.status, 1);
        }
        int32_t v[3];
        tri_nodes(a, t0, v);
        const int oc = face_code(a, owner, v);
        int nc = -1;
        if (neighbor >= 0) {
            int32_t w[3];
            tri_nodes(a, t1, w);
            nc = face_code(a, neighbor, w);
            if (nc < 0) atomicOr(d.status, 2);
        }
        if (oc < 0) atomicOr(d.status, 2);
        d.owner[f] = (int32_t)owner;
        d.neighbor[f] = (int32_t)neighbor;
        d.owner_code[f] = oc;
        d.neighbor_code[f] = nc;
        const int32_t* s = tri3 + (int64_t)t0 * 3;
        d.node_ids[f * 3 + 0] = s[0];
        d.node_ids[f * 3 + 1] = s[1];
        d.node_ids[f * 3 + 2] = s[2];
        // geometry over the sorted triple, in FP64, without fused multiply-adds
        double p[3][3];
        for (int k = 0; k < 3; ++k)
            for (int x = 0; x < 3; ++x) p[k][x] = nodes[(int64_t)s[k] * 3 + x];
        double cen[3], e1[3], e2[3];
        for (int x = 0; x < 3; ++x) {
            cen[x] = (p[0][x] + p[1][x] + p[2][x]) / 3.0;
            e1[x] = p[1][x] - p[0][x];
            e2[x] = p[2][x] - p[0][x];
        }
        double av[3] = {0.5 * (e1[1] * e2[2] - e1[2] * e2[1]), 0.5 * (e1[2] * e2[0] - e1[0] * e2[2]),
                        0.5 * (e1[0] * e2[1] - e1[1] * e2[0])};
        // owner vertex centroid: sequential sum of its vertices, then the mean
        const bool prism = owner < a.n_prism;
        const int nv = prism ? 6 : 4;
        const int32_t* cn = prism ? a.prism_conn + owner * 6 : a.tet_conn + (owner - a.n_prism) * 4;
        double cc[3];
        for (int x = 0; x < 3; ++x) {
            double acc = nodes[(int64_t)cn[0] * 3 + x];
            for (int k = 1; k < nv; ++k) acc = acc + nodes[(int64_t)cn[k] * 3 + x];
            cc[x] = acc / (double)nv;
        }
        const double dot = av[0] * (cen[0] - cc[0]) + av[1] * (cen[1] - cc[1]) + av[2] * (cen[2] - cc[2]);
        if (dot < 0.0) {
            av[0] = -av[0];
            av[1] = -av[1];
            av[2] = -av[2];
        }
        const double area = sqrt(av[0] * av[0] + av[1] * av[1] + av[2] * av[2]);
        d.area[f] = area;
        for (int x = 0; x < 3; ++x) {
            d.normal[f * 3 + x] = av[x] / area;
            d.center[f * 3 + x] = cen[x];
        }
    }
}

int grid_of(int64_t n) {
    const int64_t b = (n + kThreads - 1) / kThreads;
    return (int)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// workspace layout: tri3 | key_a | key_b | val_a | val_b | flag | uid | sort temp | scan temp
struct Layout {
    size_t tri3, key_a, key_b, val_a, val_b, flag, uid, sort_tmp, scan_tmp, sort_bytes, scan_bytes, total;
};

int layout_of(int64_t n_tri, Layout* L) {
    size_t sort_bytes = 0, scan_bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                           (int32_t*)nullptr, (int32_t*)nullptr, (int)n_tri, 0, 64) != hipSuccess)
        return -1;
    if (hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, (int32_t*)nullptr, (int32_t*)nullptr, (int)n_tri) !=
        hipSuccess)
        return -1;
    size_t o = 0;
    L->tri3 = o; o += align256((size_t)n_tri * 12);
    L->key_a = o; o += align256((size_t)n_tri * 8);
    L->key_b = o; o += align256((size_t)n_tri * 8);
    L->val_a = o; o += align256((size_t)n_tri * 4);
    L->val_b = o; o += align256((size_t)n_tri * 4);
    L->flag = o; o += align256((size_t)n_tri * 4);
    L->uid = o; o += align256((size_t)n_tri * 4);
    L->sort_tmp = o; o += align256(sort_bytes);
    L->scan_tmp = o; o += align256(scan_bytes);
    L->sort_bytes = sort_bytes;
    L->scan_bytes = scan_bytes;
    L->total = o;
    return 0;
}

int bits_for(int64_t n) {
    int b = 1;
    while (b < 31 && ((int64_t)1 << b) < n) ++b;
    return b;
}

}  // namespace

extern "C" {

int afr_face_topology_workspace(int64_t n_prism, int64_t n_tet, size_t* bytes) {
    if (!bytes || n_prism < 0 || n_tet < 0) {
        afr_set_error("afr_face_topology_workspace: bad argument");
        return -1;
    }
    const int64_t n_tri = 8 * n_prism + 4 * n_tet;
    if (n_tri >= ((int64_t)1 << 31)) {
        afr_set_error("afr_face_topology_workspace: more than 2^31 triangles");
        return -1;
    }
    Layout L;
    if (layout_of(n_tri < 1 ? 1 : n_tri, &L)) {
        afr_set_error("afr_face_topology_workspace: hipcub size query failed");
        return -1;
    }
    *bytes = L.total;
    return 0;
}

int afr_face_topology(const AfrTopoDesc* d, void* stream) {
    if (!d || d->abi_version != AFR_ABI_VERSION || d->n_nodes < 1 || d->n_nodes >= ((int64_t)1 << 31) ||
        d->n_prism < 0 || d->n_tet < 0 || !d->nodes || (d->n_prism && !d->prism_conn) || (d->n_tet && !d->tet_conn) ||
        !d->work || !d->owner || !d->neighbor || !d->owner_code || !d->neighbor_code || !d->node_ids ||
        !d->normal || !d->area || !d->center || !d->n_faces || !d->status) {
        afr_set_error("afr_face_topology: bad descriptor");
        return -1;
    }
    const int64_t n_tri = 8 * d->n_prism + 4 * d->n_tet;
    if (n_tri < 1 || n_tri >= ((int64_t)1 << 31)) {
        afr_set_error("afr_face_topology: no cells, or more than 2^31 triangles");
        return -1;
    }
    Layout L;
    if (layout_of(n_tri, &L) || d->work_bytes < L.total) {
        afr_set_error("afr_face_topology: workspace smaller than afr_face_topology_workspace asks");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    if (hipMemsetAsync(d->status, 0, sizeof(int32_t), st) != hipSuccess) {
        afr_set_error("afr_face_topology: status reset failed");
        return -1;
    }
    char* w = (char*)d->work;
    int32_t* tri3 = (int32_t*)(w + L.tri3);
    uint64_t* key_a = (uint64_t*)(w + L.key_a);
    uint64_t* key_b = (uint64_t*)(w + L.key_b);
    int32_t* val_a = (int32_t*)(w + L.val_a);
    int32_t* val_b = (int32_t*)(w + L.val_b);
    int32_t* flag = (int32_t*)(w + L.flag);
    int32_t* uid = (int32_t*)(w + L.uid);
    TopoArgs a{d->n_prism, d->n_tet, n_tri, d->prism_conn, d->tet_conn};
    const int nb = bits_for(d->n_nodes);
    const bool packed = nb <= 21;
    const int g = grid_of(n_tri);
    hipLaunchKernelGGL(k_tri_keys, dim3(g), dim3(kThreads), 0, st, a, tri3, key_a, val_a, packed ? 1 : 0);
    size_t sb = L.sort_bytes;
    hipError_t e;
    if (packed) {
        e = hipcub::DeviceRadixSort::SortPairs(w + L.sort_tmp, sb, key_a, key_b, val_a, val_b, (int)n_tri, 0,
                                               42 + nb, st);
    } else {
        // low pair first, then stably by the first id
        e = hipcub::DeviceRadixSort::SortPairs(w + L.sort_tmp, sb, key_a, key_b, val_a, val_b, (int)n_tri, 0,
                                               32 + nb, st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_first_keys, dim3(g), dim3(kThreads), 0, st, n_tri, tri3, val_b, key_a);
            sb = L.sort_bytes;
            e = hipcub::DeviceRadixSort::SortPairs(w + L.sort_tmp, sb, key_a, key_b, val_b, val_a, (int)n_tri, 0,
                                                   nb, st);
            if (e == hipSuccess) {
                // the result is in val_a: copy it where the one-pass path leaves it
                e = hipMemcpyAsync(val_b, val_a, (size_t)n_tri * 4, hipMemcpyDeviceToDevice, st);
            }
        }
    }
    if (e != hipSuccess) {
        afr_set_error("afr_face_topology: radix sort failed");
        return -1;
    }
    hipLaunchKernelGGL(k_run_flags, dim3(g), dim3(kThreads), 0, st, n_tri, tri3, val_b, flag);
    size_t cb = L.scan_bytes;
    if (hipcub::DeviceScan::InclusiveSum(w + L.scan_tmp, cb, flag, uid, (int)n_tri, st) != hipSuccess) {
        afr_set_error("afr_face_topology: scan failed");
        return -1;
    }
    hipLaunchKernelGGL(k_records, dim3(g), dim3(kThreads), 0, st, a, d->nodes, tri3, val_b, flag, uid, *d);
    if (hipGetLastError() != hipSuccess) {
        afr_set_error("afr_face_topology: launch failed");
        return -1;
    }
    return 0;
}

}  // extern "C"
