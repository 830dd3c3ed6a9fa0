// Cell-centric face slots on the GPU: the integer and normal tables of the slot form of the
// reference's flat face geometry (src/autoflowcfd/fr/face_flux_points_merge.py:83-124,
// face_flux_points.py:164-236, face_flux_points_data.py:145-217), the device counterpart of the
// host builder's combinatorial part.
//
// Every real cube face of every cell is the primary side of exactly one face record; a slot is
// (cell, real cube face) -> that record, the role (owner or neighbour), the other cell(s), the
// boundary code and row, the true normal signed for the role.  The cross-interpolation rows depend
// only on (own type, own face, other type, other face, corner correspondence), so slots are keyed
// on that and one representative per key is located (on the host, a handful of points); the key's
// rank among the distinct keys is the slot's row block.
//
//   k_group    per record: the records of a prism quad cube face form a group per side; its
//              primary is the smallest record index, its partner (interior records only) the
//              other one: atomic min / max per (cell, slot), independent of order
//   k_assign   per record: primary sides claim their (cell, slot), counted
//   k_slots    per slot: checks the count, writes the slot tables and the key (corner
//              correspondence by shared node ids, or by position for periodic images)
//   then a stable radix sort of (key, slot) pairs, run flags and a scan number the distinct keys;
//   k_blocks writes each keyed slot's block and each block's first slot (its representative).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <climits>
#include <string>

#include "../../include/afr.h"

void afr_set_error(const std::string& msg);

namespace {

constexpr int kThreads = 256;
constexpr int kSlots = 5;
constexpr uint32_t kNoKey = 0xFFFFu;   // keys are < 2^16: ((1*6+5)*2+1)*6+5 = 143, times 256, + 255

__constant__ int kPrismFace[6][4] = {{0, 2, 5, 3}, {1, 2, 5, 4}, {0, 1, 4, 3}, {-1, -1, -1, -1},
                                     {0, 1, 2, -1}, {3, 4, 5, -1}};
__constant__ int kTetFace[6][4] = {{0, 2, 3, -1}, {1, 2, 3, -1}, {0, 1, 3, -1}, {-1, -1, -1, -1},
                                   {0, 1, 2, -1}, {-1, -1, -1, -1}};
// real cube faces per slot (prism: a-, a+, b-, c-, c+; tet: a-, a+, b-, c-) and back
__constant__ int kSlotCode[2][5] = {{0, 1, 2, 4, 5}, {0, 1, 2, 4, -1}};
__constant__ int kCodeSlot[2][6] = {{0, 1, 2, -1, 3, 4}, {0, 1, 2, -1, 3, -1}};

struct SlotWork {
    int32_t *omin, *omax, *omin_i, *omax_i, *nmin, *nmax, *cnt, *role;
    uint32_t *key_a, *key_b;
    int32_t *val_a, *flag, *uid;
};

__device__ __forceinline__ bool prism_quad(int64_t cell, int code, int64_t n_prism) {
    return cell >= 0 && cell < n_prism && code >= 0 && code <= 2;
}

__device__ __forceinline__ int slot_of(int64_t cell, int code, int64_t n_prism) {
    if (code < 0 || code > 5) return -1;
    return kCodeSlot[cell < n_prism ? 0 : 1][code];
}

__global__ void __launch_bounds__(kThreads) k_init(int64_t n_slots, SlotWork w, int32_t* __restrict__ rec) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n_slots; i += (int64_t)gridDim.x * kThreads) {
        w.omin[i] = INT_MAX;
        w.omin_i[i] = INT_MAX;
        w.nmin[i] = INT_MAX;
        w.omax[i] = -1;
        w.omax_i[i] = -1;
        w.nmax[i] = -1;
        w.cnt[i] = 0;
        w.role[i] = 0;
        rec[i] = -1;
    }
}

__global__ void __launch_bounds__(kThreads) k_group(AfrSlotDesc d, SlotWork w) {
    for (int64_t f = (int64_t)blockIdx.x * kThreads + threadIdx.x; f < d.n_records; f += (int64_t)gridDim.x * kThreads) {
        const int64_t o = d.owner[f], nb = d.neighbor[f];
        const int oc = d.owner_code[f], nc = d.neighbor_code[f];
        if (prism_quad(o, oc, d.n_prism)) {
            const int64_t k = o * kSlots + slot_of(o, oc, d.n_prism);
            atomicMin(&w.omin[k], (int32_t)f);
            atomicMax(&w.omax[k], (int32_t)f);
            if (nb >= 0) {
                atomicMin(&w.omin_i[k], (int32_t)f);
                atomicMax(&w.omax_i[k], (int32_t)f);
            }
        }
        if (nb >= 0 && prism_quad(nb, nc, d.n_prism)) {
            const int64_t k = nb * kSlots + slot_of(nb, nc, d.n_prism);
            atomicMin(&w.nmin[k], (int32_t)f);
            atomicMax(&w.nmax[k], (int32_t)f);
        }
    }
}

__global__ void __launch_bounds__(kThreads) k_assign(AfrSlotDesc d, SlotWork w) {
    for (int64_t f = (int64_t)blockIdx.x * kThreads + threadIdx.x; f < d.n_records; f += (int64_t)gridDim.x * kThreads) {
        const int64_t o = d.owner[f], nb = d.neighbor[f];
        const int oc = d.owner_code[f], nc = d.neighbor_code[f];
        const int so = slot_of(o, oc, d.n_prism);
        const bool op = !prism_quad(o, oc, d.n_prism) || f == w.omin[o * kSlots + so];
        if (op) {
            if (so < 0) {
                atomicOr(d.status, 1);
            } else {
                atomicAdd(&w.cnt[o * kSlots + so], 1);
                d.rec[o * kSlots + so] = (int32_t)f;
            }
        }
        if (nb >= 0) {
            const int sn = slot_of(nb, nc, d.n_prism);
            const bool np_ = !prism_quad(nb, nc, d.n_prism) || f == w.nmin[nb * kSlots + sn];
            if (np_) {
                if (sn < 0) {
                    atomicOr(d.status, 1);
                } else {
                    atomicAdd(&w.cnt[nb * kSlots + sn], 1);
                    d.rec[nb * kSlots + sn] = (int32_t)f;
                    w.role[nb * kSlots + sn] = 1;
                }
            }
        }
    }
}

__device__ void face_corners(const AfrSlotDesc& d, int64_t cell, int code, int32_t out[4]) {
    out[0] = out[1] = out[2] = out[3] = -1;
    if (cell < 0 || code < 0 || code > 5) return;
    const bool prism = cell < d.n_prism;
    const int* loc = prism ? kPrismFace[code] : kTetFace[code];
    const int32_t* cn = prism ? d.prism_conn + cell * 6 : d.tet_conn + (cell - d.n_prism) * 4;
    for (int m = 0; m < 4; ++m)
        if (loc[m] >= 0) out[m] = cn[loc[m]];
}

__global__ void __launch_bounds__(kThreads) k_slots(AfrSlotDesc d, SlotWork w) {
#pragma clang fp contract(off)
    const int64_t C = d.n_prism + d.n_tet;
    const int64_t n_slots = C * kSlots;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n_slots; i += (int64_t)gridDim.x * kThreads) {
        const int64_t c = i / kSlots;
        const int s = (int)(i % kSlots);
        const int ct = c < d.n_prism ? 0 : 1;
        const int expect = (ct == 0 || s < 4) ? 1 : 0;
        if (w.cnt[i] != expect) atomicOr(d.status, 1);
        const int own_code = kSlotCode[ct][s];
        const int32_t r = w.cnt[i] == 1 ? d.rec[i] : -1;
        const bool valid = r >= 0;
        const bool is_own = w.role[i] == 0;
        const bool bnd = valid && d.neighbor[r] < 0;
        d.slot_bc[i] = valid ? (is_own || bnd ? d.bc[r] : -1) : -2;
        if (d.slot_bc_row) d.slot_bc_row[i] = (valid && bnd && d.bc_row) ? d.bc_row[r] : -1;
        // an unused slot carries record 0's normal and translation, as the host tables do
        const int64_t r0 = valid ? r : 0;
        const double sg = is_own ? 1.0 : -1.0;
        double tr[3];
        for (int x = 0; x < 3; ++x) {
            d.tn[i * 3 + x] = d.normal[r0 * 3 + x] * sg;
            tr[x] = d.translation[r0 * 3 + x] * sg;
            d.trans[i * 3 + x] = tr[x];
        }
        int64_t other0 = -1, other1 = -1;
        int ocode0 = -1, ocode1 = -1;
        if (valid) {
            other0 = is_own ? d.neighbor[r] : d.owner[r];
            ocode0 = is_own ? d.neighbor_code[r] : d.owner_code[r];
            int32_t part = -1;
            if (is_own) {
                const int64_t o = d.owner[r];
                const int oc = d.owner_code[r];
                if (!bnd && prism_quad(o, oc, d.n_prism)) {
                    const int64_t k = o * kSlots + slot_of(o, oc, d.n_prism);
                    if (w.omin_i[k] == r && w.omax_i[k] != r) part = w.omax_i[k];
                }
            } else {
                const int64_t nb = d.neighbor[r];
                const int nc = d.neighbor_code[r];
                if (prism_quad(nb, nc, d.n_prism)) {
                    const int64_t k = nb * kSlots + slot_of(nb, nc, d.n_prism);
                    if (w.nmin[k] == r && w.nmax[k] != r) part = w.nmax[k];
                }
            }
            if (part >= 0) {
                other1 = is_own ? d.neighbor[part] : d.owner[part];
                ocode1 = is_own ? d.neighbor_code[part] : d.owner_code[part];
            }
        }
        const bool interior = valid && !bnd;
        const bool split = interior && other1 >= 0 && other1 != other0;
        d.nb[i * 2 + 0] = interior ? (int32_t)other0 : -1;
        d.nb[i * 2 + 1] = split ? (int32_t)other1 : -1;
        d.other[i * 4 + 0] = (int32_t)other0;
        d.other[i * 4 + 1] = ocode0;
        d.other[i * 4 + 2] = (int32_t)other1;
        d.other[i * 4 + 3] = ocode1;
        const bool keyed = interior && !split;
        d.slot_flags[i] = (keyed ? 1 : 0) | (split ? 2 : 0);
        uint32_t key = kNoKey;
        if (keyed) {
            int32_t own[4], oth[4];
            face_corners(d, c, own_code, own);
            face_corners(d, other0, ocode0, oth);
            const int ncorner = own[3] >= 0 ? 4 : 3;
            int pi[4] = {0, 0, 0, 0};
            bool ok = true;
            for (int a = 0; a < 4; ++a) {
                bool any = false;
                int first = 0;
                for (int b = 3; b >= 0; --b) {
                    if (own[a] >= 0 && own[a] == oth[b]) {
                        any = true;
                        first = b;
                    }
                }
                pi[a] = first;
                ok &= any || own[a] < 0;
            }
            if (!ok) {
                // periodic image: corners by position, own corner + translation
                double xo[4][3], xt[4][3];
                for (int a = 0; a < 4; ++a)
                    for (int x = 0; x < 3; ++x) {
                        xo[a][x] = d.nodes[(int64_t)(own[a] > 0 ? own[a] : 0) * 3 + x] + tr[x];
                        xt[a][x] = d.nodes[(int64_t)(oth[a] > 0 ? oth[a] : 0) * 3 + x];
                    }
                double worst = 0.0;
                for (int a = 0; a < 4; ++a) {
                    double best = 0.0;
                    int bj = 0;
                    for (int b = 0; b < 4; ++b) {
                        double dd = 0.0;
                        if (b == 3 && oth[3] < 0) {
                            dd = __builtin_inf();
                        } else {
                            const double e0 = xo[a][0] - xt[b][0], e1 = xo[a][1] - xt[b][1], e2 = xo[a][2] - xt[b][2];
                            dd = e0 * e0 + e1 * e1 + e2 * e2;
                        }
                        if (b == 0 || dd < best) {
                            best = dd;
                            bj = b;
                        }
                    }
                    pi[a] = bj;
                    if (a < 3 && best > worst) worst = best;
                }
                const double e0 = xo[1][0] - xo[0][0], e1 = xo[1][1] - xo[0][1], e2 = xo[1][2] - xo[0][2];
                const double scale = e0 * e0 + e1 * e1 + e2 * e2;
                if (worst > 1e-12 * scale + 1e-300) atomicOr(d.status, 2);
            }
            if (ncorner == 3) pi[3] = 0;
            const int oct = other0 >= d.n_prism ? 1 : 0;
            key = (uint32_t)((((ct * 6 + own_code) * 2 + oct) * 6 + ocode0) * 256 + pi[0] + 4 * pi[1] + 16 * pi[2] +
                             64 * pi[3]);
        }
        w.key_a[i] = key;
        w.val_a[i] = (int32_t)i;
        d.blk[i * 2 + 0] = 0;
        d.blk[i * 2 + 1] = 0;
    }
}

__global__ void __launch_bounds__(kThreads) k_key_runs(int64_t n, const uint32_t* __restrict__ key,
                                                       int32_t* __restrict__ flag) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        flag[i] = (key[i] != kNoKey && (i == 0 || key[i] != key[i - 1])) ? 1 : 0;
}

__global__ void __launch_bounds__(kThreads) k_blocks(int64_t n, const uint32_t* __restrict__ key,
                                                     const int32_t* flat, const int32_t* __restrict__ flag,
                                                     const int32_t* __restrict__ uid, AfrSlotDesc d) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        if (key[i] == kNoKey) continue;
        const int32_t b = uid[i] - 1;
        d.blk[(int64_t)flat[i] * 2] = b;
        if (flag[i]) {
            d.rep[b] = flat[i];
            d.run_start[b] = (int32_t)i;
        }
        if (i == n - 1 || key[i + 1] == kNoKey) {
            d.counts[0] = uid[i];        // distinct keys
            d.counts[1] = (int32_t)(i + 1);  // keyed slots
        }
    }
}

int grid_of(int64_t n) {
    const int64_t b = (n + kThreads - 1) / kThreads;
    return (int)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Layout {
    size_t i32[11], key_a, key_b, sort_tmp, scan_tmp, sort_bytes, scan_bytes, total;
};

int layout_of(int64_t n, Layout* L) {
    size_t sort_bytes = 0, scan_bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (int32_t*)nullptr, (int32_t*)nullptr, (int)n, 0, 16) != hipSuccess)
        return -1;
    if (hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, (int32_t*)nullptr, (int32_t*)nullptr, (int)n) !=
        hipSuccess)
        return -1;
    size_t o = 0;
    for (int k = 0; k < 11; ++k) {
        L->i32[k] = o;
        o += align256((size_t)n * 4);
    }
    L->key_a = o; o += align256((size_t)n * 4);
    L->key_b = o; o += align256((size_t)n * 4);
    L->sort_tmp = o; o += align256(sort_bytes);
    L->scan_tmp = o; o += align256(scan_bytes);
    L->sort_bytes = sort_bytes;
    L->scan_bytes = scan_bytes;
    L->total = o;
    return 0;
}

}  // namespace

extern "C" {

int afr_face_slots_workspace(int64_t n_cells, size_t* bytes) {
    if (!bytes || n_cells < 1 || n_cells * kSlots >= ((int64_t)1 << 31)) {
        afr_set_error("afr_face_slots_workspace: bad cell count");
        return -1;
    }
    Layout L;
    if (layout_of(n_cells * kSlots, &L)) {
        afr_set_error("afr_face_slots_workspace: hipcub size query failed");
        return -1;
    }
    *bytes = L.total;
    return 0;
}

int afr_face_slots(const AfrSlotDesc* d, void* stream) {
    if (!d || d->abi_version != AFR_ABI_VERSION || d->n_prism < 0 || d->n_tet < 0 || d->n_records < 1 ||
        d->n_records >= ((int64_t)1 << 31) || !d->nodes || (d->n_prism && !d->prism_conn) ||
        (d->n_tet && !d->tet_conn) || !d->owner || !d->neighbor || !d->owner_code || !d->neighbor_code || !d->bc ||
        !d->normal || !d->translation || !d->work || !d->nb || !d->blk || !d->slot_bc || !d->tn || !d->trans ||
        !d->other || !d->rec || !d->slot_flags || !d->rep || !d->run_start || !d->sorted_slot || !d->counts ||
        !d->status) {
        afr_set_error("afr_face_slots: bad descriptor");
        return -1;
    }
    const int64_t C = d->n_prism + d->n_tet;
    const int64_t n = C * kSlots;
    if (C < 1 || n >= ((int64_t)1 << 31)) {
        afr_set_error("afr_face_slots: no cells, or more than 2^31 slots");
        return -1;
    }
    Layout L;
    if (layout_of(n, &L) || d->work_bytes < L.total) {
        afr_set_error("afr_face_slots: workspace smaller than afr_face_slots_workspace asks");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    if (hipMemsetAsync(d->status, 0, sizeof(int32_t), st) != hipSuccess ||
        hipMemsetAsync(d->counts, 0, 2 * sizeof(int32_t), st) != hipSuccess) {
        afr_set_error("afr_face_slots: reset failed");
        return -1;
    }
    char* base = (char*)d->work;
    SlotWork w;
    int32_t* p[11];
    for (int k = 0; k < 11; ++k) p[k] = (int32_t*)(base + L.i32[k]);
    w.omin = p[0]; w.omax = p[1]; w.omin_i = p[2]; w.omax_i = p[3]; w.nmin = p[4]; w.nmax = p[5];
    w.cnt = p[6]; w.role = p[7]; w.val_a = p[8]; w.flag = p[9]; w.uid = p[10];
    w.key_a = (uint32_t*)(base + L.key_a);
    w.key_b = (uint32_t*)(base + L.key_b);
    const int gs = grid_of(n), gr = grid_of(d->n_records);
    hipLaunchKernelGGL(k_init, dim3(gs), dim3(kThreads), 0, st, n, w, d->rec);
    hipLaunchKernelGGL(k_group, dim3(gr), dim3(kThreads), 0, st, *d, w);
    hipLaunchKernelGGL(k_assign, dim3(gr), dim3(kThreads), 0, st, *d, w);
    hipLaunchKernelGGL(k_slots, dim3(gs), dim3(kThreads), 0, st, *d, w);
    // the slot indices in key order are an output (the host samples each block's members from it)
    int32_t* val_b = d->sorted_slot;
    size_t sb = L.sort_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(base + L.sort_tmp, sb, w.key_a, w.key_b, w.val_a, val_b, (int)n, 0, 16,
                                           st) != hipSuccess) {
        afr_set_error("afr_face_slots: radix sort failed");
        return -1;
    }
    hipLaunchKernelGGL(k_key_runs, dim3(gs), dim3(kThreads), 0, st, n, w.key_b, w.flag);
    size_t cb = L.scan_bytes;
    if (hipcub::DeviceScan::InclusiveSum(base + L.scan_tmp, cb, w.flag, w.uid, (int)n, st) != hipSuccess) {
        afr_set_error("afr_face_slots: scan failed");
        return -1;
    }
    hipLaunchKernelGGL(k_blocks, dim3(gs), dim3(kThreads), 0, st, n, w.key_b, val_b, w.flag, w.uid, *d);
    if (hipGetLastError() != hipSuccess) {
        afr_set_error("afr_face_slots: launch failed");
        return -1;
    }
    return 0;
}

}  // extern "C"
