// Device helpers shared by the forward RK4 stage kernels (puff_push.hip) and
// their adjoints (puff_adjoint.hip): cell location, undulator fields, the
// workgroup's cell box and the weighted scatter onto the mesh.
//
// Reference arithmetic restated (UKFELs/Puffin):
//   getInterps_3D   interaction/system_interpolation.f90:25-98
//   1-D branch      interaction/system_interpolation_1D.f90:24-179
//   getSource_3D    interaction/system_interpolation.f90:150-248
//   b-fields        undulator/bfields.f90:36-881 (per-stage coefficients are
//                   computed on the host, see stage_scalars in puff_ctx.hip)
#pragma once

#include "puff_common.hpp"
#include "puff_kernels.hpp"

namespace puff {

// Cell of one particle and its interpolation fractions.
template <typename R>
struct Cell {
    int ix, iy, iz0, iz1;
    R fx, fy, fz;   // x_in2, y_in2, z2_in2
};

// Node and fractions of a particle (getInterps_3D / _1D).  Returns flag bits;
// 0 means every node index is inside the mesh.
template <typename R>
__device__ __forceinline__ int locate(const MeshGeom& m, R x, R y, double z2,
                                      Cell<R>& c) {
    int flag = 0;
    const double zq = floor(z2 * m.inv_dz2);
    if (m.periodic) {
        if (!(fabs(zq) < 1.0e9)) return FLAG_Z2MESH;
        long long q = (long long)zq;
        long long r = q % m.nzp;
        if (r < 0) r += m.nzp;
        c.iz0 = (int)r;
        c.iz1 = (c.iz0 + 1 == m.nzp) ? 0 : c.iz0 + 1;
    } else {
        // temporal mesh: z2node >= NZ2 stops the reference
        // (system_interpolation.f90:70-75); z2 < 0 is refused too.
        if (!(zq >= 0.0 && zq <= (double)(m.nzp - 2))) return FLAG_Z2MESH;
        c.iz0 = (int)zq;
        c.iz1 = c.iz0 + 1;
    }
    const double locz2 = z2 - zq * m.dz2;
    c.fz = (R)(locz2 * m.inv_dz2);
    if (m.one_d) {
        c.ix = 0;
        c.iy = 0;
        c.fx = 0;
        c.fy = 0;
        return 0;
    }
    const R tx = (x + (R)m.halfx) * (R)m.inv_dx;
    const R ty = (y + (R)m.halfy) * (R)m.inv_dy;
    const R xq = floor(tx), yq = floor(ty);
    // xnode >= nspinDX .or. xnode < 1 (system_interpolation.f90:60-68)
    if (!(xq >= (R)0 && xq <= (R)(m.nx - 2))) flag |= FLAG_TRANSVERSE;
    if (!(yq >= (R)0 && yq <= (R)(m.ny - 2))) flag |= FLAG_TRANSVERSE;
    if (flag) return flag;
    c.ix = (int)xq;
    c.iy = (int)yq;
    c.fx = (x + (R)m.halfx - xq * (R)m.dx) * (R)m.inv_dx;
    c.fy = (y + (R)m.halfy - yq * (R)m.dy) * (R)m.inv_dy;
    return 0;
}

// Scaled undulator field at one particle from the stage coefficients
// (bfields.f90): see stage_scalars() on the host for the coefficient map.
template <typename R>
__device__ __forceinline__ void bfield(const StageScalars& s, R x, R y, R& bx,
                                       R& by, R& bz) {
    if (s.und_type == PUFF_UND_PLANEPOLE) {
        bx = (R)s.c[3] * y;
        by = (R)s.c[1] + (R)s.c[4] * x;
        bz = (R)s.c[2] * sinh((R)s.c[7] * y);
    } else if (s.und_type == PUFF_UND_CURVED) {
        const R kx = (R)s.c[5], ky = (R)s.c[6];
        bx = (R)s.c[0] * sinh(kx * x) * sinh(ky * y) + (R)s.c[3] * y;
        by = (R)s.c[1] * cosh(kx * x) * cosh(ky * y) + (R)s.c[4] * x;
        bz = (R)s.c[2] * cosh(kx * x) * sinh(ky * y);
    } else {
        bx = (R)s.c[0] + (R)s.c[3] * y;
        by = (R)s.c[1] + (R)s.c[4] * x;
        bz = (R)s.c[5] * x + (R)s.c[6] * y;
    }
}

// Transverse derivatives of bfield(): g = {dbx/dx, dbx/dy, dby/dx, dby/dy,
// dbz/dx, dbz/dy}.
__device__ __forceinline__ void bfield_grad(const StageScalars& s, double x, double y,
                                            double g[6]) {
    g[0] = 0.0;
    g[1] = s.c[3];
    g[2] = s.c[4];
    g[3] = 0.0;
    if (s.und_type == PUFF_UND_PLANEPOLE) {
        g[4] = 0.0;
        g[5] = s.c[2] * s.c[7] * cosh(s.c[7] * y);
    } else if (s.und_type == PUFF_UND_CURVED) {
        const double kx = s.c[5], ky = s.c[6];
        const double shx = sinh(kx * x), chx = cosh(kx * x), shy = sinh(ky * y), chy = cosh(ky * y);
        g[0] = s.c[0] * kx * chx * shy;
        g[1] += s.c[0] * ky * shx * chy;
        g[2] += s.c[1] * kx * shx * chy;
        g[3] = s.c[1] * ky * chx * shy;
        g[4] = s.c[2] * kx * shx * shy;
        g[5] = s.c[2] * ky * chx * chy;
    } else {
        g[4] = s.c[5];
        g[5] = s.c[6];
    }
}

template <typename R>
__device__ __forceinline__ void atomic_add(R* p, R v) {
    unsafeAtomicAdd(p, v);
}

// Node box of the cells of one workgroup (origin and extent in nodes, i.e.
// cells + 1), and the step's active box: one atomic per workgroup into one
// of BOX_COPIES spread copies (same-address atomics would serialise).  Must
// be reached by every thread of the workgroup.
struct TileBox {
    int bx0, by0, bz0;
    int tx, ty, tz;
    int hz;        // last cell plane
    bool any;
    int extra;     // block maximum of workgroup_box's `extra` (NX = 7)
};

// NX = 7 also reduces `extra` (a non-negative int per lane) to its block
// maximum, returned in TileBox::extra.
template <typename R, int NX = 6>
__device__ __forceinline__ TileBox workgroup_box(const MeshGeom& m, bool ok, const Cell<R>& c,
                                                 Box* box, int extra = 0) {
    __shared__ int s_red[NX][MAX_BOX_WAVES];
    __shared__ int s_bb[NX];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    // max values are reduced as min of their negatives
    int v[NX];
    v[0] = ok ? c.ix : 0x7fffffff;
    v[1] = ok ? -c.ix : 1;
    v[2] = ok ? c.iy : 0x7fffffff;
    v[3] = ok ? -c.iy : 1;
    v[4] = ok ? c.iz0 : 0x7fffffff;
    v[5] = ok ? -c.iz0 : 1;
    if constexpr (NX > 6) v[6] = -extra;
#pragma unroll
    for (int d = 0; d < NX; ++d) {
        v[d] = wave_min(v[d]);
        if (lane == 0) s_red[d][wv] = v[d];
    }
    __syncthreads();
    if (threadIdx.x < NX) {
        const int d = threadIdx.x;
        int r = s_red[d][0];
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) r = min(r, s_red[d][k]);
        s_bb[d] = (d & 1) ? -r : r;
        const bool empty = (d & 1) ? (r == 1) : (r == 0x7fffffff);
        if (d < 6 && !empty) {
            Box* b = box + (blockIdx.x % BOX_COPIES);
            if (d & 1) atomicMax(&b->hi[d >> 1], -r);
            else atomicMin(&b->lo[d >> 1], r);
        }
    }
    __syncthreads();
    TileBox t;
    t.extra = 0;
    if constexpr (NX > 6) t.extra = -s_bb[NX - 1];
    t.any = s_bb[1] >= 0;
    t.bx0 = s_bb[0];
    t.by0 = s_bb[2];
    t.bz0 = s_bb[4];
    t.hz = s_bb[5];
    t.tx = t.any ? (m.one_d ? 1 : s_bb[1] - t.bx0 + 2) : 0;
    t.ty = t.any ? (m.one_d ? 1 : s_bb[3] - t.by0 + 2) : 0;
    t.tz = t.any ? s_bb[5] - t.bz0 + 2 : 0;
    return t;
}

// Node q of a box -> (lx, ly, lz), x fastest, without integer division: the
// quotients come from the hardware reciprocal of the box extents (v_rcp_f32,
// within 1 ulp), exact while q < 2^21 (the relative error of
// (q + 1/2) * rcp(t), below 2^-22, keeps it within the distance 1/(2t) of
// (q + 1/2)/t from the next integer); boxes here hold at most a few thousand
// nodes.  (A correctly rounded 1/t is a 12-instruction sequence, twice per
// box: 6 % of a mixed-form stage kernel's VALU instructions.)
struct BoxDiv {
    int tx, ty;
    float inv_tx, inv_ty;
};

__device__ __forceinline__ BoxDiv box_div(const TileBox& t) {
    return BoxDiv{t.tx, t.ty, __builtin_amdgcn_rcpf((float)t.tx), __builtin_amdgcn_rcpf((float)t.ty)};
}

__device__ __forceinline__ void box_coords(const BoxDiv& d, int q, int& lx, int& ly, int& lz) {
    const int r = (int)(((float)q + 0.5f) * d.inv_tx);
    lx = q - r * d.tx;
    lz = (int)(((float)r + 0.5f) * d.inv_ty);
    ly = r - lz * d.ty;
}

// Union of the spread box copies.
__device__ __forceinline__ Box merged_box(const Box* box) {
    Box b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = 0x7fffffff;
        b.hi[d] = -1;
    }
    for (int k = 0; k < BOX_COPIES; ++k)
        for (int d = 0; d < 3; ++d) {
            b.lo[d] = min(b.lo[d], box[k].lo[d]);
            b.hi[d] = max(b.hi[d], box[k].hi[d]);
        }
    return b;
}

// Whether the workgroup's node box fits the LDS tile and does not wrap
// around a periodic mesh (block-uniform).
template <int TILE>
__device__ __forceinline__ bool tile_fits(const MeshGeom& m, const TileBox& t) {
    return t.any && (int64_t)t.tx * t.ty * t.tz <= TILE && !(m.periodic && t.hz + 1 >= m.nzp);
}

// Nodes of a lane's cell: first node of its lower (b0) and upper (b1)
// z2 plane, its first node in the workgroup's tile (t0), and its key
// (unique negative key on lanes without a cell).
struct CellRef {
    int b0, b1, t0, key;
};

template <typename R>
__device__ __forceinline__ CellRef cell_ref(const MeshGeom& m, bool ok, const Cell<R>& c,
                                            const TileBox& t, int lane) {
    CellRef r;
    if (m.one_d) {
        r.b0 = ok ? c.iz0 : 0;
        r.b1 = ok ? c.iz1 : 0;
        r.t0 = ok ? c.iz0 - t.bz0 : 0;
        r.key = ok ? c.iz0 : -1 - lane;
        return r;
    }
    const int ntr = m.nx * m.ny;
    r.b0 = ok ? c.iz0 * ntr + c.iy * m.nx + c.ix : 0;
    r.b1 = ok ? c.iz1 * ntr + c.iy * m.nx + c.ix : 0;
    r.t0 = ok ? ((c.iz0 - t.bz0) * t.ty + (c.iy - t.by0)) * t.tx + (c.ix - t.bx0) : 0;
    r.key = ok ? r.b0 : -1 - lane;
    return r;
}

// Weight and node of corner k of a lane's cell, in getInterps_3D's order
// (x fastest, then y, then z2): w = (1-fx | fx)(1-fy | fy)(1-fz | fz).
// 1-D: lis_GR(1) = 1 - locz2/dz2, lis_GR(2) = 1 - lis_GR(1) at nodes iz0, iz1.
template <typename R>
__device__ __forceinline__ R corner_weight(const MeshGeom& m, const Cell<R>& c, int k) {
    if (m.one_d) {
        const R w0 = (R)1 - c.fz;
        return k ? (R)1 - w0 : w0;
    }
    const R wx = (k & 1) ? c.fx : (R)1 - c.fx;
    const R wy = (k & 2) ? c.fy : (R)1 - c.fy;
    const R wz = (k & 4) ? c.fz : (R)1 - c.fz;
    return wx * wy * wz;
}

// The eight corners of a lane's 3-D cell at once, for the LDS-tile paths:
// weights w = (wx wy) wz as corner_weight forms them, and tile offsets.
// Corner label q stands for corner k = q ^ rot (rot < 8): visiting the
// labels 0..7 in order with rot = lane mod 8 sends the lanes of one cell to
// different corners at each visit, so that an instruction's LDS atomics do
// not pile onto one word.  rot = 0: getInterps_3D's order.  Formed once per
// lane: the per-corner selects of corner_weight / corner_tile under a
// lane-dependent order cost about 25 VALU instructions per corner.
template <typename R>
struct CornerSet {
    R xy[4];     // wx wy of labels 0..3 (x, then y)
    R z[2];      // wz of the labels' z plane (labels 0-3, 4-7)
    int oxy[4];  // tile offset of labels 0..3 within the z plane
    int oz[2];   // tile offset of the two z planes
};

template <typename R>
__device__ __forceinline__ CornerSet<R> corner_set(const Cell<R>& c, const TileBox& t, int rot) {
    const R ux = (R)1 - c.fx, uy = (R)1 - c.fy, uz = (R)1 - c.fz;
    const bool rx = rot & 1, ry = rot & 2, rz = rot & 4;
    const R ax0 = rx ? c.fx : ux, ax1 = rx ? ux : c.fx;
    const R ay0 = ry ? c.fy : uy, ay1 = ry ? uy : c.fy;
    CornerSet<R> s;
    s.xy[0] = ax0 * ay0;
    s.xy[1] = ax1 * ay0;
    s.xy[2] = ax0 * ay1;
    s.xy[3] = ax1 * ay1;
    s.z[0] = rz ? c.fz : uz;
    s.z[1] = rz ? uz : c.fz;
    const int ox0 = rx ? 1 : 0, ox1 = 1 - ox0;
    const int oy0 = ry ? t.tx : 0, oy1 = t.tx - oy0;
    s.oxy[0] = ox0 + oy0;
    s.oxy[1] = ox1 + oy0;
    s.oxy[2] = ox0 + oy1;
    s.oxy[3] = ox1 + oy1;
    const int txy = t.tx * t.ty;
    s.oz[0] = rz ? txy : 0;
    s.oz[1] = txy - s.oz[0];
    return s;
}

__device__ __forceinline__ int corner_node(const MeshGeom& m, const CellRef& r, int k) {
    if (m.one_d) return k ? r.b1 : r.b0;
    return ((k & 4) ? r.b1 : r.b0) + (k & 1) + ((k & 2) ? m.nx : 0);
}

__device__ __forceinline__ int corner_tile(const MeshGeom& m, const TileBox& t, const CellRef& r, int k) {
    if (m.one_d) return r.t0 + k;
    return r.t0 + (k & 1) + ((k & 2) ? t.tx : 0) + ((k & 4) ? t.tx * t.ty : 0);
}

// out[node_k] += w_k (sr, si) for every lane with a cell (sr = si = 0 on
// lanes without one).
//   use_tile: accumulation in an LDS tile over the workgroup's node box
//             (real parts, then imaginary parts; element type D), then one
//             global atomic per touched node and component.  Each lane visits the corners in a
//             rotated order so that lanes of one cell do not hit the same LDS
//             address in the same instruction.
//   else:     lanes of a wavefront that share a cell are summed first (runs
//             of equal keys), then one global atomic per run and corner.
// s_tile: the LDS tile, 2 elements of type D per node of the box.  use_tile
// must be block-uniform; every thread of the workgroup must call.
template <typename R, typename D>
__device__ __forceinline__ void scatter_cells(const MeshGeom& m, const TileBox& t, bool use_tile,
                                              bool ok, const Cell<R>& c, const CellRef& cr, R sr,
                                              R si, cplx<R>* out, D* s_tile) {
    const int lane = threadIdx.x & 63;
    const int nnb = m.one_d ? 2 : 8;
    if (use_tile) {
        const int nt = t.tx * t.ty * t.tz;
        for (int q = threadIdx.x; q < 2 * nt; q += blockDim.x) s_tile[q] = (D)0;
        __syncthreads();
        if (ok) {
            const D dsr = (D)sr, dsi = (D)si;
            if (!m.one_d) {
                const CornerSet<R> cs = corner_set<R>(c, t, lane & 7);
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const D wk = (D)(cs.xy[q & 3] * cs.z[q >> 2]);
                    const int tk = cr.t0 + cs.oxy[q & 3] + cs.oz[q >> 2];
                    atomicAdd(&s_tile[tk], wk * dsr);
                    atomicAdd(&s_tile[nt + tk], wk * dsi);
                }
            } else {
                for (int q = 0; q < nnb; ++q) {
                    const int k = (q + lane) & (nnb - 1);
                    const D wk = (D)corner_weight<R>(m, c, k);
                    const int tk = corner_tile(m, t, cr, k);
                    atomicAdd(&s_tile[tk], wk * dsr);
                    atomicAdd(&s_tile[nt + tk], wk * dsi);
                }
            }
        }
        __syncthreads();
        // flush: lanes 2j and 2j + 1 take the real and imaginary parts of
        // node j, so that one atomic instruction covers both parts of 32
        // nodes (each device-scope atomic is a memory-side request per
        // 32-byte sector and instruction: half the requests of one
        // instruction per part)
        const int ntr = m.nx * m.ny;
        const BoxDiv bd = box_div(t);
        R* outc = reinterpret_cast<R*>(out);
        for (int e = threadIdx.x; e < 2 * nt; e += blockDim.x) {
            const int q = e >> 1, part = e & 1;
            const D v = s_tile[part * nt + q];
            if (v != (D)0) {
                int lx, ly, lz;
                box_coords(bd, q, lx, ly, lz);
                const int node = (t.bz0 + lz) * ntr + (t.by0 + ly) * m.nx + (t.bx0 + lx);
                atomic_add(outc + 2 * node + part, (R)v);
            }
        }
    } else {
        const int prev = __shfl_up(cr.key, 1, 64);
        const bool head = (lane == 0) || (prev != cr.key);
        const unsigned long long heads = __ballot(head);
        const unsigned long long after = (lane == 63) ? 0ull : (heads >> (lane + 1));
        const int seg_end = after ? lane + __builtin_ctzll(after) : 63;
        const int maxlen = wave_max(head ? seg_end - lane + 1 : 1);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (k < nnb) {
                const R wk = ok ? corner_weight<R>(m, c, k) : (R)0;
                R vr = wk * sr, vi = wk * si;
                for (int o = 1; o < maxlen; o <<= 1) {
                    const R tr = __shfl_down(vr, o, 64);
                    const R ti = __shfl_down(vi, o, 64);
                    if (lane + o <= seg_end) {
                        vr += tr;
                        vi += ti;
                    }
                }
                if (ok && head) {
                    const int node = corner_node(m, cr, k);
                    atomic_add(&out[node].x, vr);
                    atomic_add(&out[node].y, vi);
                }
            }
        }
    }
}

// ---- packed fixed-point tile (mixed form) -------------------------------------

// One LDS atomic per corner for both components: each contribution becomes
// the integers qr = rint(w sr S), qi = rint(w si S) with the workgroup's
// power-of-two scale S, and qr 2^32 + qi is added to one 64-bit word of the
// tile.  S keeps every node's sums below 2^30 in magnitude (STAGE_THREADS lanes, each
// adding at most the workgroup's largest |s| to a node), so the low half
// never carries into the high half beyond what its sign extension undoes:
// the word is sum qr 2^32 + sum qi exactly, and sum qi is its low 32 bits
// read as signed.  Integer sums are exact (and independent of the atomics'
// order); each contribution is rounded once, to 2^-31 of the largest node
// sum the scale allows.  The flush divides by S and adds in FP32.
//
// In 3-D each contribution (wx wy) (wz s S) is rounded to an integer by one
// FMA with 1.5 2^23 added (|w s S| < 2^22 puts the sum in [2^23, 2^24), where
// FP32 steps by 1, so the FMA's single rounding is rint), whose bit pattern
// is 0x4B400000 + the integer: the two patterns of a contribution, as one
// 64-bit word, minus 0x4B4000004B400000 are qr 2^32 + qi (modulo 2^64).
constexpr float PACK_MAGIC = 12582912.0f;                      // 1.5 2^23
constexpr unsigned long long PACK_MAGIC2 = 0x4B4000004B400000ull;   // its pattern, twice

// log2 of S for the workgroup's largest |s| (its bit pattern, max_bits).
__device__ __forceinline__ int packed_scale_log2(int max_bits) {
    const float mx = __int_as_float(max_bits);
    if (!(mx > 0.0f)) return 0;
    int e = 0;
    frexpf(mx, &e);      // mx < 2^e
    return 30 - STAGE_THREADS_LOG2 - e;   // STAGE_THREADS mx S < 2^30
}

__device__ __forceinline__ void scatter_cells_packed(const MeshGeom& m, const TileBox& t, bool ok,
                                                     const Cell<float>& c, const CellRef& cr, float sr,
                                                     float si, int log2S, cplx<float>* out,
                                                     unsigned long long* s_tile) {
    const int lane = threadIdx.x & 63;
    const int nnb = m.one_d ? 2 : 8;
    const int nt = t.tx * t.ty * t.tz;
    for (int q = threadIdx.x; q < nt; q += blockDim.x) s_tile[q] = 0ull;
    __syncthreads();
    if (ok) {
        const float ssr = ldexpf(sr, log2S), ssi = ldexpf(si, log2S);
        if (!m.one_d) {
            const CornerSet<float> cs = corner_set<float>(c, t, lane & 7);
            const float zr[2] = {cs.z[0] * ssr, cs.z[1] * ssr};
            const float zi[2] = {cs.z[0] * ssi, cs.z[1] * ssi};
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const float vr = fmaf(cs.xy[q & 3], zr[q >> 2], PACK_MAGIC);
                const float vi = fmaf(cs.xy[q & 3], zi[q >> 2], PACK_MAGIC);
                const unsigned long long wq =
                    (((unsigned long long)__float_as_uint(vr) << 32) | __float_as_uint(vi)) - PACK_MAGIC2;
                atomicAdd(&s_tile[cr.t0 + cs.oxy[q & 3] + cs.oz[q >> 2]], wq);
            }
        } else {
            for (int q = 0; q < nnb; ++q) {
                const int k = (q + lane) & (nnb - 1);
                const float wk = corner_weight<float>(m, c, k);
                const long long qr = (long long)(int)rintf(wk * ssr);
                const long long qi = (long long)(int)rintf(wk * ssi);
                atomicAdd(&s_tile[corner_tile(m, t, cr, k)], (unsigned long long)(qr * 4294967296LL + qi));
            }
        }
    }
    __syncthreads();
    const int ntr = m.nx * m.ny;
    const BoxDiv bd = box_div(t);
    const float inv = ldexpf(1.0f, -log2S);
    float* outc = reinterpret_cast<float*>(out);
    // lanes 2j and 2j + 1: real and imaginary part of node j (as in
    // scatter_cells: one atomic instruction per 32 nodes' both parts)
    for (int e = threadIdx.x; e < 2 * nt; e += blockDim.x) {
        const int q = e >> 1, part = e & 1;
        const long long w = (long long)s_tile[q];
        const long long lo = (long long)(int)(unsigned int)(w & 0xffffffffLL);
        const long long v = part ? lo : (w - lo) >> 32;   // exact: w - lo is a multiple of 2^32
        if (v != 0) {
            int lx, ly, lz;
            box_coords(bd, q, lx, ly, lz);
            const int node = (t.bz0 + lz) * ntr + (t.by0 + ly) * m.nx + (t.bx0 + lx);
            atomic_add(outc + 2 * node + part, (float)v * inv);
        }
    }
}

// ---- deterministic deposit (FP64 form; see DetState) -------------------------

// Fixed-point pair of one contribution: t = v 2^F is exact (a power-of-two
// scaling) and so is t - rint(t) (Sterbenz: |t - rint(t)| <= 1/2 with
// rint(t) within a factor 2 of t, or rint(t) = 0), so hi + lo 2^-DET_LO_BITS
// is v 2^F rounded once, at 2^-DET_LO_BITS.  Two's-complement words: integer
// atomics add them modulo 2^64, which is exact while the sums stay in range.
__device__ __forceinline__ void det_split(double v, double scale, unsigned long long& hi,
                                          unsigned long long& lo) {
    const double t = v * scale;
    const double r = rint(t);
    hi = (unsigned long long)(long long)r;
    lo = (unsigned long long)(long long)rint((t - r) * (double)(1ull << DET_LO_BITS));
}

// out[node_k] += w_k (sr, si) in fixed point for every lane with a cell.
// The same two paths as scatter_cells (LDS tile over the workgroup's node
// box, or runs of lanes that share a cell), with integer adds throughout, so
// neither the order of the LDS atomics nor that of the global ones changes a
// bit.  s_tile: DET_WORDS x nt words (Re hi | Re lo | Im hi | Im lo); out:
// DET_WORDS words per mesh node.
__device__ __forceinline__ void scatter_cells_det(const MeshGeom& m, const TileBox& t, bool use_tile,
                                                  bool ok, const Cell<double>& c, const CellRef& cr,
                                                  double sr, double si, double scale,
                                                  unsigned long long* out, unsigned long long* s_tile) {
    const int lane = threadIdx.x & 63;
    const int nnb = m.one_d ? 2 : 8;
    if (use_tile) {
        const int nt = t.tx * t.ty * t.tz;
        for (int q = threadIdx.x; q < DET_WORDS * nt; q += blockDim.x) s_tile[q] = 0ull;
        __syncthreads();
        if (ok) {
            auto add = [&](double wk, int tk) {
                unsigned long long h, l;
                det_split(wk * sr, scale, h, l);
                atomicAdd(&s_tile[tk], h);
                atomicAdd(&s_tile[nt + tk], l);
                det_split(wk * si, scale, h, l);
                atomicAdd(&s_tile[2 * nt + tk], h);
                atomicAdd(&s_tile[3 * nt + tk], l);
            };
            if (!m.one_d) {
                const CornerSet<double> cs = corner_set<double>(c, t, lane & 7);
#pragma unroll
                for (int q = 0; q < 8; ++q) add(cs.xy[q & 3] * cs.z[q >> 2], cr.t0 + cs.oxy[q & 3] + cs.oz[q >> 2]);
            } else {
                for (int q = 0; q < nnb; ++q) {
                    const int k = (q + lane) & (nnb - 1);
                    add(corner_weight<double>(m, c, k), corner_tile(m, t, cr, k));
                }
            }
        }
        __syncthreads();
        const int64_t ntr = (int64_t)m.nx * m.ny;
        const BoxDiv bd = box_div(t);
        // lanes DET_WORDS j .. DET_WORDS j + 3 take the words of node j (one
        // 32-byte sector): one atomic instruction covers 16 nodes
        for (int e = threadIdx.x; e < DET_WORDS * nt; e += blockDim.x) {
            const int q = e / DET_WORDS, j = e % DET_WORDS;
            const unsigned long long w = s_tile[j * nt + q];
            if (w) {
                int lx, ly, lz;
                box_coords(bd, q, lx, ly, lz);
                const int64_t node = (int64_t)(t.bz0 + lz) * ntr + (int64_t)(t.by0 + ly) * m.nx + (t.bx0 + lx);
                atomicAdd(out + DET_WORDS * node + j, w);
            }
        }
    } else {
        const int prev = __shfl_up(cr.key, 1, 64);
        const bool head = (lane == 0) || (prev != cr.key);
        const unsigned long long heads = __ballot(head);
        const unsigned long long after = (lane == 63) ? 0ull : (heads >> (lane + 1));
        const int seg_end = after ? lane + __builtin_ctzll(after) : 63;
        const int maxlen = wave_max(head ? seg_end - lane + 1 : 1);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (k < nnb) {
                const double wk = ok ? corner_weight<double>(m, c, k) : 0.0;
                unsigned long long w[DET_WORDS];
                det_split(wk * sr, scale, w[0], w[1]);
                det_split(wk * si, scale, w[2], w[3]);
                for (int o = 1; o < maxlen; o <<= 1) {
#pragma unroll
                    for (int j = 0; j < DET_WORDS; ++j) {
                        const unsigned long long u = __shfl_down(w[j], o, 64);
                        if (lane + o <= seg_end) w[j] += u;
                    }
                }
                if (ok && head) {
                    unsigned long long* o = out + DET_WORDS * (int64_t)corner_node(m, cr, k);
#pragma unroll
                    for (int j = 0; j < DET_WORDS; ++j)
                        if (w[j]) atomicAdd(o + j, w[j]);
                }
            }
        }
    }
}

}  // namespace puff
