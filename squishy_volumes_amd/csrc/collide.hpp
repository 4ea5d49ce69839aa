// collide.hpp - the collide phase and the twin nodes of mesh colliders, as
// device functions for the transfer kernel.
//
// collide_particle restates CpuState::collide
// (rust/crates/cpu/src/phase/collide.rs:21-206) for one particle: the closest
// triangle of each collider within the forget distance sets the particle's
// collider bits (near and side, rust/crates/util/src/collider_bits.rs:9-32),
// and a particle found on the other side of a collider than its bits
// remember gets friction, damping and a push back to the surface.
//
// Twin nodes: the reference keys grid nodes by (node id, collider bits)
// (GridKey, rust/crates/cpu/src/grid_nodes.rs:15-19) and melds each key with
// the compatible keys at its node (meld_grid.rs:16-69).  Here every particle
// scatters into the plain grid, so that a node's plain entry holds the sum
// over all its keys; a particle near a collider also scatters into its twin
// (node, bits) in a hash table.  A key with no near bit is compatible with
// every key and melds to the node's total: the plain grid's value.  A twin
// melds to the total minus the twins incompatible with it.
#pragma once

#include "internal.hpp"

namespace svamd {

constexpr float COLLIDE_EPS = 1e-5f;  // NORMALIZATION_EPS, rust/crates/util/src/consts.rs:11

__device__ __forceinline__ float3 f3(const float4 v) { return make_float3(v.x, v.y, v.z); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 operator-(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 operator*(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float len3(float3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ bool is_zero3(float3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }

// distance_to_segment, rust/crates/mesh_util/src/mesh.rs:264-275.
__device__ __forceinline__ float segment_distance(float3 p, float3 s, float3 e) {
    const float3 seg = e - s;
    const float along = dot3(p - s, seg) / dot3(seg, seg);
    if (along < 0.f) return len3(p - s);
    if (along < 1.f) return len3(p - s - seg * along);
    return len3(p - e);
}

// distance_to_triangle, mesh.rs:277-309.
__device__ __forceinline__ float triangle_distance(float3 p, float3 a, float3 b, float3 c, float3 n) {
    const float3 ab = a - b, bc = b - c, ca = c - a;
    const bool sa = dot3(n, cross3(bc, c - p)) > 0.f;
    const bool sb = dot3(n, cross3(ca, a - p)) > 0.f;
    const bool sc = dot3(n, cross3(ab, b - p)) > 0.f;
    if (sa && sb && sc) return fabsf(dot3(p - a, n));
    float d = 3.402823466e38f;
    if (!sa) d = fminf(d, segment_distance(p, b, c));
    if (!sb) d = fminf(d, segment_distance(p, c, a));
    if (!sc) d = fminf(d, segment_distance(p, a, b));
    return d;
}

struct SegmentResult {
    float distance;
    float3 to_p, normal;
};

// segment_distance_result, mesh.rs:234-262.
__device__ __forceinline__ SegmentResult segment_result(float3 p, float3 s, float3 e, float3 sn, float3 segn,
                                                        float3 en) {
    const float3 seg = e - s;
    const float along = dot3(p - s, seg) / dot3(seg, seg);
    if (along < 0.f) return {len3(p - s), p - s, sn};
    if (along < 1.f) {
        const float3 d = p - s - seg * along;
        return {len3(d), d, segn};
    }
    return {len3(p - e), p - e, en};
}

// The block of the dense grid holding position x (its node-space cell), or
// all ones outside the grid.
__device__ __forceinline__ uint32_t block_of_position(const float x[3], const GridGeom &G) {
    uint32_t b[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float fl = floorf(div_by(x[a], G.h, G.inv_h));
        if (!(fl > -1.0e9f && fl < 1.0e9f)) return 0xffffffffu;
        const int c = (int)fl - G.origin[a];
        if (c < 0 || (c >> 2) >= G.nb[a]) return 0xffffffffu;
        b[a] = (uint32_t)(c >> 2);
    }
    return b[0] + (uint32_t)G.nb[0] * (b[1] + (uint32_t)G.nb[1] * b[2]);
}

// The collide phase for one live particle at x with velocity v and collider
// bits `bits` (both updated).  Candidates are the triangles listed for the
// particle's block, ascending; per collider the first triangle of the
// smallest distance wins, as in the reference's scan.
// (Inlined: as a call it made the transfer kernel spill 736 bytes per lane
// around it and run at one wave per SIMD.)
__device__ __forceinline__ void collide_particle(const MeshDev &M, const GridGeom &G, float dt, const float xp[3],
                                                 float v[3], uint32_t &bits) {
    const uint32_t B = block_of_position(xp, G);
    const uint32_t k0 = B == 0xffffffffu ? 0u : M.blk_off[B];
    const uint32_t k1 = B == 0xffffffffu ? 0u : M.blk_off[B + 1];
    if (k0 == k1) {
        bits = 0u;
        return;
    }
    const float3 p = make_float3(xp[0], xp[1], xp[2]);
    float3 vel = make_float3(v[0], v[1], v[2]);
    for (uint32_t col = 0; col < 16u; ++col) {
        uint32_t best = 0xffffffffu;
        if (col < M.ncolliders) {
            float best_d = 3.402823466e38f;
            for (uint32_t k = k0; k < k1; ++k) {
                const uint32_t t = M.blk_tri[k];
                if (M.tri_coll[t] != col) continue;
                const float3 n = f3(M.tnorm[t]);
                if (is_zero3(n)) continue;
                const float d = triangle_distance(p, f3(M.vpos[M.tri[3 * t]]), f3(M.vpos[M.tri[3 * t + 1]]),
                                                  f3(M.vpos[M.tri[3 * t + 2]]), n);
                if (d >= M.forget) continue;
                if (d < best_d) {
                    best_d = d;
                    best = t;
                }
            }
        }
        const uint32_t both = 0x00010001u << col;
        if (best == 0xffffffffu) {
            bits &= ~both;  // set(None)
            continue;
        }
        const uint32_t ia = M.tri[3 * best], ib = M.tri[3 * best + 1], ic = M.tri[3 * best + 2];
        const float3 n = f3(M.tnorm[best]);
        const float3 a = f3(M.vpos[ia]), b = f3(M.vpos[ib]), c = f3(M.vpos[ic]);
        const uint32_t oab = M.opp[3 * best], obc = M.opp[3 * best + 1], oca = M.opp[3 * best + 2];
        const float3 zero = make_float3(0.f, 0.f, 0.f);
        const float3 ab_n = oab != 0xffffffffu ? n + f3(M.tnorm[oab]) : zero;
        const float3 bc_n = obc != 0xffffffffu ? n + f3(M.tnorm[obc]) : zero;
        const float3 ca_n = oca != 0xffffffffu ? n + f3(M.tnorm[oca]) : zero;
        const float3 ab = a - b, bc = b - c, ca = c - a;
        const float area2 = dot3(n, cross3(ca, ab));
        const float a_bary = dot3(n, cross3(bc, c - p)) / area2;
        const float b_bary = dot3(n, cross3(ca, a - p)) / area2;
        const float c_bary = dot3(n, cross3(ab, b - p)) / area2;
        SegmentResult r;
        if (a_bary > 0.f && b_bary > 0.f && c_bary > 0.f) {
            const float s = dot3(p - a, n);
            r = {fabsf(s), n * s, n};
        } else {
            const float3 an = f3(M.vnorm[ia]), bn = f3(M.vnorm[ib]), cn = f3(M.vnorm[ic]);
            r = segment_result(p, a, b, an, ab_n, bn);
            const SegmentResult r1 = segment_result(p, b, c, bn, bc_n, cn);
            if (r1.distance < r.distance) r = r1;
            const SegmentResult r2 = segment_result(p, c, a, cn, ca_n, an);
            if (r2.distance < r.distance) r = r2;
        }
        if (is_zero3(r.normal)) {
            bits &= ~both;
            continue;
        }
        const bool new_side = 0.f <= dot3(r.to_p, r.normal);
        const uint32_t near = 0x00010000u << col;
        if (!(bits & near)) {
            if (r.distance < M.accept) bits = (bits & ~both) | (new_side ? both : near);
            continue;
        }
        if (((bits & (1u << col)) != 0u) == new_side) continue;
        if (r.distance > COLLIDE_EPS) {
            const float3 cv = f3(M.vvel[ia]) * a_bary + f3(M.vvel[ib]) * b_bary + f3(M.vvel[ic]) * c_bary;
            const float3 rel = vel - cv;
            const float3 cn = r.to_p * (1.0f / r.distance);
            const float3 vn = cn * dot3(rel, cn);
            const float3 vt = rel - vn;
            const float tn = len3(vt);
            if (tn > COLLIDE_EPS) {
                const float3 tangent = vt * (1.0f / tn);
                vel = vel - tangent * fminf(M.tfric[best] * r.distance / dt, tn);
            }
            vel = vel - vn * fminf(M.tdamp[best], 1.f);
        }
        vel = vel - r.to_p * (1.0f / dt);
    }
    v[0] = vel.x;
    v[1] = vel.y;
    v[2] = vel.z;
}

__device__ __forceinline__ uint32_t twin_hash(uint32_t node) {
    uint32_t h = node * 0x9E3779B1u;
    return h ^ (h >> 15);
}

// The accumulator of twin (node, bits), inserted when missing; null (and
// TABLE_TRIES_EXCEEDED) when the table or the node's slots are full.  A key
// once set stays for the substep, so a plain load that already shows this
// node (or these bits) settles the probe without an atomic; one that shows an
// empty key (perhaps stale) falls back to the compare-and-swap.
__device__ __forceinline__ float4 *twin_slot(TwinEntry *tab, uint32_t mask, uint32_t node, uint32_t bits,
                                             DeviceScalars *ds) {
    const uint32_t h = twin_hash(node);
    for (int probe = 0; probe < TWIN_PROBES; ++probe) {
        TwinEntry &e = tab[(h + (uint32_t)probe) & mask];
        uint32_t old = e.node;
        if (old == TWIN_EMPTY) old = atomicCAS(&e.node, TWIN_EMPTY, node);
        if (old != TWIN_EMPTY && old != node) continue;
#pragma unroll
        for (int s = 0; s < TWIN_SLOTS; ++s) {
            uint32_t ob = e.bits[s];
            if (ob == 0u) ob = atomicCAS(&e.bits[s], 0u, bits);
            if (ob == 0u || ob == bits) return &e.mp[s];
        }
        break;
    }
    atomicOr(&ds->status, MPM_STATUS_TABLE_TRIES_EXCEEDED | (KID_P2G << 16));
    return nullptr;
}

// The melded velocity of twin (node, bits) (TABLE_ENTRY_MISSING and zero when
// absent).
__device__ __forceinline__ float4 twin_velocity(const TwinEntry *tab, uint32_t mask, uint32_t node, uint32_t bits,
                                                DeviceScalars *ds) {
    const uint32_t h = twin_hash(node);
    for (int probe = 0; probe < TWIN_PROBES; ++probe) {
        const TwinEntry &e = tab[(h + (uint32_t)probe) & mask];
        const uint32_t en = e.node;
        if (en == TWIN_EMPTY) break;
        if (en != node) continue;
#pragma unroll
        for (int s = 0; s < TWIN_SLOTS; ++s)
            if (e.bits[s] == bits) return e.mp[s];
        break;
    }
    atomicOr(&ds->status, MPM_STATUS_TABLE_ENTRY_MISSING | (KID_G2P << 16));
    return make_float4(0.f, 0.f, 0.f, 0.f);
}

// Scatter of one payload (the transfer kernel's 16 floats: f, m | q | h A
// columns) into the twins (node, bits) of its 27 stencil nodes, cell = the
// global cell of the stencil's first node.
__device__ __forceinline__ void twin_scatter(TwinEntry *tab, uint32_t mask, const GridGeom &G, DeviceScalars *ds,
                                             const int cell[3], const float pl[16], uint32_t bits) {
    if (cell[0] < 0 || cell[1] < 0 || cell[2] < 0 || cell[0] + 2 >= 4 * G.nb[0] || cell[1] + 2 >= 4 * G.nb[1] ||
        cell[2] + 2 >= 4 * G.nb[2])
        return;  // the plain scatter reports OUT_OF_GRID
    float wx[3], wy[3], wz[3];
    bspline_weights(pl[0], wx);
    bspline_weights(pl[1], wy);
    bspline_weights(pl[2], wz);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                const float wt = wx[i] * wy[j] * wz[k];
                const uint32_t x = (uint32_t)(cell[0] + i), y = (uint32_t)(cell[1] + j), z = (uint32_t)(cell[2] + k);
                const uint32_t node = (x >> 2) + (uint32_t)G.nb[0] * ((y >> 2) + (uint32_t)G.nb[1] * (z >> 2));
                float4 *acc = twin_slot(tab, mask, node * 64u + (x & 3u) + ((y & 3u) << 2) + ((z & 3u) << 4), bits, ds);
                if (!acc) return;
                float *g = reinterpret_cast<float *>(acc);
                unsafeAtomicAdd(g + 0, wt * pl[3]);
#pragma unroll
                for (int r = 0; r < 3; ++r)
                    unsafeAtomicAdd(g + 1 + r, wt * (pl[4 + r] + (float)i * pl[7 + r] + (float)j * pl[10 + r] +
                                                     (float)k * pl[13 + r]));
            }
}

}  // namespace svamd
