// collide.hip - per-substep kernels of mesh colliders: the mesh interpolated
// to the substep, and the meld of the twin nodes.
#include <algorithm>

#include "collide.hpp"
#include "device_math.hpp"
#include "internal.hpp"

namespace svamd {

namespace {

// nalgebra's try_normalize(eps): zero (None) when the norm is <= eps.
__device__ __forceinline__ float3 try_normalize3(float3 v, float eps) {
    const float n = len3(v);
    return n <= eps ? make_float3(0.f, 0.f, 0.f) : v * (1.0f / n);
}

// interpolate_input's mesh part (rust/crates/cpu/src/phase/interpolate_input.rs:37-56,
// 82-94) for the substep of parameter entry ds->param_index: vertex positions
// factor_a a + factor_b b (f32, as the reference), triangle normals from the
// interpolated vertices, frictions and dampings.
__global__ __launch_bounds__(256) void k_mesh_interp(MeshDev m, const SubstepParams *params, DeviceScalars *ds) {
    if (ds->halt) return;
    const float fb = params[ds->param_index].factor_b;
    const float fa = 1.0f - fb;
    auto pos = [&](uint32_t v) {
        return make_float3(fa * m.pos_a[3 * v] + fb * m.pos_b[3 * v], fa * m.pos_a[3 * v + 1] + fb * m.pos_b[3 * v + 1],
                           fa * m.pos_a[3 * v + 2] + fb * m.pos_b[3 * v + 2]);
    };
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m.nvert + m.ntri; i += stride) {
        if (i < m.nvert) {
            const float3 p = pos(i);
            m.vpos[i] = make_float4(p.x, p.y, p.z, 0.f);
        } else {
            const uint32_t t = i - m.nvert;
            const float3 a = pos(m.tri[3 * t]), b = pos(m.tri[3 * t + 1]), c = pos(m.tri[3 * t + 2]);
            const float3 n = try_normalize3(cross3(b - a, c - a), COLLIDE_EPS);
            m.tnorm[t] = make_float4(n.x, n.y, n.z, 0.f);
            m.tfric[t] = fa * m.fric_a[t] + fb * m.fric_b[t];
            m.tdamp[t] = fa * m.damp_a[t] + fb * m.damp_b[t];
        }
    }
}

// nalgebra's Vector::angle: the clamped arc cosine of the normalised dot
// product, 0 for a zero vector.
__device__ __forceinline__ float angle3(float3 a, float3 b) {
    const float n1 = len3(a), n2 = len3(b);
    if (n1 == 0.f || n2 == 0.f) return 0.f;
    return acosf(fminf(fmaxf(dot3(a, b) / (n1 * n2), -1.f), 1.f));
}

// Vertex normals: the angle-weighted sum of the normals of the vertex's
// triangles, normalised (interpolate_input.rs:57-81).
__global__ __launch_bounds__(256) void k_vertex_normals(MeshDev m, const DeviceScalars *ds) {
    if (ds->halt) return;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < m.nvert; v += gridDim.x * blockDim.x) {
        const float3 p = f3(m.vpos[v]);
        float3 sum = make_float3(0.f, 0.f, 0.f);
        for (uint32_t k = m.vtx_off[v]; k < m.vtx_off[v + 1]; ++k) {
            const uint32_t t = m.vtx_tri[k];
            uint32_t o[2], no = 0;
            for (int e = 0; e < 3; ++e)
                if (m.tri[3 * t + e] != v && no < 2) o[no++] = m.tri[3 * t + e];
            if (no < 2) continue;  // a degenerate triangle naming the vertex twice
            sum = sum + f3(m.tnorm[t]) * angle3(f3(m.vpos[o[0]]) - p, f3(m.vpos[o[1]]) - p);
        }
        const float3 n = try_normalize3(sum, COLLIDE_EPS);
        m.vnorm[v] = make_float4(n.x, n.y, n.z, 0.f);
    }
}

__device__ __forceinline__ bool compatible(uint32_t a, uint32_t b) {
    return (((a & b) >> 16) & (a ^ b)) == 0u;  // collider_bits::compatible, collider_bits.rs:34-38
}

// The twins of the substep that was just scattered (table of the epoch's
// parity): each twin's mass and momentum are the node's total (the plain
// grid, before the grid update divides and clears it) minus the twins
// incompatible with it (meld_grid.rs:41-59); then v = p / m and the grid
// options.  The other table, read by the gather that has run, is cleared.
__global__ __launch_bounds__(256) void k_twin_meld(TwinTables t, const float4 *__restrict__ grid, GridGeom G,
                                                   const SubstepParams *params, DeviceScalars *ds) {
    if (ds->halt) return;
    const SubstepParams &S = params[ds->grid_param];
    TwinEntry *cur = t.tab[S.epoch & 1u];
    TwinEntry *other = t.tab[(S.epoch + 1u) & 1u];
    const uint32_t cap = t.mask + 1u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) {
        TwinEntry &o = other[i];
        o.node = TWIN_EMPTY;
#pragma unroll
        for (int s = 0; s < TWIN_SLOTS; ++s) {
            o.bits[s] = 0u;
            o.mp[s] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        TwinEntry &e = cur[i];
        const uint32_t node = e.node;
        if (node == TWIN_EMPTY) continue;
        const float4 total = grid[node];
        uint32_t bits[TWIN_SLOTS];
        float4 mp[TWIN_SLOTS];
#pragma unroll
        for (int s = 0; s < TWIN_SLOTS; ++s) {
            bits[s] = e.bits[s];
            mp[s] = e.mp[s];
        }
        const uint32_t blk = node >> 6, loc = node & 63u;
        const int lx = 4 * (int)(blk % (uint32_t)G.nb[0]) + (int)(loc & 3u);
        const int ly = 4 * (int)((blk / (uint32_t)G.nb[0]) % (uint32_t)G.nb[1]) + (int)((loc >> 2) & 3u);
        const int lz = 4 * (int)(blk / ((uint32_t)G.nb[0] * (uint32_t)G.nb[1])) + (int)(loc >> 4);
#pragma unroll
        for (int s = 0; s < TWIN_SLOTS; ++s) {
            if (!bits[s]) continue;
            float m = total.x, p[3] = {total.y, total.z, total.w};
#pragma unroll
            for (int j = 0; j < TWIN_SLOTS; ++j) {
                if (j == s || !bits[j] || compatible(bits[s], bits[j])) continue;
                m -= mp[j].x;
                p[0] -= mp[j].y;
                p[1] -= mp[j].z;
                p[2] -= mp[j].w;
            }
            float v[3] = {0.f, 0.f, 0.f};
            if (m > 0.f) {
#pragma unroll
                for (int q = 0; q < 3; ++q) v[q] = p[q] / m;
                grid_options(S, G, lx, ly, lz, v);
            }
            e.mp[s] = make_float4(m, v[0], v[1], v[2]);
        }
    }
}

__global__ __launch_bounds__(256) void k_twin_clear(TwinEntry *tab, uint32_t cap) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) {
        tab[i].node = TWIN_EMPTY;
#pragma unroll
        for (int s = 0; s < TWIN_SLOTS; ++s) {
            tab[i].bits[s] = 0u;
            tab[i].mp[s] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
}

uint32_t groups_for(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 4096)); }

}  // namespace

hipError_t launch_mesh_interp(const MeshDev &m, const SubstepParams *params, DeviceScalars *ds, hipStream_t stream) {
    if (m.ntri == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mesh_interp, dim3(groups_for((uint64_t)m.nvert + m.ntri)), dim3(256), 0, stream, m, params, ds);
    hipLaunchKernelGGL(k_vertex_normals, dim3(groups_for(m.nvert)), dim3(256), 0, stream, m, ds);
    return hipGetLastError();
}

hipError_t launch_twin_meld(const TwinTables &t, const float4 *grid, const GridGeom &G, const SubstepParams *params,
                            DeviceScalars *ds, hipStream_t stream) {
    hipLaunchKernelGGL(k_twin_meld, dim3(groups_for((uint64_t)t.mask + 1)), dim3(256), 0, stream, t, grid, G, params, ds);
    return hipGetLastError();
}

hipError_t launch_twin_clear(const TwinTables &t, hipStream_t stream) {
    for (int k = 0; k < 2; ++k)
        hipLaunchKernelGGL(k_twin_clear, dim3(groups_for((uint64_t)t.mask + 1)), dim3(256), 0, stream, t.tab[k],
                           t.mask + 1);
    return hipGetLastError();
}

}  // namespace svamd
