// mpm_oracle.hpp - CPU restatement of the squishy_volumes CPU substep.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine: it
// is compiled into oracle/liboracle.so, which only the tests, the smoke check
// and the CPU-baseline leg of the benchmark load.  The engine never links it.
//
// It restates, phase by phase, CpuState::produce_next_state
// (rust/crates/cpu/src/cpu_state.rs:146-197) with the phase order of
// rust/crates/cpu/src/phase/mod.rs:27-75, templated on the real type so that
// the same code runs in float (the reference's precision) and double (the
// FP64 run the engine is scored against).  Each function cites the reference
// lines it follows.  Extensions without a reference counterpart (fixed-
// corotated and snow materials, grid gravity, grid boundary conditions) are
// marked as such and pinned by closed forms in the tests.
//
// Deviations from the reference that only change summation order:
//  * the sort is stable (the reference uses par_sort_unstable_by_key);
//  * grid nodes are indexed through a dense box instead of an FxHashMap, and a
//    node's contributors are visited in particle order (the reference pushes
//    them under a Mutex in scheduling order).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <vector>

#include <omp.h>

#include "../include/mpm_engine.h"

namespace mpm_oracle {

// ---------------------------------------------------------------- math types
template <class R>
struct V3 {
    R x = 0, y = 0, z = 0;
    V3() = default;
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    R &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    V3 operator+(const V3 &o) const { return {x + o.x, y + o.y, z + o.z}; }
    V3 operator-(const V3 &o) const { return {x - o.x, y - o.y, z - o.z}; }
    V3 operator*(R s) const { return {x * s, y * s, z * s}; }
    V3 operator/(R s) const { return {x / s, y / s, z / s}; }
    V3 &operator+=(const V3 &o) { x += o.x; y += o.y; z += o.z; return *this; }
    V3 &operator-=(const V3 &o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
    V3 &operator*=(R s) { x *= s; y *= s; z *= s; return *this; }
};
template <class R> V3<R> operator*(R s, const V3<R> &v) { return v * s; }
template <class R> V3<R> cross(const V3<R> &a, const V3<R> &b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> R dot(const V3<R> &a, const V3<R> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// 3x3 matrix, m[row][col].
template <class R>
struct M3 {
    R m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    static M3 identity() { M3 r; r.m[0][0] = r.m[1][1] = r.m[2][2] = 1; return r; }
    static M3 diag(R a, R b, R c) { M3 r; r.m[0][0] = a; r.m[1][1] = b; r.m[2][2] = c; return r; }
    V3<R> col(int j) const { return {m[0][j], m[1][j], m[2][j]}; }
    void set_col(int j, const V3<R> &v) { m[0][j] = v.x; m[1][j] = v.y; m[2][j] = v.z; }
    M3 operator+(const M3 &o) const { M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = m[i][j] + o.m[i][j]; return r; }
    M3 operator-(const M3 &o) const { M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = m[i][j] - o.m[i][j]; return r; }
    M3 operator*(R s) const { M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = m[i][j] * s; return r; }
    M3 &operator+=(const M3 &o) { for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] += o.m[i][j]; return *this; }
    M3 &operator*=(R s) { for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] *= s; return *this; }
    M3 operator*(const M3 &o) const {
        M3 r;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                R s = 0;
                for (int k = 0; k < 3; ++k) s += m[i][k] * o.m[k][j];
                r.m[i][j] = s;
            }
        return r;
    }
    V3<R> operator*(const V3<R> &v) const {
        return {m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
                m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z};
    }
    M3 transpose() const { M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = m[j][i]; return r; }
    R trace() const { return m[0][0] + m[1][1] + m[2][2]; }
    R determinant() const {
        return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
               m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    }
    R norm_squared() const { R s = 0; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) s += m[i][j] * m[i][j]; return s; }
};
template <class R> M3<R> outer(const V3<R> &a, const V3<R> &b) {
    M3<R> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = a[i] * b[j];
    return r;
}

// ------------------------------------------------- weights (kernels.rs:17-92)
// Quadratic B-spline, rust/crates/cpu/src/kernels.rs:17-26.
template <class R> R kernel_quadratic(R x) {
    x = std::abs(x);
    if (x < R(0.5)) return R(0.75) - x * x;
    if (x < R(1.5)) return R(0.5) * (R(1.5) - x) * (R(1.5) - x);
    return 0;
}

// position_to_shift_quadratic, rust/crates/cpu/src/kernels.rs:46-49.
template <class R> void position_to_shift_quadratic(const V3<R> &p, R h, int s[3]) {
    for (int a = 0; a < 3; ++a) s[a] = (int)std::floor(p[a] / h - R(0.5));
}

// ---------------------------------------------- constitutive math (elastic.rs)
// Lame parameters, rust/crates/util/src/elastic.rs:52-64.
template <class R> R lame_mu(R E, R nu) { return E / R(2) / (R(1) + nu); }
template <class R> R lame_lambda(R E, R nu) { return E * nu / (R(1) + nu) / (R(1) - R(2) * nu); }

// partial_invariant_3_by_position_gradient, elastic.rs:110-118:
// columns f1 x f2, f2 x f0, f0 x f1.
template <class R> M3<R> cofactor(const M3<R> &F) {
    M3<R> c;
    c.set_col(0, cross(F.col(1), F.col(2)));
    c.set_col(1, cross(F.col(2), F.col(0)));
    c.set_col(2, cross(F.col(0), F.col(1)));
    return c;
}

// first_piola_stress_neo_hookean, elastic.rs:299-311 with the partials of
// :282-291: P = (mu/2) * 2F + ((lambda ln J - mu) / J) * cof(F).
template <class R> M3<R> piola_neo_hookean(R mu, R lambda, const M3<R> &F) {
    R J = F.determinant();
    R d3 = (lambda * std::log(J) - mu) / J;
    return F * R(2) * (mu / R(2)) + cofactor(F) * d3;
}

// try_elastic_energy_neo_hookean, elastic.rs:259-272 (+ :246-257).
template <class R> bool energy_neo_hookean(R mu, R lambda, const M3<R> &F, R &out) {
    R J = F.determinant();
    if (!(J > 0)) return false;
    R lnJ = std::log(J);
    out = mu / R(2) * (F.norm_squared() - R(3)) - mu * lnJ + lambda / R(2) * lnJ * lnJ;
    return true;
}

template <class R> R powi(R x, int n) {
    // f32::powi / f64::powi: integer power by repeated multiplication.
    bool neg = n < 0;
    unsigned e = neg ? (unsigned)(-n) : (unsigned)n;
    R r = 1, b = x;
    while (e) { if (e & 1u) r *= b; b *= b; e >>= 1; }
    return neg ? R(1) / r : r;
}

// first_piola_stress_inviscid, elastic.rs:582-592 with :537-544.
template <class R> M3<R> piola_inviscid(R K, int exponent, const M3<R> &F) {
    R J = F.determinant();
    R d3 = K * (R(1) - R(1) / powi(J, exponent));
    return cofactor(F) * d3;
}

// elastic_energy_inviscid_by_invariant, elastic.rs:524-533.
template <class R> R energy_inviscid(R K, int exponent, const M3<R> &F) {
    R J = F.determinant();
    R at_rest = K * (R(1) - R(1) / (R(1) - R(exponent)));
    return K * (J - powi(J, 1 - exponent) / (R(1) - R(exponent))) - at_rest;
}

// cauchy_stress_general_viscosity, elastic.rs:659-677.
template <class R> M3<R> cauchy_viscosity(R dynamic, R bulk, const M3<R> &C) {
    M3<R> strain = (C + C.transpose()) * R(0.5);
    M3<R> r = strain * (R(2) * dynamic);
    R div = C.trace();
    for (int i = 0; i < 3; ++i) r.m[i][i] += bulk * div;
    return r;
}

// ---------------------------------------------------------------- 3x3 SVD
// Signed SVD F = U diag(s) V^T with U, V proper rotations; s[0] >= s[1] >= |s[2]|,
// s[2] < 0 iff det F < 0.  Cyclic Jacobi on F^T F, then a Givens QR of F V,
// written from the textbook method (Golub & Van Loan, 8.5).  The engine's SVD
// follows the same method (in float, with an FP64 fallback); what makes this
// one an independent check is its comparison with numpy.linalg.svd on the
// reference's sampling of F (tests/test_oracle_math.py).
template <class R> void jacobi_rotate(M3<R> &A, M3<R> &V, int p, int q) {
    R apq = A.m[p][q];
    if (apq == 0) return;
    R theta = (A.m[q][q] - A.m[p][p]) / (R(2) * apq);
    R t = (theta >= 0 ? R(1) : R(-1)) / (std::abs(theta) + std::sqrt(theta * theta + R(1)));
    R c = R(1) / std::sqrt(t * t + R(1));
    R s = t * c;
    // A <- G^T A G with G the rotation in the (p, q) plane.
    for (int k = 0; k < 3; ++k) {
        R akp = A.m[k][p], akq = A.m[k][q];
        A.m[k][p] = c * akp - s * akq;
        A.m[k][q] = s * akp + c * akq;
    }
    for (int k = 0; k < 3; ++k) {
        R apk = A.m[p][k], aqk = A.m[q][k];
        A.m[p][k] = c * apk - s * aqk;
        A.m[q][k] = s * apk + c * aqk;
    }
    for (int k = 0; k < 3; ++k) {
        R vkp = V.m[k][p], vkq = V.m[k][q];
        V.m[k][p] = c * vkp - s * vkq;
        V.m[k][q] = s * vkp + c * vkq;
    }
}

template <class R> void svd3(const M3<R> &F, M3<R> &U, V3<R> &s, M3<R> &V) {
    M3<R> A = F.transpose() * F;
    V = M3<R>::identity();
    const int sweeps = sizeof(R) == 8 ? 30 : 12;
    for (int it = 0; it < sweeps; ++it) {
        R off = A.m[0][1] * A.m[0][1] + A.m[0][2] * A.m[0][2] + A.m[1][2] * A.m[1][2];
        R diag = A.m[0][0] * A.m[0][0] + A.m[1][1] * A.m[1][1] + A.m[2][2] * A.m[2][2];
        if (off <= std::numeric_limits<R>::min() || off <= diag * std::numeric_limits<R>::epsilon() *
                                                                std::numeric_limits<R>::epsilon())
            break;
        jacobi_rotate(A, V, 0, 1);
        jacobi_rotate(A, V, 0, 2);
        jacobi_rotate(A, V, 1, 2);
    }
    if (V.determinant() < 0) V.set_col(2, V.col(2) * R(-1));
    M3<R> B = F * V;
    // Sort columns of B by decreasing norm, keeping V a rotation.
    R n[3];
    for (int j = 0; j < 3; ++j) n[j] = dot(B.col(j), B.col(j));
    auto swap_cols = [&](int a, int b) {
        V3<R> t = B.col(a); B.set_col(a, B.col(b)); B.set_col(b, t * R(-1));
        t = V.col(a); V.set_col(a, V.col(b)); V.set_col(b, t * R(-1));
        std::swap(n[a], n[b]);
    };
    if (n[0] < n[1]) swap_cols(0, 1);
    if (n[0] < n[2]) swap_cols(0, 2);
    if (n[1] < n[2]) swap_cols(1, 2);
    // Givens QR: B = Q R, Q a rotation, R upper triangular.
    U = M3<R>::identity();
    auto givens = [&](int i, int j) {  // zero B(j, i) using rows i and j
        R a = B.m[i][i], b = B.m[j][i];
        R r = std::sqrt(a * a + b * b);
        R c = r > 0 ? a / r : R(1), sn = r > 0 ? b / r : R(0);
        for (int k = 0; k < 3; ++k) {
            R bi = B.m[i][k], bj = B.m[j][k];
            B.m[i][k] = c * bi + sn * bj;
            B.m[j][k] = -sn * bi + c * bj;
        }
        for (int k = 0; k < 3; ++k) {  // U <- U G^T
            R ui = U.m[k][i], uj = U.m[k][j];
            U.m[k][i] = c * ui + sn * uj;
            U.m[k][j] = -sn * ui + c * uj;
        }
    };
    givens(0, 1);
    givens(0, 2);
    givens(1, 2);
    s = {B.m[0][0], B.m[1][1], B.m[2][2]};
    // The first two diagonal entries are made non-negative by rotating pairs.
    for (int j = 0; j < 2; ++j)
        if (s[j] < 0) {
            s[j] = -s[j];
            s[2] = -s[2];
            U.set_col(j, U.col(j) * R(-1));
            U.set_col(2, U.col(2) * R(-1));
        }
}

// ---- plasticity of the advance (advance_particles.rs:43-86) ------------------
// nalgebra's SVD has non-negative singular values; for det F < 0 its U V^T is
// a reflection.  In terms of the signed SVD: |s|, and the sign g = sgn s[2]
// folded back as diag(1, 1, g).  Both projections are functions of the
// singular values, so the result does not depend on which SVD was taken.

// Drucker-Prager return map in log strain (advance_particles.rs:50-71).
template <class R> M3<R> sand_project(const M3<R> &F, R mu, R lambda, R alpha) {
    M3<R> U, V;
    V3<R> s;
    svd3(F, U, s, V);
    const R g = s[2] < 0 ? R(-1) : R(1);
    V3<R> e(std::log(std::abs(s[0])), std::log(std::abs(s[1])), std::log(std::abs(s[2])));
    R e_tr = e.x + e.y + e.z;
    V3<R> e_hat = e - V3<R>(e_tr / 3, e_tr / 3, e_tr / 3);
    R e_hat_norm = std::sqrt(dot(e_hat, e_hat));
    if (e_tr < 0 && e_hat_norm > 0) {
        R delta_gamma = e_hat_norm + (R(3) * lambda + R(2) * mu) / R(2) / mu * e_tr * alpha;
        if (delta_gamma > 0) {
            V3<R> H = e - e_hat * (delta_gamma / e_hat_norm);
            return U * M3<R>::diag(std::exp(H.x), std::exp(H.y), g * std::exp(H.z)) * V.transpose();
        }
        return F;
    }
    return U * M3<R>::diag(R(1), R(1), g) * V.transpose();
}

// Fluid: every singular value <- (s1 s2 s3)^(1/3) (advance_particles.rs:76-86),
// i.e. |det F|^(1/3) times the orthogonal polar factor of F.
template <class R> M3<R> fluid_project(const M3<R> &F) {
    M3<R> U, V;
    V3<R> s;
    svd3(F, U, s, V);
    const R g = s[2] < 0 ? R(-1) : R(1);
    const R c = std::cbrt(std::abs(s[0] * s[1] * s[2]));
    return U * M3<R>::diag(c, c, g * c) * V.transpose();
}

// Extension (no reference counterpart): separating slip on a sphere.  A node
// at xi inside the sphere loses the velocity component along the inward
// normal; its tangential component and any outward motion stay.
template <class R> V3<R> sphere_slip(const V3<R> &xi, const V3<R> &center, R radius, V3<R> v) {
    const V3<R> d = xi - center;
    const R r2 = d.x * d.x + d.y * d.y + d.z * d.z;
    if (r2 < radius * radius && r2 > 0) {
        const R vn = (v.x * d.x + v.y * d.y + v.z * d.z) / r2;
        if (vn < 0) v = v - d * vn;
    }
    return v;
}

// ---- mesh geometry (rust/crates/mesh_util/src/mesh.rs:228-309) -------------
// NORMALIZATION_EPS, rust/crates/util/src/consts.rs:11.
constexpr double NORMALIZATION_EPS = 1e-5;

template <class R> R norm3(const V3<R> &v) { return std::sqrt(dot(v, v)); }

// nalgebra's try_normalize(eps): None (here: zero) when |v| <= eps.
template <class R> V3<R> try_normalize(const V3<R> &v, R eps) {
    const R n = norm3(v);
    return n <= eps ? V3<R>() : v / n;
}

// nalgebra's Vector::angle: the clamped arc cosine of the normalised dot
// product, 0 when either vector is zero.
template <class R> R angle_between(const V3<R> &a, const V3<R> &b) {
    const R n1 = norm3(a), n2 = norm3(b);
    if (n1 == R(0) || n2 == R(0)) return R(0);
    R c = dot(a, b) / (n1 * n2);
    c = std::min(std::max(c, R(-1)), R(1));
    return std::acos(c);
}

template <class R> struct DistanceResult {
    R distance;
    V3<R> to_p, normal;
};

// segment_distance_result, mesh.rs:234-262.
template <class R>
DistanceResult<R> segment_distance_result(const V3<R> &p, const V3<R> &start, const V3<R> &end,
                                          const V3<R> &start_normal, const V3<R> &segment_normal,
                                          const V3<R> &end_normal) {
    const V3<R> segment = end - start;
    const R along = dot(p - start, segment) / dot(segment, segment);
    if (along < R(0)) return {norm3(p - start), p - start, start_normal};
    if (along < R(1)) {
        const V3<R> d = p - start - segment * along;
        return {norm3(d), d, segment_normal};
    }
    return {norm3(p - end), p - end, end_normal};
}

// distance_to_segment, mesh.rs:264-275.
template <class R> R distance_to_segment(const V3<R> &p, const V3<R> &start, const V3<R> &end) {
    const V3<R> segment = end - start;
    const R along = dot(p - start, segment) / dot(segment, segment);
    if (along < R(0)) return norm3(p - start);
    if (along < R(1)) return norm3(p - start - segment * along);
    return norm3(p - end);
}

// distance_to_triangle, mesh.rs:277-309: the distance to the plane inside the
// triangle's prism, else to the nearest edge facing the point.
template <class R>
R distance_to_triangle(const V3<R> &p, const V3<R> &a, const V3<R> &b, const V3<R> &c, const V3<R> &n) {
    const V3<R> ab = a - b, bc = b - c, ca = c - a;
    const bool sa = dot(n, cross(bc, c - p)) > R(0);
    const bool sb = dot(n, cross(ca, a - p)) > R(0);
    const bool sc = dot(n, cross(ab, b - p)) > R(0);
    if (sa && sb && sc) return std::abs(dot(p - a, n));
    R distance = std::numeric_limits<R>::max();
    if (!sa) distance = std::min(distance, distance_to_segment(p, b, c));
    if (!sb) distance = std::min(distance, distance_to_segment(p, c, a));
    if (!sc) distance = std::min(distance, distance_to_segment(p, a, b));
    return distance;
}

// collider_bits, rust/crates/util/src/collider_bits.rs:9-32.
inline bool bits_near(uint32_t bits, int c) { return (bits & (0x00010000u << c)) != 0; }
inline bool bits_side(uint32_t bits, int c) { return (bits & (0x00000001u << c)) != 0; }
inline void bits_set(uint32_t &bits, int c, int side /* -1: None, 0: false, 1: true */) {
    bits &= ~(0x00010001u << c);
    if (side == 1) bits |= 0x00010001u << c;
    else if (side == 0) bits |= 0x00010000u << c;
}

// Topology::triangle_opposites (mesh.rs:183-226): per edge (ab, bc, ca) the
// other triangle sharing it, or all ones.
inline std::vector<uint32_t> triangle_opposites(const std::vector<uint32_t> &tri) {
    const size_t nt = tri.size() / 3;
    std::vector<std::pair<uint64_t, uint32_t>> edges;
    edges.reserve(3 * nt);
    auto key = [](uint32_t a, uint32_t b) { return a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a; };
    for (size_t t = 0; t < nt; ++t)
        for (int e = 0; e < 3; ++e) edges.push_back({key(tri[3 * t + e], tri[3 * t + (e + 1) % 3]), (uint32_t)t});
    std::sort(edges.begin(), edges.end());
    std::vector<uint32_t> opp(3 * nt, 0xffffffffu);
    for (size_t t = 0; t < nt; ++t)
        for (int e = 0; e < 3; ++e) {
            const uint64_t k = key(tri[3 * t + e], tri[3 * t + (e + 1) % 3]);
            auto it = std::lower_bound(edges.begin(), edges.end(), std::make_pair(k, 0u));
            for (; it != edges.end() && it->first == k; ++it)
                if (it->second != (uint32_t)t) {
                    opp[3 * t + e] = it->second;
                    break;
                }
        }
    return opp;
}

// ---- extensions: fixed-corotated and snow (no reference counterpart) --------
// Fixed-corotated (Stomakhin et al. 2012): Psi = mu |F - R|^2 + lambda/2 (J-1)^2,
// tau = P F^T = 2 mu (F - R) F^T + lambda (J - 1) J I.
template <class R> M3<R> kirchhoff_fixed_corotated(R mu, R lambda, const M3<R> &F, R &energy) {
    M3<R> U, V;
    V3<R> s;
    svd3(F, U, s, V);
    M3<R> Rot = U * V.transpose();
    R J = s[0] * s[1] * s[2];
    M3<R> tau = (F - Rot) * F.transpose() * (R(2) * mu);
    for (int i = 0; i < 3; ++i) tau.m[i][i] += lambda * (J - R(1)) * J;
    energy = mu * ((s[0] - 1) * (s[0] - 1) + (s[1] - 1) * (s[1] - 1) + (s[2] - 1) * (s[2] - 1)) +
             lambda / R(2) * (J - 1) * (J - 1);
    return tau;
}
template <class R> M3<R> piola_fixed_corotated(R mu, R lambda, const M3<R> &F) {
    M3<R> U, V;
    V3<R> s;
    svd3(F, U, s, V);
    M3<R> Rot = U * V.transpose();
    R J = F.determinant();
    return (F - Rot) * (R(2) * mu) + cofactor(F) * (lambda * (J - R(1)));
}

// Kirchhoff stress tau = P F^T of a particle, the quantity P2G consumes
// (rust/crates/cpu/src/phase/scatter_momentum.rs:52-84).
template <class R>
M3<R> first_piola(uint32_t flags, const mpm_particle_parameters &p, R jp, R hardening, const M3<R> &F) {
    if (flags & MPM_IS_FLUID) return piola_inviscid<R>(p.bulk_modulus, p.exponent, F);
    if (flags & (MPM_FIXED_COROTATED | MPM_SNOW)) {
        R h = (flags & MPM_SNOW) ? std::exp(hardening * (R(1) - jp)) : R(1);
        return piola_fixed_corotated<R>(R(p.mu) * h, R(p.lambda) * h, F);
    }
    return piola_neo_hookean<R>(p.mu, p.lambda, F);
}

// ----------------------------------------------------------------- the state
template <class R>
struct GridKey {
    int id[3];
    uint32_t collider_bits;
};

// collider_bits::compatible, rust/crates/util/src/collider_bits.rs:34-38.
inline bool collider_compatible(uint32_t a, uint32_t b) {
    uint32_t mask = (a & b) >> 16;
    uint32_t diff = a ^ b;
    return (mask & diff) == 0;
}

struct FrameInputCopy {
    uint64_t frame = 0;
    float gravity_a[3] = {0, 0, 0}, gravity_b[3] = {0, 0, 0};
    bool has_b = false;
    std::vector<uint32_t> flags_a, flags_b;
    std::vector<float> goals_a, goals_b;
    // the collider mesh's frames a and b (b = a when absent)
    std::vector<float> vertices_a, vertices_b, frictions_a, frictions_b, dampings_a, dampings_b;
};

// The collider mesh's topology (Topology, rust/crates/mesh_util/src/mesh.rs).
struct MeshTopology {
    std::vector<uint32_t> triangles;           // 3 per triangle
    std::vector<uint32_t> triangle_colliders;
    std::vector<uint32_t> opposites;           // 3 per triangle (ab, bc, ca)
    std::vector<uint32_t> vertex_offsets, vertex_triangles;  // triangles at each vertex, ascending
    uint32_t num_vertices = 0;
    size_t num_triangles() const { return triangle_colliders.size(); }
};

// The topology from triangles and their colliders: opposites, and the
// triangles at each vertex in ascending order (Topology::vertex_triangle_lists).
inline MeshTopology make_topology(uint32_t num_vertices, const uint32_t *triangles, const uint32_t *colliders,
                                  uint32_t num_triangles) {
    MeshTopology m;
    m.num_vertices = num_vertices;
    m.triangles.assign(triangles, triangles + 3 * (size_t)num_triangles);
    m.triangle_colliders.assign(colliders, colliders + num_triangles);
    m.opposites = triangle_opposites(m.triangles);
    m.vertex_offsets.assign(num_vertices + 1, 0);
    for (uint32_t t = 0; t < num_triangles; ++t)
        for (int e = 0; e < 3; ++e) ++m.vertex_offsets[m.triangles[3 * t + e] + 1];
    for (uint32_t v = 0; v < num_vertices; ++v) m.vertex_offsets[v + 1] += m.vertex_offsets[v];
    m.vertex_triangles.assign(m.vertex_offsets[num_vertices], 0);
    std::vector<uint32_t> fill(num_vertices, 0);
    for (uint32_t t = 0; t < num_triangles; ++t)
        for (int e = 0; e < 3; ++e) {
            const uint32_t v = m.triangles[3 * t + e];
            m.vertex_triangles[m.vertex_offsets[v] + fill[v]++] = t;
        }
    return m;
}


// ---- adaptive time steps (rust/crates/cpu/src/phase/limit_time_step.rs:25-223)
constexpr double SINGULAR_VALUE_SEPARATION = 1e-5;  // rust/crates/util/src/consts.rs:14

// limit_time_step_by_speed_of_sound for one particle (:36-130): h / c with
// c^2 = kappa / rho, kappa the largest diagonal-space stiffness over J
// (Neo-Hookean or inviscid, elastic.rs:276-295, :314-339, :555-570, :630-656),
// from nalgebra's singular values (non-negative, descending).
template <class R> R sound_limit(bool fluid, const mpm_particle_parameters &p, const M3<R> &F, R h) {
    M3<R> U, V;
    V3<R> sv;
    svd3(F, U, sv, V);
    R s[3] = {std::abs(sv[0]), std::abs(sv[1]), std::abs(sv[2])};
    std::sort(s, s + 3, [](R a, R b) { return a > b; });
    const R x = s[0], y = s[1], z = s[2];
    const R J = x * y * z;
    const R g[3] = {y * z, x * z, x * y};
    const R Hm[3][3] = {{0, z, y}, {z, 0, x}, {y, x, 0}};
    R d1, d2, diag, first[3], M[3][3];
    if (fluid) {
        const R K = R(p.bulk_modulus);
        const int gamma = p.exponent;
        d1 = K * (R(1) - R(1) / powi(J, gamma));
        d2 = R(gamma) * K / powi(J, gamma + 1);
        diag = 0;
        for (int i = 0; i < 3; ++i) first[i] = d1 * g[i];
    } else {
        const R mu = R(p.mu), lambda = R(p.lambda);
        d1 = (lambda * std::log(J) - mu) / J;
        d2 = (lambda * (R(1) - std::log(J)) + mu) / (J * J);
        diag = mu;
        for (int i = 0; i < 3; ++i) first[i] = (mu / R(2)) * (R(2) * s[i]) + d1 * g[i];
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[i][j] = (i == j ? diag : R(0)) + d2 * g[i] * g[j] + d1 * Hm[i][j];
    const R sep = R(SINGULAR_VALUE_SEPARATION);
    const R kx = std::abs(x - y) < sep ? (first[0] + x * M[0][0] - y * M[1][0]) / R(2) / x
                                       : (x * first[0] - y * first[1]) / (x * x - y * y);
    const R ky = std::abs(y - z) < sep ? (first[1] + y * M[1][1] - z * M[2][1]) / R(2) / y
                                       : (y * first[1] - z * first[2]) / (y * y - z * z);
    const R kz = std::abs(z - x) < sep ? (first[2] + x * M[2][2] - x * M[0][2]) / R(2) / z
                                       : (z * first[2] - x * first[0]) / (z * z - x * x);
    const R kappa = std::max({x * x * M[0][0], y * y * M[1][1], z * z * M[2][2], y * y * kx, y * z * ky, z * x * kz}) / J;
    const R rho = R(p.mass) / R(p.initial_volume) / J;
    const R c2 = kappa / rho;
    return c2 > R(0) ? h / std::sqrt(c2) : std::numeric_limits<R>::infinity();
}

// limit_time_step_by_isolated_particles for one particle (:132-185).
template <class R> R isolated_limit(bool fluid, const mpm_particle_parameters &p, const M3<R> &F, R h) {
    const R m = R(p.mass), vol = R(p.initial_volume);
    if (!fluid) {
        const R xi = R(3) / h / h;
        return std::sqrt(m / (vol * xi * (R(1) - R(1) / R(2)) * (R(p.mu) + R(3) / R(2) * R(p.lambda))));
    }
    const R rho0 = m / vol;
    const R j = F.determinant();
    const R K = R(p.bulk_modulus);
    const int gamma = p.exponent;
    const R first = K * (R(1) - R(1) / powi(j, gamma));
    if (std::abs(j - R(1)) > R(SINGULAR_VALUE_SEPARATION)) return h / j * std::sqrt(rho0 * (j - R(1)) / (R(6) * first * R(3)));
    const R second = R(gamma) * K / powi(j, gamma + 1);
    return h * std::sqrt(rho0 / (R(6) * second * R(3)));
}

// AdaptiveTimeStepState (rust/crates/cpu/src/adaptive_time_step_state.rs:13-70), f32
// as in the reference; a missing limit is f32::MAX.
struct AdaptiveState {
    float max_time_step = std::numeric_limits<float>::max();
    float by_velocity = std::numeric_limits<float>::max(), by_deformation = std::numeric_limits<float>::max();
    float by_sound = std::numeric_limits<float>::max(), by_isolated = std::numeric_limits<float>::max();
    std::vector<float> prior;
    float without_prior() const { return std::min({max_time_step, by_velocity, by_deformation, by_sound, by_isolated}); }
    float allowed() const {
        float a = without_prior();
        for (float v : prior) a = std::min(a, v);
        return a;
    }
    void push_current_limit() {
        if (prior.size() > 10) prior.erase(prior.begin());
        prior.push_back(without_prior());
    }
};
inline float limit_value(double v) {
    return std::isfinite(v) ? (float)v : std::numeric_limits<float>::max();
}

template <class R>
struct State {
    // CpuState (rust/crates/cpu/src/cpu_state.rs:15-23) and Particles
    // (rust/crates/cpu/src/particles.rs:13-31).
    double time = 0;
    mpm_consts consts{};
    std::vector<uint32_t> sort_map, reverse_sort_map, flags, collider_bits;
    std::vector<mpm_particle_parameters> parameters;
    std::vector<V3<R>> initial_positions, positions, velocities;
    std::vector<M3<R>> position_gradients, velocity_gradients;
    std::vector<R> elastic_energies, plastic_jp;

    // GridNodes (rust/crates/cpu/src/grid_nodes.rs:22-33), in CSR form.
    std::vector<GridKey<R>> keys;
    std::vector<uint32_t> contrib_offsets, contributors;
    std::vector<R> masses;
    std::vector<V3<R>> node_velocities;       // momentum after scatter, velocity after meld
    std::vector<int32_t> lookup;              // dense node id -> node index (the FxHashMap)
    int lookup_lo[3] = {0, 0, 0}, lookup_dim[3] = {0, 0, 0};

    std::vector<uint32_t> lookup_count;       // twins (keys) at each dense node id

    // InterpolatedInput (rust/crates/cpu/src/interpolated_input.rs:12-22).
    V3<R> gravity;
    std::vector<V3<R>> goal_positions;
    bool have_goals = false;
    FrameInputCopy input;

    // The collider mesh and its interpolation for the current substep.
    MeshTopology mesh;
    std::vector<V3<R>> vertex_positions, vertex_normals, vertex_velocities, triangle_normals;
    std::vector<R> triangle_frictions, triangle_dampings;

    // The node with id and collider bits (GridNodes::map), or -1.
    int32_t lookup_node(const int id[3], uint32_t bits = 0) const {
        int64_t idx = 0;
        for (int a = 2; a >= 0; --a) {
            int c = id[a] - lookup_lo[a];
            if (c < 0 || c >= lookup_dim[a]) return -1;
            idx = idx * lookup_dim[a] + c;
        }
        const int32_t first = lookup[idx];
        if (first < 0) return -1;
        for (uint32_t k = 0; k < lookup_count[idx]; ++k)
            if (keys[first + k].collider_bits == bits) return first + (int32_t)k;
        return -1;
    }

    // Adaptive time steps (CpuRunParameters.adaptive_time_steps): the step of
    // each phase is the adaptive state's allowed step at that phase.
    bool adaptive = false;
    AdaptiveState ats;

    R h() const { return R(consts.grid_node_size); }
    float step() const { return adaptive ? ats.allowed() : consts.time_step; }
    R dt() const { return R(step()); }

    // limit_time_step_before_force (limit_time_step.rs:25-34).
    void limit_time_step_before_force() {
        const size_t n = positions.size();
        const R hh = h();
        R snd = std::numeric_limits<R>::infinity(), iso = std::numeric_limits<R>::infinity();
        for (size_t i = 0; i < n; ++i) {
            if (flags[i] & MPM_TOMBSTONED) continue;
            const bool fluid = flags[i] & MPM_IS_FLUID;
            snd = std::min(snd, sound_limit<R>(fluid, parameters[i], position_gradients[i], hh));
            iso = std::min(iso, isolated_limit<R>(fluid, parameters[i], position_gradients[i], hh));
        }
        ats.by_sound = limit_value((double)snd);
        ats.by_isolated = limit_value((double)iso);
        ats.push_current_limit();
    }

    // limit_time_step_before_integrate (limit_time_step.rs:187-222).
    void limit_time_step_before_integrate() {
        const size_t n = positions.size();
        R vmax = 0, cmax = 0;
        bool any = false;
        for (size_t i = 0; i < n; ++i) {
            if (flags[i] & MPM_TOMBSTONED) continue;
            any = true;
            vmax = std::max(vmax, std::sqrt(dot(velocities[i], velocities[i])));
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) cmax = std::max(cmax, std::abs(velocity_gradients[i].m[r][c]));
        }
        ats.by_velocity = any && vmax != R(0) ? limit_value((double)(R(0.5) * h() / vmax)) : std::numeric_limits<float>::max();
        ats.by_deformation = any ? limit_value((double)(R(0.2) / std::max(cmax, R(1e-8)))) : std::numeric_limits<float>::max();
        ats.push_current_limit();
    }

    // ---- InterpolateInput, rust/crates/cpu/src/phase/interpolate_input.rs:18-35
    // with FrameInput::frame_factor, rust/crates/xpu/src/frame_input.rs:266-278.
    int interpolate_input() {
        double frame_time = time * (double)consts.frames_per_second;
        uint64_t frame_low = (uint64_t)std::floor(frame_time);
        if (frame_low != input.frame) return MPM_ERR_WRONG_FRAME;
        float factor_b = (float)std::fmod(frame_time, 1.0);
        float factor_a = 1.f - factor_b;
        const float *gb = input.has_b ? input.gravity_b : input.gravity_a;
        // The reference interpolates in f32 (Vector3<f32>).
        gravity = V3<R>(R(factor_a * input.gravity_a[0] + factor_b * gb[0]),
                        R(factor_a * input.gravity_a[1] + factor_b * gb[1]),
                        R(factor_a * input.gravity_a[2] + factor_b * gb[2]));
        have_goals = !input.flags_a.empty();
        if (have_goals) {
            size_t n = sort_map.size();
            goal_positions.resize(n);
            const std::vector<float> &b = input.has_b ? input.goals_b : input.goals_a;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < (int64_t)n; ++i)
                goal_positions[i] = V3<R>(R(factor_a * input.goals_a[3 * i] + factor_b * b[3 * i]),
                                          R(factor_a * input.goals_a[3 * i + 1] + factor_b * b[3 * i + 1]),
                                          R(factor_a * input.goals_a[3 * i + 2] + factor_b * b[3 * i + 2]));
        }
        interpolate_mesh(factor_a, factor_b);
        return MPM_OK;
    }

    // The mesh half of interpolate_input (interpolate_input.rs:37-94): vertex
    // positions, frictions and dampings interpolated in f32 as the reference
    // does; triangle normals; vertex normals as the angle-weighted sums of
    // their triangles' normals (interpolate_input.rs:57-81); and the vertex
    // velocities (b - a) * fps (frame_input.rs:333-347).
    void interpolate_mesh(float factor_a, float factor_b) {
        const size_t nv = mesh.num_vertices, nt = mesh.num_triangles();
        if (nt == 0) return;
        if (input.vertices_a.size() != 3 * nv) throw std::runtime_error("the frame input holds no collider mesh");
        const std::vector<float> &pb = input.vertices_b.empty() ? input.vertices_a : input.vertices_b;
        vertex_positions.resize(nv);
        vertex_velocities.resize(nv);
        for (size_t v = 0; v < nv; ++v)
            for (int a = 0; a < 3; ++a) {
                vertex_positions[v][a] = R(factor_a * input.vertices_a[3 * v + a] + factor_b * pb[3 * v + a]);
                vertex_velocities[v][a] = input.vertices_b.empty() ? R(0)
                    : R((pb[3 * v + a] - input.vertices_a[3 * v + a]) * (float)consts.frames_per_second);
            }
        const std::vector<float> &fb = input.frictions_b.empty() ? input.frictions_a : input.frictions_b;
        const std::vector<float> &db = input.dampings_b.empty() ? input.dampings_a : input.dampings_b;
        triangle_normals.resize(nt);
        triangle_frictions.resize(nt);
        triangle_dampings.resize(nt);
        for (size_t t = 0; t < nt; ++t) {
            const V3<R> &a = vertex_positions[mesh.triangles[3 * t]];
            const V3<R> &b = vertex_positions[mesh.triangles[3 * t + 1]];
            const V3<R> &c = vertex_positions[mesh.triangles[3 * t + 2]];
            triangle_normals[t] = try_normalize(cross(b - a, c - a), R(NORMALIZATION_EPS));
            triangle_frictions[t] = R(factor_a * input.frictions_a[t] + factor_b * fb[t]);
            triangle_dampings[t] = R(factor_a * input.dampings_a[t] + factor_b * db[t]);
        }
        vertex_normals.assign(nv, V3<R>());
        for (size_t v = 0; v < nv; ++v) {
            V3<R> sum;
            for (uint32_t k = mesh.vertex_offsets[v]; k < mesh.vertex_offsets[v + 1]; ++k) {
                const uint32_t t = mesh.vertex_triangles[k];
                uint32_t others[2], no = 0;
                for (int e = 0; e < 3; ++e)
                    if (mesh.triangles[3 * t + e] != (uint32_t)v) others[no++] = mesh.triangles[3 * t + e];
                const V3<R> &p = vertex_positions[v];
                sum += triangle_normals[t] * angle_between(vertex_positions[others[0]] - p, vertex_positions[others[1]] - p);
            }
            vertex_normals[v] = try_normalize(sum, R(NORMALIZATION_EPS));
        }
    }

    // ---- Sort, rust/crates/cpu/src/phase/sort.rs:17-114.
    void sort() {
        const size_t n = positions.size();
        const R hh = h();
        std::vector<int> key(3 * n);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; ++i)
            for (int a = 0; a < 3; ++a) key[3 * i + a] = (int)std::floor(positions[i][a] / hh - R(0.5));
        std::vector<uint32_t> perm(n);
        std::iota(perm.begin(), perm.end(), 0u);
        std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) {
            for (int c = 0; c < 3; ++c) {
                if (key[3 * a + c] != key[3 * b + c]) return key[3 * a + c] < key[3 * b + c];
            }
            return false;
        });
        auto apply = [&](auto &v) {
            auto lookup_copy = v;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < (int64_t)n; ++i) v[i] = lookup_copy[perm[i]];
        };
        // The arrays that move with the particles (sort.rs:57-69); the elastic
        // energies are not moved because advance_particles overwrites them.
        apply(flags);
        apply(positions);
        apply(initial_positions);
        apply(sort_map);
        apply(parameters);
        apply(position_gradients);
        apply(velocities);
        apply(velocity_gradients);
        apply(collider_bits);
        apply(plastic_jp);
        reverse_sort_map.assign(n, 0);
        for (size_t cur = 0; cur < n; ++cur) reverse_sort_map[sort_map[cur]] = (uint32_t)cur;
    }

    // ---- Collide, rust/crates/cpu/src/phase/collide.rs:21-206.  Candidate
    // triangles are all triangles in ascending order (the reference's BVH
    // query returns a superset of the triangles within the forget distance,
    // in ascending order, and the distance test discards the rest).
    void collide() {
        const size_t n = positions.size();
        const size_t nt = mesh.num_triangles();
        const R hh = h(), t = dt();
        const R forget = R(consts.grid_node_size * 2.2f), accept = R(consts.grid_node_size * 2.0f);
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            if (flags[i] & MPM_TOMBSTONED) continue;
            if (nt == 0) {
                collider_bits[i] = 0;
                continue;
            }
            (void)hh;
            const V3<R> p = positions[i];
            V3<R> velocity = velocities[i];
            uint32_t bits = collider_bits[i];
            uint32_t closest[16];
            R min_distance[16];
            for (int c = 0; c < 16; ++c) {
                closest[c] = 0xffffffffu;
                min_distance[c] = std::numeric_limits<R>::max();
            }
            for (size_t tri = 0; tri < nt; ++tri) {
                const V3<R> &nrm = triangle_normals[tri];
                if (nrm.x == R(0) && nrm.y == R(0) && nrm.z == R(0)) continue;
                const V3<R> &a = vertex_positions[mesh.triangles[3 * tri]];
                const V3<R> &b = vertex_positions[mesh.triangles[3 * tri + 1]];
                const V3<R> &c = vertex_positions[mesh.triangles[3 * tri + 2]];
                const R distance = distance_to_triangle(p, a, b, c, nrm);
                if (distance >= forget) continue;
                const uint32_t col = mesh.triangle_colliders[tri];
                if (distance < min_distance[col]) {
                    min_distance[col] = distance;
                    closest[col] = (uint32_t)tri;
                }
            }
            for (int col = 0; col < 16; ++col) {
                if (closest[col] == 0xffffffffu) {
                    bits_set(bits, col, -1);
                    continue;
                }
                const uint32_t ct = closest[col];
                const uint32_t ia = mesh.triangles[3 * ct], ib = mesh.triangles[3 * ct + 1], ic = mesh.triangles[3 * ct + 2];
                const V3<R> &nrm = triangle_normals[ct];
                const V3<R> &a = vertex_positions[ia], &b = vertex_positions[ib], &c = vertex_positions[ic];
                const uint32_t o_ab = mesh.opposites[3 * ct], o_bc = mesh.opposites[3 * ct + 1], o_ca = mesh.opposites[3 * ct + 2];
                const V3<R> ab_n = o_ab != 0xffffffffu ? nrm + triangle_normals[o_ab] : V3<R>();
                const V3<R> bc_n = o_bc != 0xffffffffu ? nrm + triangle_normals[o_bc] : V3<R>();
                const V3<R> ca_n = o_ca != 0xffffffffu ? nrm + triangle_normals[o_ca] : V3<R>();
                const V3<R> ab = a - b, bc = b - c, ca = c - a;
                const R area2 = dot(nrm, cross(ca, ab));
                const R a_bary = dot(nrm, cross(bc, c - p)) / area2;
                const R b_bary = dot(nrm, cross(ca, a - p)) / area2;
                const R c_bary = dot(nrm, cross(ab, b - p)) / area2;
                DistanceResult<R> res;
                if (a_bary > R(0) && b_bary > R(0) && c_bary > R(0)) {
                    res = {std::abs(dot(p - a, nrm)), nrm * dot(p - a, nrm), nrm};
                } else {
                    // min_by over ab, bc, ca: the first of equal distances
                    DistanceResult<R> r3[3] = {
                        segment_distance_result(p, a, b, vertex_normals[ia], ab_n, vertex_normals[ib]),
                        segment_distance_result(p, b, c, vertex_normals[ib], bc_n, vertex_normals[ic]),
                        segment_distance_result(p, c, a, vertex_normals[ic], ca_n, vertex_normals[ia])};
                    res = r3[0];
                    for (int k = 1; k < 3; ++k)
                        if (r3[k].distance < res.distance) res = r3[k];
                }
                if (res.normal.x == R(0) && res.normal.y == R(0) && res.normal.z == R(0)) {
                    bits_set(bits, col, -1);
                    continue;
                }
                const bool new_side = R(0) <= dot(res.to_p, res.normal);
                if (!bits_near(bits, col)) {
                    if (res.distance < accept) bits_set(bits, col, new_side ? 1 : 0);
                    continue;
                }
                if (bits_side(bits, col) == new_side) continue;
                if (res.distance > R(NORMALIZATION_EPS)) {
                    const V3<R> collider_velocity = vertex_velocities[ia] * a_bary + vertex_velocities[ib] * b_bary +
                                                    vertex_velocities[ic] * c_bary;
                    const V3<R> relative = velocity - collider_velocity;
                    const V3<R> contact_normal = res.to_p / res.distance;
                    const V3<R> normal_velocity = contact_normal * dot(relative, contact_normal);
                    const V3<R> tangential = relative - normal_velocity;
                    const R tn = norm3(tangential);
                    if (tn > R(NORMALIZATION_EPS)) {
                        const V3<R> tangent = tangential / tn;
                        velocity -= tangent * std::min(triangle_frictions[ct] * res.distance / t, tn);
                    }
                    velocity -= normal_velocity * std::min(triangle_dampings[ct], R(1));
                }
                velocity -= res.to_p / t;
            }
            velocities[i] = velocity;
            collider_bits[i] = bits;
        }
    }

    // ---- ExternalForce, rust/crates/cpu/src/phase/external_force.rs:17-50.
    void external_force() {
        const size_t n = positions.size();
        const R t = dt();
        const bool grid_gravity = consts.options & MPM_OPT_GRID_GRAVITY;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            if (flags[i] & MPM_TOMBSTONED) continue;
            uint32_t orig = sort_map[i];
            const uint32_t *fb = input.has_b ? input.flags_b.data() : input.flags_a.data();
            if (have_goals && (input.flags_a[orig] & MPM_HAS_GOAL) && (fb[orig] & MPM_HAS_GOAL)) {
                velocities[i] = (goal_positions[orig] - positions[i]) / t;
            } else if (!grid_gravity) {
                velocities[i] += gravity * t;
            }
        }
    }

    // ---- UpdateGridNodes, rust/crates/cpu/src/phase/update_grid_nodes.rs:31-156.
    void update_grid_nodes() {
        const size_t n = positions.size();
        const R hh = h();
        int lo[3] = {std::numeric_limits<int>::max(), std::numeric_limits<int>::max(), std::numeric_limits<int>::max()};
        int hi[3] = {std::numeric_limits<int>::min(), std::numeric_limits<int>::min(), std::numeric_limits<int>::min()};
        std::vector<int> shift(3 * n);
        for (size_t i = 0; i < n; ++i) {
            if (flags[i] & MPM_TOMBSTONED) continue;
            position_to_shift_quadratic(positions[i], hh, &shift[3 * i]);
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], shift[3 * i + a]);
                hi[a] = std::max(hi[a], shift[3 * i + a] + 2);
            }
        }
        keys.clear();
        contrib_offsets.assign(1, 0);
        contributors.clear();
        if (lo[0] > hi[0]) {
            lookup.clear();
            for (int a = 0; a < 3; ++a) lookup_dim[a] = 0;
            return;
        }
        int64_t total = 1;
        for (int a = 0; a < 3; ++a) {
            lookup_lo[a] = lo[a];
            lookup_dim[a] = hi[a] - lo[a] + 1;
            total *= lookup_dim[a];
        }
        if (total > (int64_t)1 << 31) throw std::runtime_error("oracle grid box too large");
        // Keys are (node id, collider bits) (GridKey, grid_nodes.rs:15-19):
        // contributions with bits 0 are counted per node id, the rarer ones
        // with other bits are listed and sorted.
        std::vector<uint32_t> count((size_t)total, 0);
        struct Tagged {
            int64_t f;
            uint32_t bits, particle;
            bool operator<(const Tagged &o) const {
                return f != o.f ? f < o.f : (bits != o.bits ? bits < o.bits : particle < o.particle);
            }
        };
        std::vector<Tagged> tagged;
        auto flat = [&](const int id[3]) {
            return ((int64_t)(id[2] - lo[2]) * lookup_dim[1] + (id[1] - lo[1])) * lookup_dim[0] + (id[0] - lo[0]);
        };
        for (size_t i = 0; i < n; ++i) {
            if (flags[i] & MPM_TOMBSTONED) continue;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c) {
                        int id[3] = {shift[3 * i] + a, shift[3 * i + 1] + b, shift[3 * i + 2] + c};
                        if (collider_bits[i] == 0) ++count[flat(id)];
                        else tagged.push_back({flat(id), collider_bits[i], (uint32_t)i});
                    }
        }
        std::sort(tagged.begin(), tagged.end());
        lookup.assign((size_t)total, -1);
        lookup_count.assign((size_t)total, 0);
        // Node order: node id, then collider bits (the reference's order is the hash map's).
        size_t ti = 0;
        for (int64_t f = 0; f < total; ++f) {
            const bool plain = count[f] > 0;
            const bool twins = ti < tagged.size() && tagged[ti].f == f;
            if (!plain && !twins) continue;
            GridKey<R> k;
            k.id[0] = (int)(f % lookup_dim[0]) + lo[0];
            k.id[1] = (int)((f / lookup_dim[0]) % lookup_dim[1]) + lo[1];
            k.id[2] = (int)(f / ((int64_t)lookup_dim[0] * lookup_dim[1])) + lo[2];
            lookup[f] = (int32_t)keys.size();
            if (plain) {
                k.collider_bits = 0;
                keys.push_back(k);
                contrib_offsets.push_back(contrib_offsets.back() + count[f]);
                ++lookup_count[f];
            }
            while (ti < tagged.size() && tagged[ti].f == f) {
                const uint32_t bits = tagged[ti].bits;
                uint32_t cnt = 0;
                while (ti < tagged.size() && tagged[ti].f == f && tagged[ti].bits == bits) ++cnt, ++ti;
                k.collider_bits = bits;
                keys.push_back(k);
                contrib_offsets.push_back(contrib_offsets.back() + cnt);
                ++lookup_count[f];
            }
        }
        contributors.assign(contrib_offsets.back(), 0);
        std::vector<uint32_t> fill(keys.size(), 0);
        for (size_t i = 0; i < n; ++i) {
            if (flags[i] & MPM_TOMBSTONED) continue;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c) {
                        int id[3] = {shift[3 * i] + a, shift[3 * i + 1] + b, shift[3 * i + 2] + c};
                        int32_t node = lookup_node(id, collider_bits[i]);
                        contributors[contrib_offsets[node] + fill[node]++] = (uint32_t)i;
                    }
        }
    }

    // ---- ScatterMomentum, rust/crates/cpu/src/phase/scatter_momentum.rs:22-93.
    void scatter_momentum() {
        const R hh = h();
        const R scaling = dt() * R(4) / (hh * hh);
        const size_t nn = keys.size();
        masses.assign(nn, R(0));
        node_velocities.assign(nn, V3<R>());
        const R hardening = R(consts.snow_hardening);
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t node = 0; node < (int64_t)nn; ++node) {
            const GridKey<R> &key = keys[node];
            R mass = 0;
            V3<R> velocity;
            for (uint32_t c = contrib_offsets[node]; c < contrib_offsets[node + 1]; ++c) {
                uint32_t p = contributors[c];
                V3<R> normalized = positions[p] / hh;
                V3<R> to_node_normalized = V3<R>(R(key.id[0]), R(key.id[1]), R(key.id[2])) - normalized;
                R weight = kernel_quadratic(to_node_normalized.x) * kernel_quadratic(to_node_normalized.y) *
                           kernel_quadratic(to_node_normalized.z);
                V3<R> to_node = to_node_normalized * hh;
                const mpm_particle_parameters &par = parameters[p];
                R m = R(par.mass);
                V3<R> imparted = (velocities[p] + velocity_gradients[p] * to_node) * m;
                const M3<R> &F = position_gradients[p];
                M3<R> stress = first_piola<R>(flags[p], par, plastic_jp[p], hardening, F);
                if (flags[p] & MPM_USE_VISCOSITY) {
                    M3<R> cauchy = cauchy_viscosity<R>(par.viscosity_dynamic, par.viscosity_bulk, velocity_gradients[p]);
                    imparted -= cauchy * (to_node * (scaling * F.determinant() * R(par.initial_volume)));
                }
                imparted -= stress * (F.transpose() * (to_node * (scaling * R(par.initial_volume))));
                imparted *= weight;
                mass += weight * m;
                velocity += imparted;
            }
            masses[node] = mass;
            node_velocities[node] = velocity;
        }
    }

    // ---- MeldGrid, rust/crates/cpu/src/phase/meld_grid.rs:16-69: each key
    // adds the mass and momentum of the compatible twins at its node id (the
    // keys of one id are adjacent here), then divides.  The grid extensions
    // (grid gravity and boundary conditions) follow the division.
    void meld_grid() {
        const size_t nn = keys.size();
        const R hh = h(), t = dt();
        const bool grid_gravity = consts.options & MPM_OPT_GRID_GRAVITY;
        const std::vector<R> m0 = masses;
        const std::vector<V3<R>> p0 = node_velocities;
#pragma omp parallel for schedule(static)
        for (int64_t node = 0; node < (int64_t)nn; ++node) {
            R mass = m0[node];
            V3<R> v = p0[node];
            for (int64_t other = first_twin_bits(node); other < (int64_t)nn; ++other) {
                if (keys[other].id[0] != keys[node].id[0] || keys[other].id[1] != keys[node].id[1] ||
                    keys[other].id[2] != keys[node].id[2])
                    break;
                if (other == node || !collider_compatible(keys[node].collider_bits, keys[other].collider_bits)) continue;
                mass += m0[other];
                v += p0[other];
            }
            if (mass > 0) v = v / mass;
            else v = V3<R>();
            if (mass > 0) apply_grid_extensions(keys[node].id, v, hh, t, grid_gravity);
            node_velocities[node] = v;
            masses[node] = mass;
        }
    }

    // The first key (twin) with the same node id as key `node`.
    int64_t first_twin_bits(int64_t node) const {
        while (node > 0 && keys[node - 1].id[0] == keys[node].id[0] && keys[node - 1].id[1] == keys[node].id[1] &&
               keys[node - 1].id[2] == keys[node].id[2])
            --node;
        return node;
    }

    // Extension (no reference counterpart): grid gravity, sticky plane, box
    // walls, sphere (separating slip: a node inside loses its velocity
    // component along the inward normal).
    void apply_grid_extensions(const int id[3], V3<R> &v, R hh, R t, bool grid_gravity) const {
        if (grid_gravity) v += gravity * t;
        V3<R> xi(R(id[0]) * hh, R(id[1]) * hh, R(id[2]) * hh);
        if ((consts.options & MPM_OPT_STICKY_PLANE) && xi.z < R(consts.plane_z)) v = V3<R>();
        if (consts.options & MPM_OPT_BOX_WALLS) {
            for (int a = 0; a < 3; ++a) {
                if (xi[a] < R(consts.wall_min[a]) && v[a] < 0) v[a] = 0;
                if (xi[a] > R(consts.wall_max[a]) && v[a] > 0) v[a] = 0;
            }
        }
        if (consts.options & MPM_OPT_SPHERE)
            v = sphere_slip<R>(xi, V3<R>(R(consts.sphere_center[0]), R(consts.sphere_center[1]), R(consts.sphere_center[2])),
                               R(consts.sphere_radius), v);
    }

    // ---- CollectVelocity, rust/crates/cpu/src/phase/collect_velocity.rs:19-75.
    void collect_velocity() {
        const size_t n = positions.size();
        const R hh = h();
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            if (flags[i] & MPM_TOMBSTONED) continue;
            V3<R> velocity;
            M3<R> C;
            const V3<R> &position = positions[i];
            V3<R> normalized = position / hh;
            V3<R> shift(std::floor(normalized.x - R(0.5)), std::floor(normalized.y - R(0.5)),
                        std::floor(normalized.z - R(0.5)));
            V3<R> shifted = normalized - shift;
            R w[3][3];
            for (int a = 0; a < 3; ++a)
                for (int k = 0; k < 3; ++k) w[a][k] = kernel_quadratic(shifted[a] - R(k));
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c) {
                        R weight = w[0][a] * w[1][b] * w[2][c];
                        int id[3] = {(int)shift.x + a, (int)shift.y + b, (int)shift.z + c};
                        V3<R> node_pos(R(id[0]) * hh, R(id[1]) * hh, R(id[2]) * hh);
                        V3<R> to_node = node_pos - position;
                        int32_t node = lookup_node(id, collider_bits[i]);
                        if (node < 0) throw std::runtime_error("missing node");
                        V3<R> gv = node_velocities[node];
                        velocity += gv * weight;
                        C += outer(gv * weight, to_node);
                    }
            C *= R(4) / hh / hh;
            velocities[i] = velocity;
            velocity_gradients[i] = C;
        }
    }

    // ---- AdvanceParticles, rust/crates/cpu/src/phase/advance_particles.rs:17-93.
    // Returns false (EnergyError) when a Neo-Hookean particle ends with det F <= 0.
    bool advance_particles() {
        const size_t n = positions.size();
        const R t = dt();
        bool ok = true;
        const R snow_lo = R(1) - R(consts.snow_compression), snow_hi = R(1) + R(consts.snow_stretch);
        const R hardening = R(consts.snow_hardening);
#pragma omp parallel for schedule(static) reduction(&& : ok)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            if (flags[i] & MPM_TOMBSTONED) continue;
            positions[i] += velocities[i] * t;
            M3<R> &F = position_gradients[i];
            F += velocity_gradients[i] * F * t;
            const mpm_particle_parameters &p = parameters[i];
            if (flags[i] & MPM_IS_FLUID) {
                F = fluid_project<R>(F);
                elastic_energies[i] = energy_inviscid<R>(p.bulk_modulus, p.exponent, F);
                continue;
            }
            if (flags[i] & MPM_SNOW) {
                // Extension: snow return map (Stomakhin et al. 2013).
                M3<R> U, V;
                V3<R> s;
                svd3(F, U, s, V);
                R jp = plastic_jp[i];
                for (int a = 0; a < 3; ++a) {
                    R c = std::min(std::max(s[a], snow_lo), snow_hi);
                    jp *= s[a] / c;
                    s[a] = c;
                }
                plastic_jp[i] = jp;
                F = U * M3<R>::diag(s[0], s[1], s[2]) * V.transpose();
                R hfac = std::exp(hardening * (R(1) - jp));
                R e;
                kirchhoff_fixed_corotated<R>(R(p.mu) * hfac, R(p.lambda) * hfac, F, e);
                elastic_energies[i] = e;
                continue;
            }
            if ((flags[i] & MPM_USE_SAND_ALPHA)) F = sand_project<R>(F, R(p.mu), R(p.lambda), R(p.sand_alpha));
            if (flags[i] & MPM_FIXED_COROTATED) {
                R e;
                kirchhoff_fixed_corotated<R>(R(p.mu), R(p.lambda), F, e);
                elastic_energies[i] = e;
                continue;
            }
            R e;
            if (energy_neo_hookean<R>(p.mu, p.lambda, F, e)) {
                elastic_energies[i] = e;
            } else {
                flags[i] |= MPM_FAILED;
                ok = false;
            }
        }
        return ok;
    }

    // ---- CullParticles, rust/crates/cpu/src/phase/cull_particles.rs:17-41
    // (strict inequalities: a particle on the boundary is culled).
    void cull_particles() {
        const size_t n = positions.size();
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            if (flags[i] & MPM_TOMBSTONED) continue;
            bool within = true;
            for (int a = 0; a < 3; ++a) {
                R p = positions[i][a];
                within = within && (p > R(consts.domain_min[a]) && p < R(consts.domain_max[a]));
            }
            if (!within) flags[i] |= MPM_TOMBSTONED;
        }
    }

    // Phases up to and including ScatterMomentum (the P2G half of a substep),
    // with the limits phase between UpdateGridNodes and ScatterMomentum when
    // the step adapts (phase/mod.rs:27-75).
    int run_to_scatter() {
        int r = interpolate_input();
        if (r) return r;
        sort();
        collide();
        external_force();
        update_grid_nodes();
        if (adaptive) {
            limit_time_step_before_force();
            if (step() == 0.f) return MPM_ERR_ZERO_TIME_STEP;
        }
        scatter_momentum();
        return MPM_OK;
    }

    // The rest of a substep.  Returns MPM_ERR_ENERGY when advance_particles failed.
    int run_from_meld() {
        meld_grid();
        collect_velocity();
        if (adaptive) {
            limit_time_step_before_integrate();
            if (step() == 0.f) return MPM_ERR_ZERO_TIME_STEP;
        }
        if (!advance_particles()) return MPM_ERR_ENERGY;
        cull_particles();
        time += (double)step();  // cpu_state.rs:187-190, f32 dt added in f64
        return MPM_OK;
    }

    // produce_next_state, rust/crates/cpu/src/cpu_state.rs:146-197.
    int step_to(double target_time, const volatile int32_t *cancel, uint64_t &substeps) {
        substeps = 0;
        if (!adaptive && step() == 0.f) return MPM_ERR_ZERO_TIME_STEP;
        while (time < target_time) {
            if (step() == 0.f) return MPM_ERR_ZERO_TIME_STEP;
            if (cancel && *cancel) return MPM_ERR_CANCELED;
            int r = run_to_scatter();
            if (r) return r;
            r = run_from_meld();
            if (r) return r;
            ++substeps;
        }
        return MPM_OK;
    }
};

}  // namespace mpm_oracle
