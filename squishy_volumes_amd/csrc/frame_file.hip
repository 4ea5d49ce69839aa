// frame_file.hip - the reference's on-disk frame format on the host side:
// an IoState (time, particles, optional grid nodes) as 32 magic bytes, 64
// version bytes and the serde encoding of IoState by bincode 1.3's default
// functions (little-endian, fixed-width integers, u64 sequence lengths, a u8
// tag before an Option's value, a u32 variant index before an enum's
// fields, fixed-size arrays without a length).  Layout sources:
//   IoState            rust/crates/file_frame/src/io_state.rs:11-18, magic :20-29
//   magic and version  rust/crates/file_util/src/lib.rs:27-29, 36-44, 63-71
//   Particles          rust/crates/file_frame/src/particles.rs:92-109
//   ParticleParameters rust/crates/file_frame/src/particles.rs:55-90
//   GridNodes          rust/crates/file_frame/src/grid_nodes.rs:9-15
// The engine's SoA view maps to it as ParticleFlags::from(&ParticleParameters)
// does (particles.rs:35-53): viscosity is Some iff MPM_USE_VISCOSITY, the
// parameters are Fluid iff MPM_IS_FLUID, sand_alpha is Some iff
// MPM_USE_SAND_ALPHA.  Jp (plastic_jp) has no place in the format.
#include <cstring>
#include <initializer_list>

#include "../../include/mpm_engine.h"

namespace {

constexpr char kMagic[33] = "Squishy Volumes Frame File Magic";
constexpr uint64_t kMagicLen = 32, kVersionLen = 64;
// The version of squishy_volumes_file_util whose readers accept the file
// (rust/crates/file_util/Cargo.toml).
constexpr char kVersion[] = "0.3.4";
constexpr uint64_t kNone = ~0ull;

uint64_t parameters_size(uint32_t flags) {
    uint64_t s = 4 + 4 + 1 + 4;                          // mass, volume, Option tag, variant
    if (flags & MPM_USE_VISCOSITY) s += 8;                // ViscosityParameters
    if (flags & MPM_IS_FLUID) return s + 8;               // exponent, bulk_modulus
    return s + 8 + 1 + ((flags & MPM_USE_SAND_ALPHA) ? 4 : 0);  // mu, lambda, Option<f32>
}

struct Writer {
    uint8_t *p;
    template <class T>
    void put(T v) {
        std::memcpy(p, &v, sizeof(T));
        p += sizeof(T);
    }
    void bytes(const void *src, uint64_t n) {
        if (n) std::memcpy(p, src, n);
        p += n;
    }
};

struct Reader {
    const uint8_t *p, *end;
    bool ok = true;
    template <class T>
    T get() {
        T v{};
        if ((uint64_t)(end - p) < sizeof(T)) {
            ok = false;
            return v;
        }
        std::memcpy(&v, p, sizeof(T));
        p += sizeof(T);
        return v;
    }
    const uint8_t *take(uint64_t n) {
        if ((uint64_t)(end - p) < n) {
            ok = false;
            return nullptr;
        }
        const uint8_t *q = p;
        p += n;
        return q;
    }
};

bool particles_ok(const mpm_particles_view *v) {
    return v && (v->count == 0 || (v->flags && v->parameters && v->elastic_energies && v->collider_bits &&
                                   v->positions && v->position_gradients && v->velocities &&
                                   v->velocity_gradients && v->initial_positions));
}

bool grid_ok(const mpm_grid_view *g) {
    return !g || g->count == 0 || (g->node_ids && g->collider_bits && g->masses && g->velocities);
}

// A length-prefixed sequence of n elements of `width` bytes from src.
void put_seq(Writer &w, const void *src, uint64_t n, uint64_t width) {
    w.put<uint64_t>(n);
    w.bytes(src, n * width);
}

// Reads a sequence's length and checks it against the expected count (and
// that its bytes are there); returns its bytes.
const uint8_t *get_seq(Reader &r, uint64_t n, uint64_t width) {
    if (r.get<uint64_t>() != n) r.ok = false;
    return r.ok ? r.take(n * width) : nullptr;
}

// A sequence of n elements of `width` bytes into dst.
void copy_seq(void *dst, Reader &r, uint64_t n, uint64_t width) {
    const uint8_t *src = get_seq(r, n, width);
    if (src && n) std::memcpy(dst, src, n * width);
}

}  // namespace

extern "C" {

uint64_t mpm_frame_file_size(const mpm_particles_view *particles, const mpm_grid_view *grid) {
    if (!particles_ok(particles) || !grid_ok(grid)) return 0;
    const uint64_t n = particles->count;
    uint64_t s = kMagicLen + kVersionLen + 8;  // time
    s += 8 + 4 * n;                            // flags
    s += 8;
    for (uint64_t i = 0; i < n; ++i) s += parameters_size(particles->flags[i]);
    s += 8 + 4 * n + 8 + 4 * n;                // elastic energies, collider bits
    s += 3 * (8 + 12 * n) + 2 * (8 + 36 * n);  // positions, velocities, initial positions; F, C
    s += 1;                                    // Option<GridNodes>
    if (grid) s += 4 * 8 + grid->count * (12 + 4 + 4 + 12);
    return s;
}

int mpm_frame_encode(const mpm_particles_view *particles, const mpm_grid_view *grid, uint8_t *out, uint64_t capacity,
                     uint64_t *written) {
    if (!particles_ok(particles) || !grid_ok(grid) || !written) return MPM_ERR_INVALID_ARGUMENT;
    const uint64_t size = mpm_frame_file_size(particles, grid);
    *written = size;
    if (!out || capacity < size) return MPM_ERR_CAPACITY;
    const uint64_t n = particles->count;
    Writer w{out};
    w.bytes(kMagic, kMagicLen);
    uint8_t version[kVersionLen] = {0};
    std::memcpy(version, kVersion, sizeof(kVersion) - 1);
    w.bytes(version, kVersionLen);
    w.put<double>(particles->time);
    put_seq(w, particles->flags, n, 4);
    w.put<uint64_t>(n);
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t f = particles->flags[i];
        const mpm_particle_parameters &q = particles->parameters[i];
        w.put<float>(q.mass);
        w.put<float>(q.initial_volume);
        if (f & MPM_USE_VISCOSITY) {
            w.put<uint8_t>(1);
            w.put<float>(q.viscosity_dynamic);
            w.put<float>(q.viscosity_bulk);
        } else {
            w.put<uint8_t>(0);
        }
        if (f & MPM_IS_FLUID) {
            w.put<uint32_t>(1);  // SpecificParticleParameters::Fluid
            w.put<int32_t>(q.exponent);
            w.put<float>(q.bulk_modulus);
        } else {
            w.put<uint32_t>(0);  // SpecificParticleParameters::Solid
            w.put<float>(q.mu);
            w.put<float>(q.lambda);
            if (f & MPM_USE_SAND_ALPHA) {
                w.put<uint8_t>(1);
                w.put<float>(q.sand_alpha);
            } else {
                w.put<uint8_t>(0);
            }
        }
    }
    put_seq(w, particles->elastic_energies, n, 4);
    put_seq(w, particles->collider_bits, n, 4);
    put_seq(w, particles->positions, n, 12);
    put_seq(w, particles->position_gradients, n, 36);
    put_seq(w, particles->velocities, n, 12);
    put_seq(w, particles->velocity_gradients, n, 36);
    put_seq(w, particles->initial_positions, n, 12);
    if (grid) {
        w.put<uint8_t>(1);
        put_seq(w, grid->node_ids, grid->count, 12);
        put_seq(w, grid->collider_bits, grid->count, 4);
        put_seq(w, grid->masses, grid->count, 4);
        put_seq(w, grid->velocities, grid->count, 12);
    } else {
        w.put<uint8_t>(0);
    }
    return (uint64_t)(w.p - out) == size ? MPM_OK : MPM_ERR_INVALID_ARGUMENT;
}

int mpm_frame_decode_counts(const uint8_t *data, uint64_t size, uint64_t *particles, uint64_t *grid_nodes) {
    if (!data || !particles || !grid_nodes) return MPM_ERR_INVALID_ARGUMENT;
    Reader r{data, data + size};
    const uint8_t *magic = r.take(kMagicLen);
    const uint8_t *version = r.take(kVersionLen);
    if (!r.ok || std::memcmp(magic, kMagic, kMagicLen) != 0) return MPM_ERR_INVALID_ARGUMENT;
    uint8_t expected[kVersionLen] = {0};
    std::memcpy(expected, kVersion, sizeof(kVersion) - 1);
    if (std::memcmp(version, expected, kVersionLen) != 0) return MPM_ERR_INVALID_ARGUMENT;
    r.get<double>();
    const uint64_t n = r.get<uint64_t>();
    if (!r.ok || n > (size / 4)) return MPM_ERR_INVALID_ARGUMENT;
    const uint8_t *flags = r.take(4 * n);
    if (!r.ok || r.get<uint64_t>() != n) return MPM_ERR_INVALID_ARGUMENT;
    // The parameters' layout follows from the flags: check each record's
    // option tags and variant against them, then skip it.
    for (uint64_t i = 0; i < n && r.ok; ++i) {
        uint32_t f;
        std::memcpy(&f, flags + 4 * i, 4);
        const uint8_t *q = r.take(parameters_size(f));
        if (!q) break;
        const uint8_t visc = q[8];
        uint32_t variant;
        std::memcpy(&variant, q + 9 + ((f & MPM_USE_VISCOSITY) ? 8 : 0), 4);
        bool match = visc == ((f & MPM_USE_VISCOSITY) ? 1 : 0) && variant == ((f & MPM_IS_FLUID) ? 1u : 0u);
        if (match && !(f & MPM_IS_FLUID)) {
            const uint8_t sand = q[9 + ((f & MPM_USE_VISCOSITY) ? 8 : 0) + 4 + 8];
            match = sand == ((f & MPM_USE_SAND_ALPHA) ? 1 : 0);
        }
        if (!match) return MPM_ERR_INVALID_ARGUMENT;
    }
    for (uint64_t width : {4ull, 4ull, 12ull, 36ull, 12ull, 36ull, 12ull}) get_seq(r, n, width);
    const uint8_t tag = r.get<uint8_t>();
    if (!r.ok || tag > 1) return MPM_ERR_INVALID_ARGUMENT;
    *particles = n;
    *grid_nodes = kNone;
    if (tag == 1) {
        const uint64_t m = r.get<uint64_t>();
        if (!r.ok || m > size / 12) return MPM_ERR_INVALID_ARGUMENT;
        r.take(12 * m);
        for (uint64_t width : {4ull, 4ull, 12ull}) get_seq(r, m, width);
        if (!r.ok) return MPM_ERR_INVALID_ARGUMENT;
        *grid_nodes = m;
    }
    return r.p == r.end ? MPM_OK : MPM_ERR_INVALID_ARGUMENT;
}

int mpm_frame_decode(const uint8_t *data, uint64_t size, mpm_particles_view *particles, mpm_grid_view *grid) {
    uint64_t n = 0, m = 0;
    int rc = mpm_frame_decode_counts(data, size, &n, &m);
    if (rc != MPM_OK) return rc;
    if (!particles_ok(particles) || particles->count != n) return MPM_ERR_INVALID_ARGUMENT;
    const bool has_grid = m != kNone;
    if (has_grid && grid && grid->capacity < m) {
        grid->count = m;  // the capacity wanted
        return MPM_ERR_CAPACITY;
    }
    if (has_grid && grid && m > 0 && !(grid->node_ids && grid->collider_bits && grid->masses && grid->velocities))
        return MPM_ERR_INVALID_ARGUMENT;
    Reader r{data + kMagicLen + kVersionLen, data + size};
    particles->time = r.get<double>();
    copy_seq(particles->flags, r, n, 4);
    r.get<uint64_t>();
    for (uint64_t i = 0; i < n; ++i) {
        mpm_particle_parameters q{};
        q.mass = r.get<float>();
        q.initial_volume = r.get<float>();
        if (r.get<uint8_t>() == 1) {
            q.viscosity_dynamic = r.get<float>();
            q.viscosity_bulk = r.get<float>();
        }
        if (r.get<uint32_t>() == 1) {
            q.exponent = r.get<int32_t>();
            q.bulk_modulus = r.get<float>();
        } else {
            q.mu = r.get<float>();
            q.lambda = r.get<float>();
            if (r.get<uint8_t>() == 1) q.sand_alpha = r.get<float>();
        }
        particles->parameters[i] = q;
    }
    copy_seq(particles->elastic_energies, r, n, 4);
    copy_seq(particles->collider_bits, r, n, 4);
    copy_seq(particles->positions, r, n, 12);
    copy_seq(particles->position_gradients, r, n, 36);
    copy_seq(particles->velocities, r, n, 12);
    copy_seq(particles->velocity_gradients, r, n, 36);
    copy_seq(particles->initial_positions, r, n, 12);
    if (particles->plastic_jp)
        for (uint64_t i = 0; i < n; ++i) particles->plastic_jp[i] = 1.0f;
    r.get<uint8_t>();
    if (grid) grid->count = 0;
    if (has_grid && grid) {
        grid->count = m;
        copy_seq(grid->node_ids, r, m, 12);
        copy_seq(grid->collider_bits, r, m, 4);
        copy_seq(grid->masses, r, m, 4);
        copy_seq(grid->velocities, r, m, 12);
    }
    return r.ok ? MPM_OK : MPM_ERR_INVALID_ARGUMENT;
}

}  // extern "C"
