// Seeded synthetic point clouds for the parity tests and the benchmark.
//
// Host-only C++ (OpenMP).  Every generator is a pure function of its seed: the PRNG is
// xoshiro256** seeded through splitmix64, Gaussian noise uses Box-Muller, and each
// sample or ray owns its own PRNG stream (seeded from (seed, ray index)), so the result
// does not depend on the OpenMP thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {  // xoshiro256**
    uint64_t s[4];
    Rng(uint64_t seed, uint64_t stream)
    {
        uint64_t x = seed * 0x100000001B3ull ^ (stream + 0x632BE59BD9B4E019ull);
        for (auto& v : s) v = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double gauss()
    {
        double u1 = uniform();
        double u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct Box {
    double lo[3], hi[3];
};
struct Cyl {  // vertical cylinder
    double cx, cy, r, z0, z1;
};

// ray/box slab test; returns the entry distance > 1e-9 or +inf
inline double hit_box(const double* o, const double* d, const Box& b)
{
    double t0 = 0.0, t1 = INFINITY;
    for (int a = 0; a < 3; a++) {
        if (std::fabs(d[a]) < 1e-15) {
            if (o[a] < b.lo[a] || o[a] > b.hi[a]) return INFINITY;
            continue;
        }
        double inv = 1.0 / d[a];
        double ta = (b.lo[a] - o[a]) * inv, tb = (b.hi[a] - o[a]) * inv;
        if (ta > tb) std::swap(ta, tb);
        t0 = std::max(t0, ta);
        t1 = std::min(t1, tb);
        if (t0 > t1) return INFINITY;
    }
    return t0 > 1e-9 ? t0 : INFINITY;
}

inline double hit_cyl(const double* o, const double* d, const Cyl& c)
{
    double best = INFINITY;
    double ox = o[0] - c.cx, oy = o[1] - c.cy;
    double A = d[0] * d[0] + d[1] * d[1];
    if (A > 1e-18) {
        double B = 2 * (ox * d[0] + oy * d[1]);
        double C = ox * ox + oy * oy - c.r * c.r;
        double disc = B * B - 4 * A * C;
        if (disc >= 0) {
            double sq = std::sqrt(disc);
            for (double t : {(-B - sq) / (2 * A), (-B + sq) / (2 * A)}) {
                double z = o[2] + t * d[2];
                if (t > 1e-9 && z >= c.z0 && z <= c.z1) {
                    best = std::min(best, t);
                    break;
                }
            }
        }
    }
    if (std::fabs(d[2]) > 1e-15) {  // caps
        for (double zc : {c.z0, c.z1}) {
            double t = (zc - o[2]) / d[2];
            if (t <= 1e-9) continue;
            double x = ox + t * d[0], y = oy + t * d[1];
            if (x * x + y * y <= c.r * c.r) best = std::min(best, t);
        }
    }
    return best;
}

// distance to the inside walls of an axis-aligned room containing o
inline double hit_room(const double* o, const double* d, const Box& room)
{
    double best = INFINITY;
    for (int a = 0; a < 3; a++) {
        if (d[a] > 1e-15)
            best = std::min(best, (room.hi[a] - o[a]) / d[a]);
        else if (d[a] < -1e-15)
            best = std::min(best, (room.lo[a] - o[a]) / d[a]);
    }
    return best;
}

}  // namespace

extern "C" {

// Plain small scan: a 2000x400 ground plane at z=0, three boxes with sides U[50,300]
// standing on it, a sphere of radius 150 resting on it; area-uniform samples with
// isotropic N(0, sigma^2) noise.
int scene_plane_boxes_sphere(uint64_t seed, int64_t n, double sigma, float* out)
{
    Rng g(seed, 0xABCDEFull);
    Box boxes[3];
    for (auto& b : boxes) {
        double sx = g.uniform(50, 300), sy = g.uniform(50, 300), sz = g.uniform(50, 300);
        double cx = g.uniform(-1000 + sx, 1000 - sx), cy = g.uniform(-200 + sy / 2, 200 - sy / 2);
        b = Box{{cx - sx / 2, cy - sy / 2, 0.0}, {cx + sx / 2, cy + sy / 2, sz}};
    }
    const double R = 150;
    const double scx = g.uniform(-850, 850), scy = g.uniform(-50, 50), scz = R;
    // surface areas: plane, 5 faces per box, sphere
    std::vector<double> area;
    area.push_back(2000.0 * 400.0);
    for (auto& b : boxes) {
        double ex = b.hi[0] - b.lo[0], ey = b.hi[1] - b.lo[1], ez = b.hi[2] - b.lo[2];
        area.push_back(ex * ey);  // top
        area.push_back(ex * ez);  // -y
        area.push_back(ex * ez);  // +y
        area.push_back(ey * ez);  // -x
        area.push_back(ey * ez);  // +x
    }
    area.push_back(4 * M_PI * R * R);
    std::vector<double> cum(area.size());
    double tot = 0;
    for (size_t i = 0; i < area.size(); i++) cum[i] = (tot += area[i]);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        Rng r(seed, (uint64_t)i);
        double u = r.uniform() * tot;
        size_t s = std::lower_bound(cum.begin(), cum.end(), u) - cum.begin();
        if (s >= area.size()) s = area.size() - 1;
        double p[3];
        if (s == 0) {
            p[0] = r.uniform(-1000, 1000);
            p[1] = r.uniform(-200, 200);
            p[2] = 0;
        } else if (s == area.size() - 1) {
            double z = r.uniform(-1, 1), phi = r.uniform(0, 2 * M_PI), rr = std::sqrt(1 - z * z);
            p[0] = scx + R * rr * std::cos(phi);
            p[1] = scy + R * rr * std::sin(phi);
            p[2] = scz + R * z;
        } else {
            const Box& b = boxes[(s - 1) / 5];
            int f = (int)((s - 1) % 5);
            double a = r.uniform(), c = r.uniform();
            double lx = b.lo[0] + a * (b.hi[0] - b.lo[0]), ly = b.lo[1] + a * (b.hi[1] - b.lo[1]);
            double lz = b.lo[2] + c * (b.hi[2] - b.lo[2]);
            switch (f) {
                case 0: p[0] = lx; p[1] = b.lo[1] + c * (b.hi[1] - b.lo[1]); p[2] = b.hi[2]; break;
                case 1: p[0] = lx; p[1] = b.lo[1]; p[2] = lz; break;
                case 2: p[0] = lx; p[1] = b.hi[1]; p[2] = lz; break;
                case 3: p[0] = b.lo[0]; p[1] = ly; p[2] = lz; break;
                default: p[0] = b.hi[0]; p[1] = ly; p[2] = lz; break;
            }
        }
        for (int j = 0; j < 3; j++) out[3 * i + j] = (float)(p[j] + sigma * r.gauss());
    }
    return 0;
}

struct Room {
    Box walls;
    std::vector<Box> boxes;
    std::vector<Cyl> cyls;
};

// A 20 x 15 x 4 m room with n_obj random boxes / vertical cylinders standing on the floor,
// keeping 0.6 m around the origin free.
static Room make_room(uint64_t seed, int n_obj)
{
    Rng g(seed, 0x5EEDull);
    Room rm{Box{{-8.0, -6.0, -1.5}, {12.0, 9.0, 2.5}}, {}, {}};
    const Box& room = rm.walls;
    for (int i = 0; i < n_obj; i++) {
        bool is_box = g.uniform() < 0.5;
        double h = g.uniform(0.3, 2.5);
        for (int tries = 0; tries < 1000; tries++) {
            double cx = g.uniform(room.lo[0] + 1, room.hi[0] - 1), cy = g.uniform(room.lo[1] + 1, room.hi[1] - 1);
            double sx = g.uniform(0.3, 2.0), sy = g.uniform(0.3, 2.0), r = g.uniform(0.1, 0.8);
            double ext = is_box ? 0.5 * std::sqrt(sx * sx + sy * sy) : r;
            if (std::sqrt(cx * cx + cy * cy) < ext + 0.6) continue;  // keep the scanner free
            if (is_box)
                rm.boxes.push_back(
                    Box{{cx - sx / 2, cy - sy / 2, room.lo[2]}, {cx + sx / 2, cy + sy / 2, room.lo[2] + h}});
            else
                rm.cyls.push_back(Cyl{cx, cy, r, room.lo[2], room.lo[2] + h});
            break;
        }
    }
    return rm;
}

static void cast_room(const Room& rm, uint64_t noise_seed, int n_az, int n_el, double sigma, const double o[3],
                      float* out)
{
    const int64_t total = (int64_t)n_az * n_el;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < total; i++) {
        int ia = (int)(i / n_el), ie = (int)(i % n_el);
        double az = (ia + 0.5) * (2 * M_PI / n_az);
        double el = (-60.0 + (ie + 0.5) * (150.0 / n_el)) * (M_PI / 180.0);
        double d[3] = {std::cos(el) * std::cos(az), std::cos(el) * std::sin(az), std::sin(el)};
        double t = hit_room(o, d, rm.walls);
        for (auto& b : rm.boxes) t = std::min(t, hit_box(o, d, b));
        for (auto& c : rm.cyls) t = std::min(t, hit_cyl(o, d, c));
        Rng r(noise_seed, (uint64_t)i);
        t += sigma * r.gauss();
        for (int j = 0; j < 3; j++) out[3 * i + j] = (float)(o[j] + d[j] * t);
    }
}

// Terrestrial-scan-like cloud: rays from a scanner at the origin on a uniform
// (azimuth 0..360, elevation -60..+90 deg) grid of n_az x n_el cell centres into the room;
// radial N(0, sigma^2) noise.  Writes n_az*n_el points.
int scene_room_scan(uint64_t seed, int n_az, int n_el, int n_obj, double sigma, float* out)
{
    const Room rm = make_room(seed, n_obj);
    const double o[3] = {0, 0, 0};
    cast_room(rm, seed, n_az, n_el, sigma, o, out);
    return 0;
}

// A scanner position in the room of `seed`: uniform over the floor area 0.5 m away from the
// walls and every object, height U[-0.5, 0.5] m (pose_seed picks it).
int scene_room_pose(uint64_t seed, int n_obj, uint64_t pose_seed, double* out)
{
    const Room rm = make_room(seed, n_obj);
    Rng g(pose_seed, 0x9053ull);
    for (int tries = 0; tries < 100000; tries++) {
        const double x = g.uniform(rm.walls.lo[0] + 0.5, rm.walls.hi[0] - 0.5);
        const double y = g.uniform(rm.walls.lo[1] + 0.5, rm.walls.hi[1] - 0.5);
        const double z = g.uniform(-0.5, 0.5);
        bool free = true;
        for (auto& b : rm.boxes)
            if (x > b.lo[0] - 0.5 && x < b.hi[0] + 0.5 && y > b.lo[1] - 0.5 && y < b.hi[1] + 0.5) free = false;
        for (auto& c : rm.cyls)
            if (std::hypot(x - c.cx, y - c.cy) < c.r + 0.5) free = false;
        if (!free) continue;
        out[0] = x;
        out[1] = y;
        out[2] = z;
        return 0;
    }
    return 1;
}

// The room scanned from `pos` (as scene_room_scan), noise drawn from noise_seed.
int scene_room_scan_at(uint64_t seed, int n_az, int n_el, int n_obj, double sigma, const double* pos,
                       uint64_t noise_seed, float* out)
{
    const Room rm = make_room(seed, n_obj);
    cast_room(rm, noise_seed, n_az, n_el, sigma, pos, out);
    return 0;
}

// Noisy sphere + torus: a unit-radius-scaled sphere of radius R at the origin and a
// torus (major 2R, minor 0.5R) around it; area-uniform samples, isotropic noise of
// sigma_rel * R.
int scene_sphere_torus(uint64_t seed, int64_t n, double R, double sigma_rel, float* out)
{
    const double a = 2.0 * R, b = 0.5 * R;
    const double As = 4 * M_PI * R * R, At = 4 * M_PI * M_PI * a * b;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        Rng r(seed, (uint64_t)i);
        double p[3];
        if (r.uniform() * (As + At) < As) {
            double z = r.uniform(-1, 1), phi = r.uniform(0, 2 * M_PI), rr = std::sqrt(1 - z * z);
            p[0] = R * rr * std::cos(phi);
            p[1] = R * rr * std::sin(phi);
            p[2] = R * z;
        } else {
            // area-uniform on the torus: accept theta with density ~ (a + b cos theta)
            double th, ph;
            for (;;) {
                th = r.uniform(0, 2 * M_PI);
                if (r.uniform() * (a + b) <= a + b * std::cos(th)) break;
            }
            ph = r.uniform(0, 2 * M_PI);
            p[0] = (a + b * std::cos(th)) * std::cos(ph);
            p[1] = (a + b * std::cos(th)) * std::sin(ph);
            p[2] = b * std::sin(th);
        }
        for (int j = 0; j < 3; j++) out[3 * i + j] = (float)(p[j] + sigma_rel * R * r.gauss());
    }
    return 0;
}

}  // extern "C"
