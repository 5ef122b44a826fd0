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

// ---- outdoor scene: Perlin heightfield + box buildings ------------------------------------

// 2D gradient noise (Perlin's improved noise, quintic fade) over a permutation of 0..255 drawn
// from the seed; eight unit gradient directions.
struct Perlin {
    int perm[512];
    explicit Perlin(uint64_t seed)
    {
        Rng g(seed, 0x9E71ull);
        int p[256];
        for (int i = 0; i < 256; i++) p[i] = i;
        for (int i = 255; i > 0; i--) std::swap(p[i], p[(int)(g.next() % (uint64_t)(i + 1))]);
        for (int i = 0; i < 512; i++) perm[i] = p[i & 255];
    }
    static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
    static double grad(int h, double x, double y)
    {
        static const double c = 0.70710678118654752;
        static const double gx[8] = {1, -1, 0, 0, c, -c, c, -c}, gy[8] = {0, 0, 1, -1, c, c, -c, -c};
        return gx[h & 7] * x + gy[h & 7] * y;
    }
    double noise(double x, double y) const
    {
        const double fx = std::floor(x), fy = std::floor(y);
        const int xi = (int)((int64_t)fx & 255), yi = (int)((int64_t)fy & 255);
        const double xf = x - fx, yf = y - fy;
        const double u = fade(xf), v = fade(yf);
        const int a = perm[xi] + yi, b = perm[xi + 1] + yi;
        const double n00 = grad(perm[a], xf, yf), n10 = grad(perm[b], xf - 1, yf);
        const double n01 = grad(perm[a + 1], xf, yf - 1), n11 = grad(perm[b + 1], xf - 1, yf - 1);
        const double x0 = n00 + u * (n10 - n00), x1 = n01 + u * (n11 - n01);
        return x0 + v * (x1 - x0);
    }
    // fractal sum: 5 octaves from a 128 m wavelength, gain 1/2
    double fbm(double x, double y) const
    {
        double s = 0, amp = 1, f = 1.0 / 128.0;
        for (int o = 0; o < 5; o++) {
            s += amp * noise(x * f + 17.3 * o, y * f - 9.1 * o);
            amp *= 0.5;
            f *= 2;
        }
        return s;
    }
};

// A 200 x 200 m area centred on the origin: terrain z = H(x, y), bilinear over a 0.5 m grid of
// Perlin fBm samples (+-3 m), and box buildings standing in it (tops at most 25 m above the
// terrain's lowest point, so the scene is at most 30 m high).
struct Outdoor {
    static constexpr double HALF = 100.0, CELL = 0.5;
    static constexpr int G = 400;  // cells per side
    std::vector<double> h;         // (G + 1)^2 samples
    double slope = 0;              // largest |dH/dx| + |dH/dy| over the bilinear patches
    std::vector<Box> blds;

    double at(int i, int j) const { return h[(size_t)j * (G + 1) + i]; }
    double height(double x, double y) const
    {
        double u = (x + HALF) / CELL, v = (y + HALF) / CELL;
        u = std::min(std::max(u, 0.0), (double)G);
        v = std::min(std::max(v, 0.0), (double)G);
        int i = std::min((int)u, G - 1), j = std::min((int)v, G - 1);
        const double a = u - i, b = v - j;
        const double h0 = at(i, j) + a * (at(i + 1, j) - at(i, j));
        const double h1 = at(i, j + 1) + a * (at(i + 1, j + 1) - at(i, j + 1));
        return h0 + b * (h1 - h0);
    }
    bool in_building(double x, double y, double margin) const
    {
        for (auto& b : blds)
            if (x > b.lo[0] - margin && x < b.hi[0] + margin && y > b.lo[1] - margin && y < b.hi[1] + margin)
                return true;
        return false;
    }
};

static Outdoor make_outdoor(uint64_t seed, int n_bld)
{
    Outdoor sc;
    const Perlin pn(seed);
    const int G = Outdoor::G;
    sc.h.resize((size_t)(G + 1) * (G + 1));
    for (int j = 0; j <= G; j++)
        for (int i = 0; i <= G; i++)
            sc.h[(size_t)j * (G + 1) + i] =
                3.0 * pn.fbm(-Outdoor::HALF + i * Outdoor::CELL, -Outdoor::HALF + j * Outdoor::CELL);
    double zmin = INFINITY;
    for (double v : sc.h) zmin = std::min(zmin, v);
    for (int j = 0; j < G; j++)
        for (int i = 0; i < G; i++) {
            const double sx = std::max(std::fabs(sc.at(i + 1, j) - sc.at(i, j)),
                                       std::fabs(sc.at(i + 1, j + 1) - sc.at(i, j + 1)));
            const double sy = std::max(std::fabs(sc.at(i, j + 1) - sc.at(i, j)),
                                       std::fabs(sc.at(i + 1, j + 1) - sc.at(i + 1, j)));
            sc.slope = std::max(sc.slope, (sx + sy) / Outdoor::CELL);
        }
    Rng g(seed, 0xB1D6ull);
    for (int b = 0; b < n_bld; b++) {
        for (int tries = 0; tries < 1000; tries++) {
            const double sx = g.uniform(2, 8), sy = g.uniform(2, 8), ht = g.uniform(3, 20);
            const double cx = g.uniform(-Outdoor::HALF + sx, Outdoor::HALF - sx);
            const double cy = g.uniform(-Outdoor::HALF + sy, Outdoor::HALF - sy);
            if (std::hypot(cx, cy) < 0.5 * std::hypot(sx, sy) + 5.0) continue;  // the central scanner
            double base = INFINITY;
            for (int c = 0; c < 4; c++)
                base = std::min(base, sc.height(cx + ((c & 1) ? sx : -sx) / 2, cy + ((c & 2) ? sy : -sy) / 2));
            base -= 0.5;  // set into the ground
            const double top = std::min(base + ht, zmin + 25.0);
            if (top <= base + 1.0) continue;
            sc.blds.push_back(Box{{cx - sx / 2, cy - sy / 2, base}, {cx + sx / 2, cy + sy / 2, top}});
            break;
        }
    }
    return sc;
}

// Building lists per azimuth bin as seen from o: a ray tests only the buildings of its bin.
struct AzBins {
    static constexpr int NB = 7200;
    std::vector<std::vector<int>> bins;
    AzBins(const Outdoor& sc, const double* o) : bins(NB)
    {
        for (int b = 0; b < (int)sc.blds.size(); b++) {
            const Box& B = sc.blds[b];
            double a[4];
            for (int c = 0; c < 4; c++)
                a[c] = std::atan2(((c & 2) ? B.hi[1] : B.lo[1]) - o[1], ((c & 1) ? B.hi[0] : B.lo[0]) - o[0]);
            // the smallest arc covering the four corner directions (o is outside the footprint)
            double lo = a[0], width = 0;
            for (int c = 0; c < 4; c++) {
                double w = 0;
                for (int e = 0; e < 4; e++) {
                    double d = a[e] - a[c];
                    while (d < 0) d += 2 * M_PI;
                    w = std::max(w, d);
                }
                if (c == 0 || w < width) {
                    width = w;
                    lo = a[c];
                }
            }
            const int b0 = (int)std::floor((lo + M_PI) / (2 * M_PI) * NB) - 1;
            const int nb = (int)std::ceil(width / (2 * M_PI) * NB) + 3;
            for (int k = 0; k < nb && k < NB; k++) bins[(size_t)(((b0 + k) % NB) + NB) % NB].push_back(b);
        }
    }
    const std::vector<int>& of(double az) const
    {
        int k = (int)std::floor((az + M_PI) / (2 * M_PI) * NB);
        return bins[(size_t)((k % NB) + NB) % NB];
    }
};

// First hit of the ray o + t d (|d| = 1) with the terrain or a building, or +inf.
static double hit_outdoor(const Outdoor& sc, const AzBins& az, const double* o, const double* d)
{
    double tb = INFINITY;
    for (int b : az.of(std::atan2(d[1], d[0]))) tb = std::min(tb, hit_box(o, d, sc.blds[b]));
    // leave the area: the terrain exists inside the 200 x 200 m square only
    double tx = INFINITY;
    for (int a = 0; a < 2; a++) {
        if (d[a] > 1e-15) tx = std::min(tx, (Outdoor::HALF - o[a]) / d[a]);
        if (d[a] < -1e-15) tx = std::min(tx, (-Outdoor::HALF - o[a]) / d[a]);
    }
    const double tmax = std::min(tb, tx);
    // terrain: f(t) = height above ground along the ray; |f'| <= |dz| + slope * |dxy|, so steps of
    // f / that bound never cross the surface
    const double lip = std::fabs(d[2]) + sc.slope * std::hypot(d[0], d[1]);
    double t = 0;
    for (int it = 0; it < 100000 && t < tmax; it++) {
        const double f = o[2] + t * d[2] - sc.height(o[0] + t * d[0], o[1] + t * d[1]);
        if (f < 1e-5) return t;
        t += f / lip;
    }
    return tb;
}

// Exactly n points from rays on a uniform (azimuth 0..360, elevation -60..+90 deg) grid from o:
// the grid is sized from a pilot scan to give at least n hits, and the hits, in ray order, are
// thinned evenly to n.  Radial N(0, sigma^2) noise from the ray's own stream.
static int cast_outdoor(const Outdoor& sc, const double* o, int64_t n, double sigma, uint64_t noise_seed, float* out)
{
    if (n <= 0) return 0;
    const AzBins az(sc, o);
    auto ray = [](int64_t i, int n_az, int n_el, double* d) {
        const int ia = (int)(i / n_el), ie = (int)(i % n_el);
        const double a = (ia + 0.5) * (2 * M_PI / n_az);
        const double el = (-60.0 + (ie + 0.5) * (150.0 / n_el)) * (M_PI / 180.0);
        d[0] = std::cos(el) * std::cos(a);
        d[1] = std::cos(el) * std::sin(a);
        d[2] = std::sin(el);
    };
    // hit fraction on a pilot grid
    const int pe = 64, pa = 164;
    int64_t ph = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : ph)
    for (int64_t i = 0; i < (int64_t)pa * pe; i++) {
        double d[3];
        ray(i, pa, pe, d);
        if (std::isfinite(hit_outdoor(sc, az, o, d))) ph++;
    }
    double frac = std::max((double)ph / ((double)pa * pe), 1e-3);
    for (int attempt = 0; attempt < 8; attempt++) {
        const double rays = (double)n / frac * 1.03 + 64;
        const int n_el = std::max(4, (int)std::llround(std::sqrt(rays / 2.56)));
        const int n_az = std::max(4, (int)std::ceil(rays / n_el));
        const int64_t total = (int64_t)n_az * n_el;
        std::vector<float> tt((size_t)total);
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t i = 0; i < total; i++) {
            double d[3];
            ray(i, n_az, n_el, d);
            tt[(size_t)i] = (float)hit_outdoor(sc, az, o, d);
        }
        // hits in ray order: their ranks
        std::vector<int64_t> rank((size_t)total);
        int64_t hits = 0;
        for (int64_t i = 0; i < total; i++) {
            rank[(size_t)i] = hits;
            if (std::isfinite(tt[(size_t)i])) hits++;
        }
        if (hits < n) {
            frac = std::max((double)hits / (double)total, 1e-3) * 0.97;
            continue;
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < total; i++) {
            if (!std::isfinite(tt[(size_t)i])) continue;
            const int64_t j = rank[(size_t)i];
            // hit j is kept iff the even thinning hits -> n advances at it
            const int64_t k0 = (int64_t)((__int128)j * n / hits), k1 = (int64_t)((__int128)(j + 1) * n / hits);
            if (k1 == k0) continue;
            double d[3];
            ray(i, n_az, n_el, d);
            Rng r(noise_seed, (uint64_t)i);
            const double t = (double)tt[(size_t)i] + sigma * r.gauss();
            for (int c = 0; c < 3; c++) out[3 * k0 + c] = (float)(o[c] + d[c] * t);
        }
        return 0;
    }
    return 1;
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

// The outdoor scene of `seed` with n_bld buildings scanned from its centre (1.8 m above the
// terrain at the origin): exactly n points, radial noise sigma.
int scene_outdoor_scan(uint64_t seed, int n_bld, int64_t n, double sigma, float* out, double* scanner)
{
    const Outdoor sc = make_outdoor(seed, n_bld);
    const double o[3] = {0.0, 0.0, sc.height(0.0, 0.0) + 1.8};
    if (scanner) std::memcpy(scanner, o, sizeof(o));
    return cast_outdoor(sc, o, n, sigma, seed, out);
}

// A scanner pose in the outdoor scene: (x, y) uniform within 90 m of the centre and at least 2 m
// from every building, 1.8 m above the terrain (pose_seed picks it).
int scene_outdoor_pose(uint64_t seed, int n_bld, uint64_t pose_seed, double* out)
{
    const Outdoor sc = make_outdoor(seed, n_bld);
    Rng g(pose_seed, 0x9053ull);
    for (int tries = 0; tries < 100000; tries++) {
        const double x = g.uniform(-90.0, 90.0), y = g.uniform(-90.0, 90.0);
        if (sc.in_building(x, y, 2.0)) continue;
        out[0] = x;
        out[1] = y;
        out[2] = sc.height(x, y) + 1.8;
        return 0;
    }
    return 1;
}

// The outdoor scene scanned from `pos`: exactly n points, noise drawn from noise_seed.
int scene_outdoor_scan_at(uint64_t seed, int n_bld, const double* pos, int64_t n, double sigma, uint64_t noise_seed,
                          float* out)
{
    const Outdoor sc = make_outdoor(seed, n_bld);
    return cast_outdoor(sc, pos, n, sigma, noise_seed, out);
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
