// synth.cpp -- seeded synthetic C-alpha chains for the benchmark workloads.
//
// A chain of length L is built from successive 5-20 residue segments drawn
// from three kinds: helix (rise 1.5 A, 100 deg/residue, radius 2.3 A), strand
// (rise 3.3 A, alternating +-0.94 A so that consecutive C-alpha atoms are
// 3.8 A apart) and coil (3.8 A steps, bond angle ~N(110 deg, 10 deg), random
// dihedral).  Every placed atom must lie inside a sphere of radius
// 2.2 * L^0.38 A around the origin; a segment (or coil step) that leaves it is
// redrawn.  Homologs are rigidly moved, noisy copies of a source chain with
// random deletions and short coil insertions.
//
// Host code only; it prepares inputs and is not on the device path.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

namespace {

struct V3 { double x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 unit(V3 a) { double n = norm(a); return n > 0 ? a * (1.0 / n) : V3{1, 0, 0}; }

using Rng = std::mt19937_64;

V3 rand_unit(Rng& g) {
    std::normal_distribution<double> n(0.0, 1.0);
    for (;;) {
        V3 v{n(g), n(g), n(g)};
        double l = norm(v);
        if (l > 1e-6) return v * (1.0 / l);
    }
}

// orthonormal frame with first axis a
void frame(V3 a, V3& u, V3& v) {
    V3 t = std::fabs(a.x) < 0.9 ? V3{1, 0, 0} : V3{0, 1, 0};
    u = unit(cross(a, t));
    v = cross(a, u);
}

struct Builder {
    Rng g;
    double R;
    std::vector<V3> pts;
    explicit Builder(uint64_t seed, int L) : g(seed), R(2.2 * std::pow((double)L, 0.38)) {}
    bool inside(V3 p) const { return norm(p) <= R; }

    // next coil atom: 3.8 A bond, angle ~N(110,10) deg with the previous bond
    V3 coil_step() {
        std::normal_distribution<double> ang(110.0, 10.0);
        std::uniform_real_distribution<double> dih(-M_PI, M_PI);
        const size_t n = pts.size();
        if (n == 0) return {0, 0, 0};
        if (n == 1) return pts[0] + rand_unit(g) * 3.8;
        V3 b = unit(pts[n - 1] - pts[n - 2]);
        V3 u, v;
        frame(b, u, v);
        double th = M_PI - ang(g) * M_PI / 180.0;  // deviation from straight
        double ph = dih(g);
        V3 d = b * std::cos(th) + (u * std::cos(ph) + v * std::sin(ph)) * std::sin(th);
        return pts[n - 1] + unit(d) * 3.8;
    }
    void add_coil(int k) {
        for (int i = 0; i < k; i++) {
            V3 p{};
            bool ok = false;
            for (int t = 0; t < 60 && !ok; t++) { p = coil_step(); ok = inside(p); }
            if (!ok) {  // pull toward the centre
                V3 last = pts.empty() ? V3{0, 0, 0} : pts.back();
                V3 dir = unit(V3{0, 0, 0} - last + rand_unit(g) * 0.5);
                p = last + dir * 3.8;
            }
            pts.push_back(p);
        }
    }
    // helix or strand segment of k atoms along a random axis
    bool add_regular(int kind, int k) {
        V3 last = pts.empty() ? V3{0, 0, 0} : pts.back();
        for (int t = 0; t < 30; t++) {
            V3 a = rand_unit(g), u, v;
            frame(a, u, v);
            std::vector<V3> seg(k);
            std::uniform_real_distribution<double> ph0(0, 2 * M_PI);
            double p0 = ph0(g);
            for (int i = 0; i < k; i++) {
                if (kind == 0) {
                    double th = p0 + i * 100.0 * M_PI / 180.0;
                    seg[i] = (u * std::cos(th) + v * std::sin(th)) * 2.3 + a * (1.5 * i);
                } else {
                    seg[i] = a * (3.3 * i) + u * ((i & 1) ? 0.94 : -0.94);
                }
            }
            // attach: first atom 3.8 A from the current end, along a random direction
            V3 start = pts.empty() ? V3{0, 0, 0} : last + rand_unit(g) * 3.8;
            V3 sh = start - seg[0];
            bool ok = true;
            for (int i = 0; i < k && ok; i++) { seg[i] = seg[i] + sh; ok = inside(seg[i]); }
            if (ok) { pts.insert(pts.end(), seg.begin(), seg.end()); return true; }
        }
        return false;
    }
    void build(int L) {
        std::uniform_int_distribution<int> seglen(5, 20), kind(0, 2);
        while ((int)pts.size() < L) {
            int k = std::min(seglen(g), L - (int)pts.size());
            int t = kind(g);
            if (t == 2 || !add_regular(t, k)) add_coil(k);
        }
        pts.resize(L);
    }
};

void random_rotation(Rng& g, double R[9]) {
    // uniform rotation from a random unit quaternion
    std::normal_distribution<double> n(0.0, 1.0);
    double q[4];
    double s = 0;
    for (int i = 0; i < 4; i++) { q[i] = n(g); s += q[i] * q[i]; }
    s = 1.0 / std::sqrt(s);
    for (int i = 0; i < 4; i++) q[i] *= s;
    double a = q[0], b = q[1], c = q[2], d = q[3];
    R[0] = a * a + b * b - c * c - d * d; R[1] = 2 * (b * c - a * d); R[2] = 2 * (b * d + a * c);
    R[3] = 2 * (b * c + a * d); R[4] = a * a - b * b + c * c - d * d; R[5] = 2 * (c * d - a * b);
    R[6] = 2 * (b * d - a * c); R[7] = 2 * (c * d + a * b); R[8] = a * a - b * b - c * c + d * d;
}

}  // namespace

extern "C" {

// One chain of length L; xyz receives 3*L floats (x0,y0,z0,x1,...).
void gta_synth_chain(uint64_t seed, int L, float* xyz) {
    Builder b(seed, L);
    b.build(L);
    for (int i = 0; i < L; i++) {
        xyz[3 * i] = (float)b.pts[i].x; xyz[3 * i + 1] = (float)b.pts[i].y; xyz[3 * i + 2] = (float)b.pts[i].z;
    }
}

// A homolog of src (Lsrc atoms): random deletions (fraction del), coil
// insertions of 1-4 atoms (fraction ins of sites), Gaussian noise sigma, a
// random rotation and a translation of up to maxtrans A.  out must hold
// maxout*3 floats; srcidx (optional) receives the source index of each output
// atom (-1 for inserted atoms).  Returns the homolog length.
// n chains at once (chain k of lens[k] residues from seeds[k] at
// xyz + 3*offs[k]) on nthreads threads; each chain depends only on its seed,
// so the result equals n calls of gta_synth_chain
void gta_synth_chain_batch(const uint64_t* seeds, const int* lens, const long long* offs, int n, float* xyz,
                           int nthreads) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (int k = 0; k < n; k++) gta_synth_chain(seeds[k], lens[k], xyz + 3 * offs[k]);
}

int gta_synth_homolog(uint64_t seed, const float* src, int Lsrc, float sigma, float maxtrans,
                      float del, float ins, float* out, int maxout, int* srcidx) {
    Rng g(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::normal_distribution<double> N(0.0, 1.0);
    std::vector<V3> p;
    std::vector<int> idx;
    p.reserve(Lsrc * 2);
    for (int i = 0; i < Lsrc; i++) {
        if (U(g) < del) continue;
        V3 a{src[3 * i], src[3 * i + 1], src[3 * i + 2]};
        if (!p.empty() && U(g) < ins) {
            int k = 1 + (int)(U(g) * 4);
            V3 prev = p.back();
            for (int j = 1; j <= k; j++) {
                double t = (double)j / (k + 1);
                V3 m = prev * (1 - t) + a * t + rand_unit(g) * 1.5;
                p.push_back(m);
                idx.push_back(-1);
            }
        }
        p.push_back(a);
        idx.push_back(i);
    }
    if (p.size() < 3) {  // keep at least the first atoms
        p.clear(); idx.clear();
        for (int i = 0; i < std::min(Lsrc, 3); i++) {
            p.push_back({src[3 * i], src[3 * i + 1], src[3 * i + 2]});
            idx.push_back(i);
        }
    }
    double R[9];
    random_rotation(g, R);
    V3 t = rand_unit(g) * (U(g) * maxtrans);
    int n = std::min((int)p.size(), maxout);
    for (int i = 0; i < n; i++) {
        V3 a = p[i] + V3{N(g), N(g), N(g)} * sigma;
        V3 r{R[0] * a.x + R[1] * a.y + R[2] * a.z, R[3] * a.x + R[4] * a.y + R[5] * a.z,
             R[6] * a.x + R[7] * a.y + R[8] * a.z};
        r = r + t;
        out[3 * i] = (float)r.x; out[3 * i + 1] = (float)r.y; out[3 * i + 2] = (float)r.z;
        if (srcidx) srcidx[i] = idx[i];
    }
    return n;
}

}  // extern "C"
