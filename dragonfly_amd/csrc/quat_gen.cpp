// Host-side orientation sampling: the 600-cell refinement of SO(3) used by the reference
// (dragonfly/src/quaternion.c:27-842, quat_gen at :811-842).
//
// The sample set, its order and every floating-point operation that produces a value are
// those of the reference, so the output is bit-identical: the vertex table order
// (quaternion.c:67-113, with the even-permutation table of :31-34), the lexicographic
// edge/face/cell order (:134-221), the cell used for each point's weight (:227-297), the
// weight formula (:300-318) with its 5/6 and 35/36 factors (:323, :364), the half-sphere
// selection rule (:485-508) and the final weight normalisation (:835-839). Topology is
// computed here with exact arithmetic in Z[tau] instead of the reference's thresholded
// floating-point distances; both select exactly the nearest-neighbour pairs.
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/emc_hip.h"

namespace {

// x = (a + tau*b) / (2 * num) in each of 4 coordinates.
struct Lat {
    std::array<std::array<int, 2>, 4> c{};
};

struct Vert {
    std::array<double, 4> x{};  // double coordinates used for weights
    Lat lat;                    // exact lattice form scaled by num
};

struct Sample {
    Lat lat;
    double w;
};

inline double golden() { return (std::sqrt(5.) + 1.) / 2.; }

std::vector<Vert> make_vertices(int num) {
    static const int even_perm[12][4] = {{0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2},
                                         {1, 2, 0, 3}, {1, 0, 3, 2}, {1, 3, 2, 0},
                                         {2, 0, 1, 3}, {2, 3, 0, 1}, {2, 1, 3, 0},
                                         {3, 1, 0, 2}, {3, 0, 2, 1}, {3, 2, 1, 0}};
    const double tau = golden();
    std::vector<Vert> out;
    out.reserve(120);
    // (+-1/2)^4
    for (int code = 0; code < 16; ++code) {
        Vert v;
        for (int c = 0; c < 4; ++c) {
            int bit = (code >> (3 - c)) & 1;
            v.x[c] = bit - 0.5;
            v.lat.c[c][0] = (2 * bit - 1) * num;
        }
        out.push_back(v);
    }
    // +-1 on one axis
    for (int sign = 0; sign < 2; ++sign)
        for (int axis = 0; axis < 4; ++axis) {
            Vert v;
            v.x[axis] = 2 * sign - 1.0;
            v.lat.c[axis][0] = (2 * sign - 1) * 2 * num;
            out.push_back(v);
        }
    // even permutations of (+-tau, +-1, +-1/tau, 0) / 2
    for (int code = 0; code < 8; ++code) {
        int s0 = 2 * ((code >> 2) & 1) - 1, s1 = 2 * ((code >> 1) & 1) - 1, s2 = 2 * (code & 1) - 1;
        std::array<double, 4> base = {s0 * tau / 2., s1 * 0.5, s2 / (2 * tau), 0.};
        std::array<std::array<int, 2>, 4> bl = {{{0, s0 * num}, {s1 * num, 0}, {-s2 * num, s2 * num}, {0, 0}}};
        for (const auto &pm : even_perm) {
            Vert v;
            for (int c = 0; c < 4; ++c) {
                v.x[c] = base[pm[c]];
                v.lat.c[c] = bl[pm[c]];
            }
            out.push_back(v);
        }
    }
    return out;
}

// Inner product of two lattice points as P + Q*tau (exact; tau^2 = tau + 1).
inline std::pair<long, long> lat_dot(const Lat &a, const Lat &b) {
    long p = 0, q = 0;
    for (int c = 0; c < 4; ++c) {
        long a0 = a.c[c][0], a1 = a.c[c][1], b0 = b.c[c][0], b1 = b.c[c][1];
        p += a0 * b0 + a1 * b1;
        q += a0 * b1 + a1 * b0 + a1 * b1;
    }
    return {p, q};
}

double point_weight(const double *vq, const double *vc) {
    double dot = 0, nq = 0, nc = 0;
    for (int i = 0; i < 4; ++i) {
        nq += vq[i] * vq[i];
        nc += vc[i] * vc[i];
    }
    nq = std::sqrt(nq);
    nc = std::sqrt(nc);
    for (int i = 0; i < 4; ++i) dot += vq[i] * vc[i];
    return dot / (std::pow(nq, 4.0) * nc);
}

// Enumerate the interior points of a k-simplex (k = ids.size()-1) split num times, in the
// reference's nested-loop order: integer steps j_1..j_k >= 1 with sum <= num-1.
void simplex_points(const std::vector<Vert> &V, const std::vector<int> &ids, int num,
                    const double *vc, double factor, std::vector<Sample> &out) {
    const double tau = golden();
    const int k = (int)ids.size() - 1;
    std::array<Lat, 3> step;
    for (int i = 0; i < k; ++i)
        for (int c = 0; c < 4; ++c)
            for (int s = 0; s < 2; ++s)
                step[i].c[c][s] = (V[ids[i + 1]].lat.c[c][s] - V[ids[0]].lat.c[c][s]) / num;
    std::array<int, 3> j = {1, 1, 1};
    auto emit = [&]() {
        Sample p;
        double vq[4];
        for (int c = 0; c < 4; ++c)
            for (int s = 0; s < 2; ++s) {
                int acc = V[ids[0]].lat.c[c][s];
                for (int i = 0; i < k; ++i) acc += j[i] * step[i].c[c][s];
                p.lat.c[c][s] = acc;
            }
        for (int c = 0; c < 4; ++c) vq[c] = (p.lat.c[c][0] + tau * p.lat.c[c][1]) / (2.0 * num);
        p.w = factor * point_weight(vq, vc);
        out.push_back(p);
    };
    if (k == 1) {
        for (j[0] = 1; j[0] <= num - 1; ++j[0]) emit();
    } else if (k == 2) {
        for (j[0] = 1; j[0] <= num - 2; ++j[0])
            for (j[1] = 1; j[0] + j[1] <= num - 1; ++j[1]) emit();
    } else {
        for (j[0] = 1; j[0] <= num - 3; ++j[0])
            for (j[1] = 1; j[0] + j[1] <= num - 2; ++j[1])
                for (j[2] = 1; j[0] + j[1] + j[2] <= num - 1; ++j[2]) emit();
    }
}

void vertex_sum(const std::vector<Vert> &V, const std::vector<int> &ids, double *vc) {
    for (int c = 0; c < 4; ++c) {
        vc[c] = 0.;
        for (int id : ids) vc[c] += V[id].x[c];
    }
}

bool in_upper_half(const Lat &l) {
    for (int c = 0; c < 4; ++c)
        for (int s = 0; s < 2; ++s)
            if (l.c[c][s] != 0) return l.c[c][s] > 0;
    return false;
}

}  // namespace

extern "C" int emc_num_rot(int num_div) {
    if (num_div < 1 || num_div > 64) return EMC_E_ARG;
    return 10 * (5 * num_div * num_div * num_div + num_div);
}

extern "C" int emc_quat_gen(int num, double *quats) {
    if (num < 1 || num > 64 || quats == nullptr) return EMC_E_ARG;
    const std::vector<Vert> V = make_vertices(num);
    const int nv = (int)V.size();

    // neighbours: the largest inner product between distinct vertices
    std::pair<long, long> best{-1000000, 0};
    const double tau = golden();
    for (int a = 0; a < nv; ++a)
        for (int b = a + 1; b < nv; ++b) {
            auto d = lat_dot(V[a].lat, V[b].lat);
            if (d.first + d.second * tau > best.first + best.second * tau) best = d;
        }
    std::vector<std::vector<char>> nb(nv, std::vector<char>(nv, 0));
    for (int a = 0; a < nv; ++a)
        for (int b = 0; b < nv; ++b)
            nb[a][b] = (a != b) && lat_dot(V[a].lat, V[b].lat) == best;

    std::vector<std::array<int, 2>> edges;
    std::vector<std::array<int, 3>> faces;
    std::vector<std::array<int, 4>> cells;
    for (int a = 0; a < nv; ++a)
        for (int b = a + 1; b < nv; ++b) {
            if (!nb[a][b]) continue;
            edges.push_back({a, b});
            for (int c = b + 1; c < nv; ++c) {
                if (!(nb[a][c] && nb[b][c])) continue;
                faces.push_back({a, b, c});
                for (int d = c + 1; d < nv; ++d)
                    if (nb[a][d] && nb[b][d] && nb[c][d]) cells.push_back({a, b, c, d});
            }
        }
    if (edges.size() != 720 || faces.size() != 1200 || cells.size() != 600) return EMC_E_ARG;

    std::vector<Sample> pts_v(nv), pts_e, pts_f, pts_c;
    std::vector<char> done(nv, 0);
    for (const auto &cl : cells)
        for (int id : cl) {
            if (done[id]) continue;
            done[id] = 1;
            double vc[4];
            vertex_sum(V, {cl[0], cl[1], cl[2], cl[3]}, vc);
            pts_v[id].lat = V[id].lat;
            pts_v[id].w = 5. / 6 * point_weight(V[id].x.data(), vc);
        }
    if (num > 1)
        for (const auto &e : edges) {
            int a = e[0], b = e[1], c1 = -1, c2 = -1;
            for (int x = 0; x < nv && c1 < 0; ++x)
                if (x != b && nb[a][x] && nb[b][x]) c1 = x;
            for (int x = c1 + 1; x < nv && c2 < 0; ++x)
                if (x != b && nb[a][x] && nb[b][x] && nb[c1][x]) c2 = x;
            double vc[4];
            vertex_sum(V, {a, b, c1, c2}, vc);
            simplex_points(V, {a, b}, num, vc, 35. / 36, pts_e);
        }
    if (num > 2)
        for (const auto &f : faces) {
            int d = -1;
            for (int x = 0; x < nv && d < 0; ++x)
                if (x != f[1] && x != f[2] && nb[f[0]][x] && nb[f[1]][x] && nb[f[2]][x]) d = x;
            double vc[4];
            vertex_sum(V, {f[0], f[1], f[2], d}, vc);
            simplex_points(V, {f[0], f[1], f[2]}, num, vc, 1.0, pts_f);
        }
    if (num > 3)
        for (const auto &cl : cells) {
            double vc[4];
            vertex_sum(V, {cl[0], cl[1], cl[2], cl[3]}, vc);
            simplex_points(V, {cl[0], cl[1], cl[2], cl[3]}, num, vc, 1.0, pts_c);
        }

    const int nrot = emc_num_rot(num);
    int at = 0;
    for (const auto *set : {&pts_v, &pts_e, &pts_f, &pts_c})
        for (const Sample &p : *set) {
            if (!in_upper_half(p.lat)) continue;
            if (at >= nrot) return EMC_E_ARG;
            double q[4], n2 = 0;
            for (int c = 0; c < 4; ++c) {
                q[c] = (p.lat.c[c][0] + tau * p.lat.c[c][1]) / (2.0 * num);
                n2 += q[c] * q[c];
            }
            n2 = std::sqrt(n2);
            for (int c = 0; c < 4; ++c) quats[at * 5 + c] = q[c] / n2;
            quats[at * 5 + 4] = p.w;
            ++at;
        }
    if (at != nrot) return EMC_E_ARG;
    double total = 0.;
    for (int r = 0; r < nrot; ++r) total += quats[r * 5 + 4];
    total = 1. / total;
    for (int r = 0; r < nrot; ++r) quats[r * 5 + 4] *= total;
    return nrot;
}

// voronoi_subset (dragonfly/src/quaternion.c:844-862): nearest coarse orientation of every
// fine one, by 1 - (q_f . q_c)^2, the first minimum kept.
extern "C" int emc_voronoi_subset(int num_coarse, const double *qcoarse, int num_fine,
                                  const double *qfine, int *nearest) {
    if (num_coarse <= 0 || num_fine < 0 || !qcoarse || !qfine || !nearest) return -1;
    for (int i = 0; i < num_fine; ++i) {
        const double *f = qfine + 5 * (size_t)i;
        double dmin = 2.;
        int best = 0;
        for (int j = 0; j < num_coarse; ++j) {
            const double *c = qcoarse + 5 * (size_t)j;
            double dot = 0.;
            for (int k = 0; k < 4; ++k) dot += f[k] * c[k];
            const double dist = 1. - dot * dot;
            if (dist < dmin) {
                dmin = dist;
                best = j;
            }
        }
        nearest[i] = best;
    }
    return 0;
}
