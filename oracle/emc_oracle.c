/*
 * CPU oracle for one EMC iteration -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py load this
 * library, and only as the checker; the product never links or calls it.
 *
 * Restates the reference Cython engine (duaneloh/Dragonfly, dragonfly/src/) loop for loop,
 * including the per-thread keep lists, schedule(static,1) orientation split, thread-private
 * volumes and the linear keep-list searches, so that with one thread it reproduces the
 * reference's results and with many threads it times like the reference.
 *
 * How each part is pinned (tests/ hold the checks):
 *   orc_quat_gen        bit-exact vs the compiled reference quat_gen (oracle/_ref) and
 *                       committed golden hashes; closed-form count 10(5n^3+n); unit norm;
 *                       weight sum 1; half-sphere selection rule.
 *   orc_make_rot_quat   vs compiled reference; orthonormality; identity quaternion.
 *   orc_slice_gen3d     bit-exact vs compiled reference slice_gen3d; affine model is
 *                       reproduced exactly; out-of-grid value DBL_MIN.
 *   orc_slice_merge3d   bit-exact vs compiled reference slice_merge3d; weights of an
 *                       in-grid pixel sum to 1.
 *   orc_symmetrize_friedel  bit-exact vs compiled reference; Friedel invariance.
 *   orc_maximize        the reference's max_*.c need mpi.h/gsl/hdf5 and do not build here,
 *                       so this is pinned by an independent dense brute force on tiny inputs
 *                       (tests/test_oracle_maximize.py), by closed forms (single orientation,
 *                       constant model) and by invariants (posterior sums to 1, rescale
 *                       identity, scale identity).
 *   orc_update_model    Friedel invariance, divide-by-weight and rms closed forms.
 *   num_modes > 1       (combined orientations r = rotation * num_modes + mode) pinned by the
 *                       identical-modes reduction to the one-mode oracle: every mode a copy of
 *                       one model gives P / M per combined column, rmax = M * rmax_1, the same
 *                       rescale, likelihood and info (the log M term cancels the split),
 *                       model2 / M and the one-mode inter_weight in every mode, and
 *                       occupancies 1 / M; with modes made distinct, by the dense brute force.
 *   background path     (with_bg and need_scaling) orc_calc_powder against a dense sum; the
 *                       likelihood log(phi W + b B), posterior, likelihood and info against
 *                       a dense numpy formulation; a zero background gives the background-free
 *                       posterior and scales fcounts / psum_d; every view of
 *                       orc_update_tomogram_bg is zero, the negligible-background ratio or a
 *                       root (|G| < 1e-4) of the gradient, and every scale of
 *                       orc_update_scale_bg a root (|G| < 1e-3) or a 1e-5 bracket of one
 *                       (tests/test_oracle_background.py). Which root the bracketed search
 *                       lands on is not pinned beyond that tolerance: the reference's own
 *                       search moves by up to 4e-3 of the largest view under 1e-14
 *                       perturbations of its inputs.
 */
#include "emc_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* dragonfly/src/max_internal.h:6-8 */
#define ORC_PROB_MIN 1.e-6
#define ORC_PDIFF_THRESH 14.
#define ORC_MAX_EXP_START -1.e100

/* ------------------------------------------------------------------------------------ */
/* Orientation sampling: dragonfly/src/quaternion.c:27-842                               */
/* ------------------------------------------------------------------------------------ */

/* One 600-cell vertex: double coordinates and the exact (a + tau*b) lattice form. */
typedef struct {
    double x[4];
    int lat[4][2];
} orc_vertex;

typedef struct {
    int lat[4][2];
    double w;
} orc_point;

static double orc_tau(void) { return (sqrt(5.) + 1.) / 2.; }

/* Vertex enumeration order of quaternion.c:67-113 (16 half-integer, 8 axis, 96 even
 * permutations of (tau, 1, 1/tau, 0)/2 in the permutation order of :31-34). */
static void orc_vertices(int num, orc_vertex *v) {
    static const int perm[12][4] = {{0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2},
                                    {1, 2, 0, 3}, {1, 0, 3, 2}, {1, 3, 2, 0},
                                    {2, 0, 1, 3}, {2, 3, 0, 1}, {2, 1, 3, 0},
                                    {3, 1, 0, 2}, {3, 0, 2, 1}, {3, 2, 1, 0}};
    double tau = orc_tau();
    int n = 0, s[4];
    memset(v, 0, 120 * sizeof(orc_vertex));
    for (s[0] = 0; s[0] < 2; ++s[0])
        for (s[1] = 0; s[1] < 2; ++s[1])
            for (s[2] = 0; s[2] < 2; ++s[2])
                for (s[3] = 0; s[3] < 2; ++s[3], ++n)
                    for (int c = 0; c < 4; ++c) {
                        v[n].x[c] = s[c] - 0.5;
                        v[n].lat[c][0] = (2 * s[c] - 1) * num;
                    }
    for (int sg = 0; sg < 2; ++sg)
        for (int axis = 0; axis < 4; ++axis, ++n) {
            v[n].x[axis] = 2 * sg - 1.0;
            v[n].lat[axis][0] = (2 * sg - 1) * 2 * num;
        }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                double base[4];
                int blat[4][2] = {{0, (2 * i - 1) * num}, {(2 * j - 1) * num, 0},
                                  {-(2 * k - 1) * num, (2 * k - 1) * num}, {0, 0}};
                base[0] = (2 * i - 1) * tau / 2.;
                base[1] = (2 * j - 1) * 0.5;
                base[2] = (2 * k - 1) / (2 * tau);
                base[3] = 0;
                for (int m = 0; m < 12; ++m, ++n)
                    for (int c = 0; c < 4; ++c) {
                        v[n].x[c] = base[perm[m][c]];
                        v[n].lat[c][0] = blat[perm[m][c]][0];
                        v[n].lat[c][1] = blat[perm[m][c]][1];
                    }
            }
}

/* Squared distance between lattice points in Z[tau], returned as A + B*tau (exact). */
static void orc_lat_dist2(const int a[4][2], const int b[4][2], long *A, long *B) {
    long sa = 0, sb = 0;
    for (int c = 0; c < 4; ++c) {
        long da = a[c][0] - b[c][0], db = a[c][1] - b[c][1];
        /* (da + db tau)^2 = da^2 + db^2 (tau + 1) + 2 da db tau */
        sa += da * da + db * db;
        sb += 2 * da * db + db * db;
    }
    *A = sa;
    *B = sb;
}

/* Adjacency of the 600-cell: pairs at the minimal distance (quaternion.c:115-161). */
static void orc_adjacency(const orc_vertex *v, unsigned char adj[120][120]) {
    double tau = orc_tau(), best = 1e300;
    long bA = 0, bB = 0;
    for (int i = 0; i < 120; ++i)
        for (int j = i + 1; j < 120; ++j) {
            long A, B;
            orc_lat_dist2(v[i].lat, v[j].lat, &A, &B);
            double d = A + B * tau;
            if (d < best) { best = d; bA = A; bB = B; }
        }
    for (int i = 0; i < 120; ++i)
        for (int j = 0; j < 120; ++j) {
            long A, B;
            adj[i][j] = 0;
            if (i == j) continue;
            orc_lat_dist2(v[i].lat, v[j].lat, &A, &B);
            adj[i][j] = (A == bA && B == bB);
        }
}

/* Weight of a sample point (quaternion.c:300-318). */
static double orc_point_weight(const double *vq, const double *vc) {
    double w = 0, nq = 0, nc = 0;
    for (int i = 0; i < 4; ++i) {
        nq += pow(vq[i], 2);
        nc += pow(vc[i], 2);
    }
    nq = sqrt(nq);
    nc = sqrt(nc);
    for (int i = 0; i < 4; ++i) w += vq[i] * vc[i];
    w /= pow(nq, 4) * nc;
    return w;
}

static void orc_sum4(const orc_vertex *v, const int *ids, double *out) {
    for (int c = 0; c < 4; ++c) {
        out[c] = 0.;
        for (int k = 0; k < 4; ++k) out[c] += v[ids[k]].x[c];
    }
}

static void orc_lat_to_q(const int lat[4][2], int num, double *q) {
    double tau = orc_tau();
    for (int c = 0; c < 4; ++c) q[c] = (lat[c][0] + tau * lat[c][1]) / (2.0 * num);
}

int orc_num_rot(int num_div) { return 10 * (5 * num_div * num_div * num_div + num_div); }

/* First non-zero lattice component positive (quaternion.c:485-508). */
static int orc_upper_half(const int lat[4][2]) {
    for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 2; ++k)
            if (lat[c][k] != 0) return lat[c][k] > 0;
    return 0;
}

static int orc_emit(const orc_point *pts, int npts, int num, double *quats, int at) {
    for (int i = 0; i < npts; ++i) {
        if (!orc_upper_half(pts[i].lat)) continue;
        double q[4], nrm = 0;
        orc_lat_to_q(pts[i].lat, num, q);
        for (int c = 0; c < 4; ++c) nrm += pow(q[c], 2);
        nrm = sqrt(nrm);
        for (int c = 0; c < 4; ++c) quats[at * 5 + c] = q[c] / nrm;
        quats[at * 5 + 4] = pts[i].w;
        ++at;
    }
    return at;
}

int orc_quat_gen(int num, double *quats) {
    static unsigned char adj[120][120];
    orc_vertex v[120];
    int edges[720][2], faces[1200][3], cells[600][4];
    int ne = 0, nf = 0, nc = 0;
    double tau = orc_tau();
    if (num < 1) return -1;

    orc_vertices(num, v);
    orc_adjacency(v, adj);
    /* edges, faces, cells: cliques in lexicographic order (quaternion.c:134-221) */
    for (int a = 0; a < 120; ++a)
        for (int b = a + 1; b < 120; ++b) {
            if (!adj[a][b]) continue;
            edges[ne][0] = a; edges[ne][1] = b; ++ne;
        }
    for (int e = 0; e < ne; ++e)
        for (int c = edges[e][1] + 1; c < 120; ++c)
            if (adj[edges[e][0]][c] && adj[edges[e][1]][c]) {
                faces[nf][0] = edges[e][0]; faces[nf][1] = edges[e][1]; faces[nf][2] = c; ++nf;
            }
    for (int f = 0; f < nf; ++f)
        for (int d = faces[f][2] + 1; d < 120; ++d)
            if (adj[faces[f][0]][d] && adj[faces[f][1]][d] && adj[faces[f][2]][d]) {
                for (int k = 0; k < 3; ++k) cells[nc][k] = faces[f][k];
                cells[nc][3] = d;
                ++nc;
            }
    if (ne != 720 || nf != 1200 || nc != 600) return -2;

    int nverts = 120, nedge = 720 * (num - 1), nface = 1200 * (num - 2) * (num - 1) / 2;
    int ncell = 600 * (num - 3) * (num - 2) * (num - 1) / 6;
    if (nface < 0) nface = 0;
    if (ncell < 0) ncell = 0;
    orc_point *pv = calloc(nverts, sizeof(orc_point));
    orc_point *pe = calloc(nedge > 0 ? nedge : 1, sizeof(orc_point));
    orc_point *pf = calloc(nface > 0 ? nface : 1, sizeof(orc_point));
    orc_point *pc = calloc(ncell > 0 ? ncell : 1, sizeof(orc_point));

    /* vertex weights from the first cell holding each vertex (quaternion.c:320-357) */
    int seen[120] = {0};
    for (int c = 0; c < 600; ++c)
        for (int k = 0; k < 4; ++k) {
            int id = cells[c][k];
            if (seen[id]) continue;
            seen[id] = 1;
            double vc[4];
            memcpy(pv[id].lat, v[id].lat, sizeof(v[id].lat));
            orc_sum4(v, cells[c], vc);
            pv[id].w = 5. / 6 * orc_point_weight(v[id].x, vc);
        }
    /* edge points, cell through (a, b, smallest common neighbour, next) (:256-297, :359-393) */
    int at = 0;
    for (int e = 0; e < 720 && num > 1; ++e) {
        int a = edges[e][0], b = edges[e][1], c1 = -1, c2 = -1, ids[4];
        for (int x = 0; x < 120 && c1 < 0; ++x)
            if (x != b && adj[a][x] && adj[b][x]) c1 = x;
        for (int x = c1 + 1; x < 120 && c2 < 0; ++x)
            if (x != b && adj[a][x] && adj[b][x] && adj[c1][x]) c2 = x;
        ids[0] = a; ids[1] = b; ids[2] = c1; ids[3] = c2;
        double vc[4];
        orc_sum4(v, ids, vc);
        for (int j = 1; j < num; ++j, ++at) {
            double vq[4];
            for (int c = 0; c < 4; ++c)
                for (int k = 0; k < 2; ++k)
                    pe[at].lat[c][k] = v[a].lat[c][k] + j * ((v[b].lat[c][k] - v[a].lat[c][k]) / num);
            for (int c = 0; c < 4; ++c) vq[c] = (pe[at].lat[c][0] + tau * pe[at].lat[c][1]) / (2.0 * num);
            pe[at].w = 35. / 36 * orc_point_weight(vq, vc);
        }
    }
    /* face points, cell through the face and its smallest other vertex (:227-254, :395-431) */
    at = 0;
    for (int f = 0; f < 1200 && num > 2; ++f) {
        int a = faces[f][0], b = faces[f][1], c = faces[f][2], d = -1, ids[4];
        for (int x = 0; x < 120 && d < 0; ++x)
            if (x != b && x != c && adj[a][x] && adj[b][x] && adj[c][x]) d = x;
        ids[0] = a; ids[1] = b; ids[2] = c; ids[3] = d;
        double vc[4];
        orc_sum4(v, ids, vc);
        for (int j = 1; j < num - 1; ++j)
            for (int k = 1; k < num - j; ++k, ++at) {
                double vq[4];
                for (int q = 0; q < 4; ++q)
                    for (int s = 0; s < 2; ++s)
                        pf[at].lat[q][s] = v[a].lat[q][s] + j * ((v[b].lat[q][s] - v[a].lat[q][s]) / num)
                                           + k * ((v[c].lat[q][s] - v[a].lat[q][s]) / num);
                for (int q = 0; q < 4; ++q) vq[q] = (pf[at].lat[q][0] + tau * pf[at].lat[q][1]) / (2.0 * num);
                pf[at].w = orc_point_weight(vq, vc);
            }
    }
    /* cell interior points (:433-472) */
    at = 0;
    for (int cc = 0; cc < 600 && num > 3; ++cc) {
        const int *ids = cells[cc];
        double vc[4];
        orc_sum4(v, ids, vc);
        for (int j = 1; j < num - 2; ++j)
            for (int k = 1; k < num - 1 - j; ++k)
                for (int m = 1; m < num - j - k; ++m, ++at) {
                    double vq[4];
                    for (int q = 0; q < 4; ++q)
                        for (int s = 0; s < 2; ++s)
                            pc[at].lat[q][s] = v[ids[0]].lat[q][s]
                                + j * ((v[ids[1]].lat[q][s] - v[ids[0]].lat[q][s]) / num)
                                + k * ((v[ids[2]].lat[q][s] - v[ids[0]].lat[q][s]) / num)
                                + m * ((v[ids[3]].lat[q][s] - v[ids[0]].lat[q][s]) / num);
                    for (int q = 0; q < 4; ++q) vq[q] = (pc[at].lat[q][0] + tau * pc[at].lat[q][1]) / (2.0 * num);
                    pc[at].w = orc_point_weight(vq, vc);
                }
    }

    /* keep one of each +-q pair, in vertex/edge/face/cell order (:474-680) */
    int nrot = orc_num_rot(num);
    at = orc_emit(pv, nverts, num, quats, 0);
    at = orc_emit(pe, nedge, num, quats, at);
    at = orc_emit(pf, nface, num, quats, at);
    at = orc_emit(pc, ncell, num, quats, at);
    free(pv); free(pe); free(pf); free(pc);
    if (at != nrot) return -3;

    /* normalise weights (:835-839) */
    double total = 0.;
    for (int r = 0; r < nrot; ++r) total += quats[r * 5 + 4];
    total = 1. / total;
    for (int r = 0; r < nrot; ++r) quats[r * 5 + 4] *= total;
    return nrot;
}

void orc_voronoi_subset(int num_coarse, const double *qcoarse, int num_fine, const double *qfine,
                        int *nearest) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < num_fine; ++i) {
        double dmin = 2.;
        nearest[i] = 0;
        for (int j = 0; j < num_coarse; ++j) {
            double dot = 0.;
            for (int c = 0; c < 4; ++c) dot += qfine[i * 5 + c] * qcoarse[j * 5 + c];
            double dist = 1. - dot * dot;
            if (dist < dmin) {
                dmin = dist;
                nearest[i] = j;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------ */
/* Geometry: dragonfly/src/model.c                                                       */
/* ------------------------------------------------------------------------------------ */

/* model.c:759-786 */
void orc_make_rot_quat(const double *q, double *rot) {
    double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    double q01 = q0 * q1, q02 = q0 * q2, q03 = q0 * q3, q11 = q1 * q1, q12 = q1 * q2;
    double q13 = q1 * q3, q22 = q2 * q2, q23 = q2 * q3, q33 = q3 * q3;
    rot[0] = (1. - 2. * (q22 + q33));
    rot[1] = 2. * (q12 + q03);
    rot[2] = 2. * (q13 - q02);
    rot[3] = 2. * (q12 - q03);
    rot[4] = (1. - 2. * (q11 + q33));
    rot[5] = 2. * (q01 + q23);
    rot[6] = 2. * (q02 + q13);
    rot[7] = 2. * (q23 - q01);
    rot[8] = (1. - 2. * (q11 + q22));
}

/* Rotated voxel position of pixel t: rot * q_t + center, summed in the order of
 * model.c:20-25 / :184-189. */
static inline void orc_rotate(const double *rot, const double *qv, double center, double *out) {
    for (int i = 0; i < 3; ++i) {
        double acc = 0.;
        for (int j = 0; j < 3; ++j) acc += rot[i * 3 + j] * qv[j];
        out[i] = acc + center;
    }
}

/* model.c:11-59 */
void orc_slice_gen3d(const double *quat, const double *model, long size, int num_pix,
                     const double *qvals, const double *corr, double *slice) {
    double rot[9];
    long s2 = size * size;
    double center = (double)(size / 2);
    orc_make_rot_quat(quat, rot);
    for (long t = 0; t < num_pix; ++t) {
        double p[3];
        orc_rotate(rot, &qvals[t * 3], center, p);
        long x = (long)p[0], y = (long)p[1], z = (long)p[2];
        if (x < 0 || x > size - 2 || y < 0 || y > size - 2 || z < 0 || z > size - 2) {
            slice[t] = DBL_MIN;
            continue;
        }
        double fx = p[0] - x, fy = p[1] - y, fz = p[2] - z;
        double cx = 1. - fx, cy = 1. - fy, cz = 1. - fz;
        long x1 = (x + 1) % size, y1 = (y + 1) % size, z1 = (z + 1) % size;
        double val = cx * cy * cz * model[x * s2 + y * size + z] +
                     cx * cy * fz * model[x * s2 + y * size + z1] +
                     cx * fy * cz * model[x * s2 + y1 * size + z] +
                     cx * fy * fz * model[x * s2 + y1 * size + z1] +
                     fx * cy * cz * model[x1 * s2 + y * size + z] +
                     fx * cy * fz * model[x1 * s2 + y * size + z1] +
                     fx * fy * cz * model[x1 * s2 + y1 * size + z] +
                     fx * fy * fz * model[x1 * s2 + y1 * size + z1];
        slice[t] = val * corr[t];
    }
}

/* model.c:171-244 */
void orc_slice_merge3d(const double *quat, const double *slice, double *model2, double *weight,
                       long size, int num_pix, const double *qvals, const double *corr,
                       const uint8_t *mask) {
    double rot[9];
    long s2 = size * size;
    double center = (double)(size / 2);
    orc_make_rot_quat(quat, rot);
    for (long t = 0; t < num_pix; ++t) {
        if (mask[t] > 1) continue;
        double p[3];
        orc_rotate(rot, &qvals[t * 3], center, p);
        long x = (long)p[0], y = (long)p[1], z = (long)p[2];
        if (x < 0 || x > size - 2 || y < 0 || y > size - 2 || z < 0 || z > size - 2) continue;
        double fx = p[0] - x, fy = p[1] - y, fz = p[2] - z;
        double cx = 1. - fx, cy = 1. - fy, cz = 1. - fz;
        long x1 = (x + 1) % size, y1 = (y + 1) % size, z1 = (z + 1) % size;
        double w = slice[t] / corr[t];
        long idx[8] = {x * s2 + y * size + z, x * s2 + y * size + z1, x * s2 + y1 * size + z,
                       x * s2 + y1 * size + z1, x1 * s2 + y * size + z, x1 * s2 + y * size + z1,
                       x1 * s2 + y1 * size + z, x1 * s2 + y1 * size + z1};
        double f[8] = {cx * cy * cz, cx * cy * fz, cx * fy * cz, cx * fy * fz,
                       fx * cy * cz, fx * cy * fz, fx * fy * cz, fx * fy * fz};
        for (int k = 0; k < 8; ++k) {
            weight[idx[k]] += f[k];
            model2[idx[k]] += f[k] * w;
        }
    }
}

/* model.c:470-490 (sequential order matters on the z == center plane) */
void orc_symmetrize_friedel(double *a, const double *w, int size) {
    int lo = (size % 2 == 0) ? 1 : 0, c = size / 2;
    for (int x = lo; x < size; ++x)
        for (int y = lo; y < size; ++y)
            for (int z = lo; z <= c; ++z) {
                long v1 = (long)x * size * size + (long)y * size + z;
                long v2 = (long)(2 * c - x) * size * size + (long)(2 * c - y) * size + (2 * c - z);
                double ai = 0.5 * (a[v1] * w[v1] + a[v2] * w[v2]);
                double aw = 0.5 * (w[v1] + w[v2]);
                ai = (aw > 0) ? ai / aw : 0.;
                a[v1] = ai;
                a[v2] = ai;
            }
}

/* ------------------------------------------------------------------------------------ */
/* One iteration: dragonfly/src/maximize.c:8-56, max_prob.c, max_update.c, max_scale.c    */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    double *val;
    int *pos;
    int n, cap;
} orc_keep;

static void orc_keep_push(orc_keep *k, double v, int r) {
    if (k->n == k->cap) {
        k->cap = k->cap ? 2 * k->cap : 4;
        k->val = realloc(k->val, k->cap * sizeof(double));
        k->pos = realloc(k->pos, k->cap * sizeof(int));
    }
    k->val[k->n] = v;
    k->pos[k->n] = r;
    k->n++;
}

/* Drop entries with value <= thresh, keeping order (max_prob.c:429-448). */
static void orc_keep_prune(orc_keep *k, double thresh) {
    int o = 0;
    for (int i = 0; i < k->n; ++i)
        if (!(k->val[i] <= thresh)) {
            k->val[o] = k->val[i];
            k->pos[o] = k->pos[i];
            ++o;
        }
    k->n = o;
}

static int orc_keep_find(const orc_keep *k, int r) {
    for (int i = 0; i < k->n; ++i)
        if (k->pos[i] == r) return i;
    return -1;
}

static double orc_now(void) { return omp_get_wtime(); }

/* ---- background-aware updates (need_scaling with a detector background) ---------------- */

void orc_calc_powder(int num_pix, const uint8_t *mask, int num_data, const int *ones, const int *multi,
                     const long *ones_accum, const long *multi_accum, const int *place_ones,
                     const int *place_multi, const int *count_multi, double *powder) {
    memset(powder, 0, num_pix * sizeof(double));
    for (int d = 0; d < num_data; ++d) {
        for (int k = 0; k < ones[d]; ++k) {
            const int t = place_ones[ones_accum[d] + k];
            if (mask[t] < 1) powder[t]++;
        }
        for (int k = 0; k < multi[d]; ++k) {
            const int t = place_multi[multi_accum[d] + k];
            if (mask[t] < 1) powder[t] += count_multi[multi_accum[d] + k];
        }
    }
    for (int t = 0; t < num_pix; ++t) powder[t] /= num_data;
}

/* gradient_rt (max_update.c:232-341) for one orientation; pr[d] is the frame's posterior at
 * it (0 where not kept). Pixel states: 0 to optimise, 1 no photon seen, 128 negligible
 * background, 160 find the lower bound, 192 bracketed, 255 done. */
static void orc_gradient_rt(const orc_problem *p, const double *pr, double psum_r, uint8_t *mask,
                            const double *W, double *grad) {
    for (int t = 0; t < p->num_pix; ++t) {
        if (mask[t] < 128) {
            grad[t] = -psum_r;
            mask[t] = 1;
        } else if (mask[t] == 160) {
            grad[t] = -1e100;
        } else {
            grad[t] = 0.;
        }
    }
    for (int d = 0; d < p->num_data; ++d) {
        if (p->blacklist[d]) continue;
        const double pv = pr[d];
        if (!(pv > ORC_PROB_MIN)) continue;
        const double sc = p->scale[d], bs = p->bgscale[d];
        for (int k = 0; k < p->ones[d]; ++k) {
            const int t = p->place_ones[p->ones_accum[d] + k];
            if (mask[t] < 128) {
                grad[t] += pv * sc / (W[t] * sc + bs * p->background[t]);
                mask[t] = 0;
            } else if (mask[t] == 128) {
                grad[t] += pv;
            } else if (mask[t] == 160) {
                grad[t] = fmax(grad[t], sc / bs);
            }
        }
        for (int k = 0; k < p->multi[d]; ++k) {
            const int t = p->place_multi[p->multi_accum[d] + k];
            const int c = p->count_multi[p->multi_accum[d] + k];
            if (mask[t] < 128) {
                grad[t] += pv * c * sc / (W[t] * sc + bs * p->background[t]);
                mask[t] = 0;
            } else if (mask[t] == 128) {
                grad[t] += pv * c;
            } else if (mask[t] == 160) {
                grad[t] = fmax(grad[t], sc / bs);
            }
        }
    }
}

/* update_tomogram_bg (max_update.c:343-543): per pixel, the root in W of the gradient of
 * the background-aware log-likelihood, bracketed then found by Ridder's method. */
static void orc_update_tomogram_bg(const orc_problem *p, const double *pr, double psum_r, double scalemax,
                                   double *view) {
    const int P = p->num_pix;
    const double grad_tol = 1.e-4;
    uint8_t *mask = malloc(P);
    double *w = calloc(8 * (size_t)P, sizeof(double));
    double *W_old = w, *W_new = w + P, *W_mid = w + 2 * P, *W_latest = w + 3 * P;
    double *G_old = w + 4 * P, *G_new = w + 5 * P, *G_mid = w + 6 * P, *G_latest = w + 7 * P;
    memset(view, 0, P * sizeof(double));
    for (int t = 0; t < P; ++t) {
        if (p->mask[t] < 2) {
            if (p->background[t] / scalemax < 1.e-2 * p->powder[t]) {
                mask[t] = 128;
            } else {
                W_new[t] = 1.e-8 - p->background[t] / scalemax;
                mask[t] = 0;
            }
        } else {
            mask[t] = 255;
        }
    }
    orc_gradient_rt(p, pr, psum_r, mask, W_old, G_old);
    for (int t = 0; t < P; ++t) {
        if (fabs(G_old[t]) < grad_tol) {
            view[t] = 0.;
            mask[t] = 255;
        } else if (G_old[t] > 0.) {
            const double val = p->mean_count / P;
            W_new[t] = p->powder[t] > val ? p->powder[t] : val;
            W_new[t] *= 2.;
        }
    }
    int i, nmask = 0;
    for (i = 0; i < 10; ++i) {
        nmask = 0;
        orc_gradient_rt(p, pr, psum_r, mask, W_new, G_new);
        for (int t = 0; t < P; ++t) {
            if (mask[t] == 0) {
                if (G_old[t] < 0. && G_new[t] < 0.) {
                    mask[t] = 160;
                } else if (G_old[t] > 0. && G_new[t] > 0.) {
                    W_new[t] *= 2;
                } else {
                    mask[t] = 192;
                    nmask++;
                }
            } else if (mask[t] == 160) {
                W_new[t] = 1.e-8 - p->background[t] / G_new[t];
                mask[t] = 0;
            } else {
                nmask++;
            }
        }
        if (nmask == P) break;
    }
    for (i = 0; i < 50; ++i) {
        nmask = 0;
        for (int t = 0; t < P; ++t) {
            if (mask[t] == 255) {
                nmask++;
            } else if (mask[t] == 1) {
                mask[t] = 255;
                view[t] = 0.;
                nmask++;
            } else if (mask[t] == 128) {
                view[t] = G_old[t] / psum_r;
                mask[t] = 255;
                nmask++;
            } else if (mask[t] == 192) {
                mask[t] = 0;
                W_mid[t] = 0.5 * (W_old[t] + W_new[t]);
            } else if (mask[t] == 0) {
                W_mid[t] = 0.5 * (W_old[t] + W_new[t]);
            }
        }
        if (nmask == P) break;
        orc_gradient_rt(p, pr, psum_r, mask, W_mid, G_mid);
        for (int t = 0; t < P; ++t) {
            if (mask[t] == 1) {
                mask[t] = 255;
                view[t] = 0.;
                nmask++;
            } else if (mask[t] == 0) {
                if (fabs(G_mid[t]) < grad_tol) {
                    view[t] = W_mid[t];
                    mask[t] = 255;
                    nmask++;
                } else {
                    const double sq = sqrt(G_mid[t] * G_mid[t] - G_old[t] * G_new[t]);
                    if (sq < 1.e-15) {
                        W_latest[t] = W_mid[t];
                    } else {
                        W_latest[t] = W_mid[t] + (W_mid[t] - W_old[t]) * copysign(1.0, G_old[t]) * G_mid[t] / sq;
                        const double lo = fmin(W_old[t], W_new[t]), hi = fmax(W_old[t], W_new[t]);
                        if (W_latest[t] <= lo || W_latest[t] >= hi) W_latest[t] = W_mid[t];
                    }
                }
            }
        }
        if (nmask == P) break;
        orc_gradient_rt(p, pr, psum_r, mask, W_latest, G_latest);
        for (int t = 0; t < P; ++t) {
            if (mask[t] == 1) {
                mask[t] = 255;
                view[t] = 0.;
                nmask++;
            } else if (mask[t] == 0) {
                if (fabs(G_latest[t]) < grad_tol) {
                    view[t] = W_latest[t];
                    mask[t] = 255;
                    nmask++;
                } else if (G_mid[t] * G_latest[t] < 0.) {
                    W_old[t] = W_mid[t];
                    G_old[t] = G_mid[t];
                    W_new[t] = W_latest[t];
                    G_new[t] = G_latest[t];
                } else if (G_old[t] * G_latest[t] < 0.) {
                    W_new[t] = W_latest[t];
                    G_new[t] = G_latest[t];
                } else {
                    W_old[t] = W_latest[t];
                    G_old[t] = G_latest[t];
                }
            }
        }
        if (nmask == P) break;
    }
    free(mask);
    free(w);
}

/* The view update_tomogram_bg finds for one orientation (pr: every frame's posterior at it,
 * psum_r and scalemax as calc_psum_r gives them); for the tests' root checks. */
void orc_tomogram_bg_view(const orc_problem *p, const double *pr, double psum_r, double scalemax, double *view) {
    orc_update_tomogram_bg(p, pr, psum_r, scalemax, view);
}

/* gradient_d (max_scale.c:21-132): for every frame with mask 0, the derivative in its scale
 * of the background-aware log-likelihood, summed over its posterior (P > 1e-6) and photons
 * (mask < 1), minus psum_d. views[r] = slice_gen * rescale of combined orientation r. */
static void orc_gradient_d(const orc_problem *p, int Rn, const double *views, const double *prob,
                           const double *psum_d, const uint8_t *mask, const double *scale, double *grad) {
    const int N = p->num_data, P = p->num_pix;
    for (int d = 0; d < N; ++d)
        if (mask[d] == 0) grad[d] = 0.;
    for (int d = 0; d < N; ++d) {
        if (mask[d] > 0) continue;
        double acc = 0.;
        for (int r = 0; r < Rn; ++r) {
            const double pv = prob[(size_t)d * Rn + r];
            if (!(pv > ORC_PROB_MIN)) continue;
            const double *view = views + (size_t)r * P;
            for (int k = 0; k < p->ones[d]; ++k) {
                const int t = p->place_ones[p->ones_accum[d] + k];
                if (p->mask[t] < 1)
                    acc += pv * view[t] / (view[t] * scale[d] + p->bgscale[d] * p->background[t]);
            }
            for (int k = 0; k < p->multi[d]; ++k) {
                const int t = p->place_multi[p->multi_accum[d] + k];
                if (p->mask[t] < 1)
                    acc += pv * p->count_multi[p->multi_accum[d] + k] * view[t] /
                           (view[t] * scale[d] + p->bgscale[d] * p->background[t]);
            }
        }
        grad[d] += acc;
    }
    for (int d = 0; d < N; ++d)
        if (mask[d] == 0) grad[d] -= psum_d[d];
}

/* update_scale_bg (max_scale.c:134-288): per frame, the root in the scale of gradient_d,
 * bracketed (with the 99% early stop over all frames) then found by Ridder's method. */
static void orc_update_scale_bg(const orc_problem *p, int Rn, const double *views, const double *prob,
                                const double *psum_d, double *scale) {
    const int N = p->num_data;
    double *w = calloc(8 * (size_t)N, sizeof(double));
    double *s_old = w, *s_new = w + N, *s_mid = w + 2 * N, *s_latest = w + 3 * N;
    double *G_old = w + 4 * N, *G_new = w + 5 * N, *G_mid = w + 6 * N, *G_latest = w + 7 * N;
    uint8_t *mask = calloc(N, 1);
    for (int d = 0; d < N; ++d) {
        if (p->blacklist[d]) {
            mask[d] = 255;
            scale[d] = -1.;
            continue;
        }
        s_new[d] = (p->iteration > 1 || p->known_scale) ? 4. * scale[d] : 32.;
    }
    orc_gradient_d(p, Rn, views, prob, psum_d, mask, s_old, G_old);
    for (int d = 0; d < N; ++d)
        if ((mask[d] != 255) & (G_old[d] < 1.e-6)) {
            scale[d] = 0.001;
            mask[d] = 255;
        }
    int i, num_mask = 0;
    for (i = 0; i < 5; ++i) {
        orc_gradient_d(p, Rn, views, prob, psum_d, mask, s_new, G_new);
        num_mask = 0;
        for (int d = 0; d < N; ++d) {
            if (mask[d] != 0) {
                num_mask++;
            } else if (G_new[d] < 0.) {
                mask[d] = 192;
                num_mask++;
            } else {
                s_new[d] *= 4.;
            }
        }
        if (num_mask > ((double)0.99 * N)) break;
    }
    for (i = 0; i < 50; ++i) {
        num_mask = 0;
        for (int d = 0; d < N; ++d) {
            if (mask[d] == 255) {
                num_mask++;
            } else if (fabs(s_old[d] - s_new[d]) < 1.e-5) {
                mask[d] = 255;
                scale[d] = 0.5 * (s_old[d] + s_new[d]);
                num_mask++;
            } else if (mask[d] == 192) {
                mask[d] = 0;
                s_mid[d] = 0.5 * (s_old[d] + s_new[d]);
            } else {
                s_mid[d] = 0.5 * (s_old[d] + s_new[d]);
            }
        }
        if (num_mask == N) break;
        orc_gradient_d(p, Rn, views, prob, psum_d, mask, s_mid, G_mid);
        for (int d = 0; d < N; ++d) {
            if (mask[d] == 255) continue;
            if (fabs(G_mid[d]) < 1.e-3) {
                scale[d] = s_mid[d];
                mask[d] = 255;
                num_mask++;
                continue;
            }
            const double sq = sqrt(G_mid[d] * G_mid[d] - G_old[d] * G_new[d]);
            if (sq < 1.e-15) {
                s_latest[d] = s_mid[d];
            } else {
                s_latest[d] = s_mid[d] + (s_mid[d] - s_old[d]) * copysign(1.0, G_old[d]) * G_mid[d] / sq;
                const double lo = fmin(s_old[d], s_new[d]), hi = fmax(s_old[d], s_new[d]);
                if (s_latest[d] <= lo || s_latest[d] >= hi) s_latest[d] = s_mid[d];
            }
        }
        if (num_mask == N) break;
        orc_gradient_d(p, Rn, views, prob, psum_d, mask, s_latest, G_latest);
        for (int d = 0; d < N; ++d)
            if (mask[d] < 255) {
                if (fabs(G_latest[d]) < 1.e-3) {
                    scale[d] = s_latest[d];
                    mask[d] = 255;
                    num_mask++;
                } else if (G_mid[d] * G_latest[d] < 0.) {
                    s_old[d] = s_mid[d];
                    G_old[d] = G_mid[d];
                    s_new[d] = s_latest[d];
                    G_new[d] = G_latest[d];
                } else if (G_old[d] * G_latest[d] < 0.) {
                    s_new[d] = s_latest[d];
                    G_new[d] = G_latest[d];
                } else {
                    s_old[d] = s_latest[d];
                    G_old[d] = G_latest[d];
                }
            }
        if (num_mask == N) break;
    }
    free(w);
    free(mask);
}

double orc_maximize(const orc_problem *p, orc_result *res) {
    /* combined orientations r = rotind * M + mode (max_prob.c:80-84) */
    const int M = p->num_modes > 1 ? p->num_modes : 1;
    const int N = p->num_data, Rn = p->num_rot * M, P = p->num_pix;
    const long vol1 = p->size * p->size * p->size, vol = vol1 * M;
    const int want_psum_d = p->need_scaling && p->update_scale;
    /* background-aware updates only with need_scaling (max_prob.c:94, max_update.c:11) */
    const int bg_on = p->with_bg && p->need_scaling;
    double t0 = orc_now(), t1;
    int nthreads = p->num_threads > 0 ? p->num_threads : 1;

    /* allocate_memory (maximize.c:58-137) */
    memset(p->model2, 0, vol * sizeof(double));
    memset(p->inter_weight, 0, vol * sizeof(double));
    memset(res->u, 0, Rn * sizeof(double));
    double *max_exp_p = malloc(N * sizeof(double));
    for (int d = 0; d < N; ++d) {
        max_exp_p[d] = ORC_MAX_EXP_START;
        res->rmax[d] = 0;
        res->likelihood[d] = 0.;
        res->info[d] = 0.;
        res->p_norm[d] = 0.;
        res->max_exp[d] = 0.;
        if (want_psum_d) res->psum_d[d] = 0.;
    }
    /* the background-aware scale update reads every frame's posterior after the merge */
    double *pdense = res->prob_dense;
    if (!pdense && bg_on && want_psum_d) pdense = malloc((size_t)N * Rn * sizeof(double));
    if (pdense) memset(pdense, 0, (size_t)N * Rn * sizeof(double));
    if (res->quat_norm) memset(res->quat_norm, 0, (size_t)N * M * sizeof(double));

    /* calculate_rescale (max_prob.c:3-64) */
    double total = 0.;
    double orient_s = 0., loop_s = 0.;  /* thread-seconds: per-orientation work, whole loops */
#pragma omp parallel num_threads(nthreads) default(shared)
    {
        double priv_total = 0.;
        double *view = malloc(P * sizeof(double));
        const double l0 = orc_now();
#pragma omp for schedule(static, 1) nowait
        for (int r = 0; r < Rn; ++r) {
            const int rotind = r / M, mode = r % M;
            orc_slice_gen3d(&p->quats[rotind * 5], p->model1 + mode * vol1, p->size, P, p->qvals, p->corr,
                            view);
            for (int t = 0; t < P; ++t)
                if (p->mask[t] < 1) res->u[r] += view[t];
            priv_total += p->quats[rotind * 5 + 4] * res->u[r];
        }
        const double dl = orc_now() - l0;
#pragma omp critical(orc_total)
        {
            total += priv_total;
            orient_s += dl;
            loop_s += dl;
        }
        free(view);
    }
    res->rescale = p->mean_count / total * M;
    t1 = orc_now();
    res->seconds[0] = t1 - t0;

    double t_prob = 0, t_norm = 0;
#pragma omp parallel num_threads(nthreads) default(shared)
    {
        int tid = omp_get_thread_num();
        orc_keep *keep = calloc(N, sizeof(orc_keep));
        double *pmax = malloc(N * sizeof(double));
        int *prmax = calloc(N, sizeof(int));
        double *view = malloc(P * sizeof(double));
        double *pmodel = calloc(vol, sizeof(double));
        double *pweight = calloc(vol, sizeof(double));
        double *plike = calloc(N, sizeof(double));
        double *pinfo = calloc(N, sizeof(double));
        double *ppsum_d = calloc(N, sizeof(double));
        double *pnorm = calloc(N, sizeof(double));
        double *pqnorm = calloc((size_t)N * M, sizeof(double));
        double *pr = bg_on ? calloc(N, sizeof(double)) : NULL;
        for (int d = 0; d < N; ++d) pmax[d] = ORC_MAX_EXP_START;

        /* calculate_prob (max_prob.c:66-213) */
        double my_orient = 0., my_loop = 0., lt = orc_now();
#pragma omp for schedule(static, 1)
        for (int r = 0; r < Rn; ++r) {
            const int rotind = r / M, mode = r % M;
            const double ot = orc_now();
            orc_slice_gen3d(&p->quats[rotind * 5], p->model1 + mode * vol1, p->size, P, p->qvals, p->corr,
                            view);
            if (!bg_on)
                for (int t = 0; t < P; ++t) view[t] = log(view[t]);
            my_orient += orc_now() - ot;
            double logw = log(p->quats[rotind * 5 + 4]);
            for (int d = 0; d < N; ++d) {
                if (p->blacklist[d]) continue;
                if (p->quat_mapping) {
                    int found = 0;
                    for (long k = p->rel_off[d]; k < p->rel_off[d + 1] && !found; ++k)
                        found = p->rel_quat[k] == p->quat_mapping[r];
                    if (!found) continue;
                }
                double pval;
                if (p->need_scaling && (p->iteration > 1 || p->known_scale))
                    pval = logw - res->u[r] * p->scale[d];
                else
                    pval = logw - res->u[r] * res->rescale;
                if (p->ftype == 0 && bg_on) {
                    /* max_prob.c:134-156: log of scale * W + bgscale * background */
                    const int *po = &p->place_ones[p->ones_accum[d]];
                    for (int k = 0; k < p->ones[d]; ++k)
                        if (p->mask[po[k]] < 1) {
                            double ip = view[po[k]] * p->scale[d] + p->bgscale[d] * p->background[po[k]];
                            if (ip < 0) ip = DBL_MIN;
                            pval += log(ip);
                        }
                    const int *pm = &p->place_multi[p->multi_accum[d]];
                    const int *cm = &p->count_multi[p->multi_accum[d]];
                    for (int k = 0; k < p->multi[d]; ++k)
                        if (p->mask[pm[k]] < 1) {
                            double ip = view[pm[k]] * p->scale[d] + p->bgscale[d] * p->background[pm[k]];
                            if (ip < 0) ip = DBL_MIN;
                            pval += cm[k] * log(ip);
                        }
                } else if (p->ftype == 0) {
                    const int *po = &p->place_ones[p->ones_accum[d]];
                    for (int k = 0; k < p->ones[d]; ++k)
                        if (p->mask[po[k]] < 1) pval += view[po[k]];
                    const int *pm = &p->place_multi[p->multi_accum[d]];
                    const int *cm = &p->count_multi[p->multi_accum[d]];
                    for (int k = 0; k < p->multi[d]; ++k)
                        if (p->mask[pm[k]] < 1) pval += cm[k] * view[pm[k]];
                } else {
                    const int *fr = &p->int_frames[(size_t)d * P];
                    for (int t = 0; t < P; ++t)
                        if (p->mask[t] < 1) pval += fr[t] * view[t];
                }
                if (pval + ORC_PDIFF_THRESH / p->beta[d] > pmax[d]) orc_keep_push(&keep[d], pval, r);
                if (pval > pmax[d]) {
                    pmax[d] = pval;
                    prmax[d] = r;
                    orc_keep_prune(&keep[d], pval - ORC_PDIFF_THRESH / p->beta[d]);
                }
            }
        }
        my_loop += orc_now() - lt;
        if (tid == 0) t_prob = orc_now();

        /* normalize_prob (max_prob.c:215-265) */
#pragma omp critical(orc_maxexp)
        for (int d = 0; d < N; ++d)
            if (pmax[d] > max_exp_p[d]) {
                max_exp_p[d] = pmax[d];
                res->rmax[d] = prmax[d];
            }
#pragma omp barrier
#pragma omp single
        for (int d = 0; d < N; ++d) {
            res->max_exp[d] = max_exp_p[d];
            if (max_exp_p[d] == ORC_MAX_EXP_START) res->rmax[d] = -1;
        }
        for (int d = 0; d < N; ++d)
            for (int i = 0; i < keep[d].n; ++i)
                pnorm[d] += exp(p->beta[d] * (keep[d].val[i] - res->max_exp[d]));
#pragma omp critical(orc_psum)
        for (int d = 0; d < N; ++d) res->p_norm[d] += pnorm[d];
#pragma omp barrier
        if (tid == 0) t_norm = orc_now();

        /* update_tomogram + merge_tomogram (max_update.c:3-230) */
        lt = orc_now();
#pragma omp for schedule(static, 1) nowait
        for (int r = 0; r < Rn; ++r) {
            const int rotind = r / M, mode = r % M;
            double w_r = p->quats[rotind * 5 + 4], psum_r = 0.;
            double scalemax = -DBL_MAX;  /* calc_psum_r's return value (max_update.c:50,120-121) */
            if (pr) memset(pr, 0, N * sizeof(double));
            for (int d = 0; d < N; ++d) {
                if (p->blacklist[d]) continue;
                int ind = orc_keep_find(&keep[d], r);
                if (ind < 0) continue;
                double temp = keep[d].val[ind];
                double prob = exp(p->beta[d] * (temp - res->max_exp[d])) / res->p_norm[d];
                keep[d].val[ind] = prob;
                plike[d] += prob * (temp - p->sum_fact[d]);
                if (p->need_scaling) {
                    psum_r += prob * p->scale[d];
                    if (p->update_scale) ppsum_d[d] += prob * res->u[r] * res->rescale;
                } else {
                    psum_r += prob;
                }
                if (pdense) pdense[(size_t)d * Rn + r] = prob;
                if (pr) pr[d] = prob;
                if (prob < ORC_PROB_MIN) continue;
                if (M > 1) pqnorm[(size_t)d * M + mode] += prob;
                pinfo[d] += prob * log(prob / w_r * M);
                if (p->need_scaling && p->scale[d] > scalemax) scalemax = p->scale[d];
            }
            if (bg_on) {
                orc_update_tomogram_bg(p, pr, psum_r, scalemax, view);
                if (psum_r > 0.)
                    orc_slice_merge3d(&p->quats[rotind * 5], view, pmodel + mode * vol1, pweight + mode * vol1,
                                      p->size, P, p->qvals, p->corr, p->mask);
                continue;
            }
            memset(view, 0, P * sizeof(double));
            for (int d = 0; d < N; ++d) {
                if (p->blacklist[d]) continue;
                int ind = orc_keep_find(&keep[d], r);
                if (ind < 0) continue;
                double prob = keep[d].val[ind];
                if (!(prob > ORC_PROB_MIN)) continue;
                if (p->ftype == 0) {
                    const int *po = &p->place_ones[p->ones_accum[d]];
                    for (int k = 0; k < p->ones[d]; ++k)
                        if (p->mask[po[k]] < 2) view[po[k]] += prob;
                    const int *pm = &p->place_multi[p->multi_accum[d]];
                    const int *cm = &p->count_multi[p->multi_accum[d]];
                    for (int k = 0; k < p->multi[d]; ++k)
                        if (p->mask[pm[k]] < 2) view[pm[k]] += cm[k] * prob;
                } else {
                    const int *fr = &p->int_frames[(size_t)d * P];
                    for (int t = 0; t < P; ++t) view[t] += fr[t] * prob;
                }
            }
            if (psum_r > 0.) {
                const double ot = orc_now();
                for (int t = 0; t < P; ++t) view[t] /= psum_r;
                orc_slice_merge3d(&p->quats[rotind * 5], view, pmodel + mode * vol1, pweight + mode * vol1,
                                  p->size, P, p->qvals, p->corr, p->mask);
                my_orient += orc_now() - ot;
            }
        }
        my_loop += orc_now() - lt;
#pragma omp critical(orc_timing)
        {
            orient_s += my_orient;
            loop_s += my_loop;
        }
#pragma omp barrier

        /* combine_information_omp (max_prob.c:267-342) */
#pragma omp critical(orc_model)
        for (long x = 0; x < vol; ++x) {
            p->model2[x] += pmodel[x];
            p->inter_weight[x] += pweight[x];
        }
#pragma omp critical(orc_like_info)
        for (int d = 0; d < N; ++d) {
            res->likelihood[d] += plike[d];
            res->info[d] += pinfo[d];
        }
        if (M > 1 && res->quat_norm) {
#pragma omp critical(orc_quat_norm)
            for (size_t k = 0; k < (size_t)N * M; ++k) res->quat_norm[k] += pqnorm[k];
        }
        if (want_psum_d) {
#pragma omp critical(orc_scale)
            for (int d = 0; d < N; ++d)
                if (!p->blacklist[d]) res->psum_d[d] += ppsum_d[d];
        }
        for (int d = 0; d < N; ++d) {
            free(keep[d].val);
            free(keep[d].pos);
        }
        free(keep); free(pmax); free(prmax); free(view); free(pmodel); free(pweight);
        free(plike); free(pinfo); free(ppsum_d); free(pnorm); free(pqnorm); free(pr);
    }
    double t2 = orc_now();

    /* combine_information_mpi (max_prob.c:344-427), one rank */
    double avg = 0., mi = 0.;
    for (int d = 0; d < N; ++d)
        if (!p->blacklist[d]) {
            mi += res->info[d];
            avg += res->likelihood[d];
        }
    mi /= (N - p->num_blacklist);
    avg /= (N - p->num_blacklist);
    res->mutual_info = mi;
    res->avg_likelihood = avg;

    /* update_scale (max_scale.c:3-19) */
    if (want_psum_d && p->with_bg) {
        /* update_scale_bg takes the background whenever the detector has one (max_scale.c:9) */
        double *views = malloc((size_t)Rn * P * sizeof(double));
        for (int r = 0; r < Rn; ++r) {
            orc_slice_gen3d(&p->quats[(r / M) * 5], p->model1 + (r % M) * vol1, p->size, P, p->qvals, p->corr,
                            views + (size_t)r * P);
            for (int t = 0; t < P; ++t) views[(size_t)r * P + t] *= res->rescale;
        }
        orc_update_scale_bg(p, Rn, views, pdense, res->psum_d, p->scale);
        free(views);
    } else if (want_psum_d) {
        for (int d = 0; d < N; ++d)
            if (!p->blacklist[d]) p->scale[d] = p->fcounts[d] / res->psum_d[d];
    }
    if (pdense != res->prob_dense) free(pdense);
    free(max_exp_p);

    res->seconds[1] = t_prob - t1;
    res->seconds[2] = t_norm - t_prob;
    res->seconds[3] = t2 - t_norm;
    res->seconds[4] = orc_now() - t0;
    res->seconds[5] = orient_s;
    res->seconds[6] = loop_s;
    return avg;
}

/* EMCRecon.update_model, Cython order: symmetrize each mode, then divide, rms and blend over
 * every mode's volume (recon.pyx:112-149) */
double orc_update_model_modes(double *model1, double *model2, const double *w, long size, double alpha,
                              int num_modes) {
    const long vol1 = size * size * size, vol = vol1 * num_modes;
    for (int m = 0; m < num_modes; ++m) orc_symmetrize_friedel(model2 + m * vol1, w + m * vol1, (int)size);
    for (long x = 0; x < vol; ++x)
        if (w[x] > 0.) model2[x] /= w[x];
    double change = 0.;
    for (long x = 0; x < vol; ++x) {
        double diff = model2[x] - model1[x];
        change += diff * diff;
        if (alpha > 0.)
            model1[x] = alpha * model1[x] + (1 - alpha) * model2[x];
        else
            model1[x] = model2[x];
    }
    return sqrt(change / num_modes / vol1);
}

double orc_update_model(double *model1, double *model2, const double *w, long size, double alpha) {
    return orc_update_model_modes(model1, model2, w, size, alpha, 1);
}
