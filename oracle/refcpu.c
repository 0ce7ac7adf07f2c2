/*
 * refcpu.c -- ORACLE / TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the float32 loops of fullrmc's hot-path Cython
 * extensions.  It is the checker the HIP path is compared against and the
 * CPU baseline timed beside the GPU; nothing in the product path links it.
 *
 * Semantics followed (reference files, relative to the fullrmc tree):
 *   round()                          Extensions/pairs_distances.pyx:31-32
 *   PBC distance, atom vs all        Extensions/pairs_distances.pyx:343-386
 *   IBC distance, atom vs all        Extensions/pairs_distances.pyx:438-471
 *   PBC/IBC distance, point vs all   Extensions/pairs_distances.pyx:288-340, 410-433
 *   differences (point / index)      Extensions/pairs_distances.pyx:135-253
 *   from-to differences              Extensions/pairs_distances.pyx:73-132
 *   _single_pairs_histograms         Extensions/pairs_histograms.pyx:36-68
 *   multiple_pairs_histograms_coords Extensions/pairs_histograms.pyx:150-217
 *   multiple_pairs_histograms_dists  Extensions/pairs_histograms.pyx:225-281
 *   get_reciprocal_basis (with the   Extensions/boundary_conditions_collection.pyx:45-80
 *     basis[0,1] typo at :59)
 *   transform_coordinates            Extensions/boundary_conditions_collection.pyx:88-110
 *   box_coordinates_real_distances   Extensions/boundary_conditions_collection.pyx:119-165
 *   _single_atomic_distances_dists   Extensions/atomic_distances.pyx:42-118
 *   multiple_atomic_distances_coords Extensions/atomic_distances.pyx:326-410
 *   multiple_atomic_distances_dists  Extensions/atomic_distances.pyx:419-491
 *   _pair_elements_stats             Extensions/atomic_distances.pyx:632-656
 *
 * Float semantics: every expression is evaluated left to right in IEEE
 * binary32 with round-to-nearest-even, no fused multiply-add (the recipe
 * builds with -ffp-contract=off and no -march), sqrt and floor/ceil through
 * libm's double functions exactly as the Cython-generated C does.
 *
 * Bin index quirk: the reference writes h[a,b,binIndex] with no bounds
 * check; binIndex can equal histSize for a distance just below maxDistance.
 * Inside the (nel,nel,H) C-ordered array that aliases h[a,b+1,0]; at the
 * very last slot it would write past the array.  This restatement keeps the
 * in-array aliasing (flat indexing) and counts past-the-end writes (and NaN
 * distances, whose int cast is undefined in C) in *overflow instead.
 *
 * Histograms are counted in int32 (all reference increments are +1.0f on
 * integer-valued float32 bins, exact below 2^24) and added to the caller's
 * float32 arrays at the end, matching the reference's arithmetic exactly.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline float ref_round(float num) {
    /* pairs_distances.pyx:31-32: floor(num+0.5) if num>0 else ceil(num-0.5);
       math.h floor/ceil take and return double: exact for float inputs. */
    return num > 0.0f ? (float)floor((double)(num + 0.5f)) : (float)ceil((double)(num - 0.5f));
}

static inline float ref_pbc_dist(float ax, float ay, float az, const float *c, const float *B) {
    float diff_x = ax - c[0];
    float diff_y = ay - c[1];
    float diff_z = az - c[2];
    float box_dx = diff_x - ref_round(diff_x);
    float box_dy = diff_y - ref_round(diff_y);
    float box_dz = diff_z - ref_round(diff_z);
    float real_dx = box_dx * B[0] + box_dy * B[3] + box_dz * B[6];
    float real_dy = box_dx * B[1] + box_dy * B[4] + box_dz * B[7];
    float real_dz = box_dx * B[2] + box_dy * B[5] + box_dz * B[8];
    return (float)sqrt((double)(real_dx * real_dx + real_dy * real_dy + real_dz * real_dz));
}

static inline float ref_ibc_dist(float ax, float ay, float az, const float *c) {
    float real_dx = ax - c[0];
    float real_dy = ay - c[1];
    float real_dz = az - c[2];
    return (float)sqrt((double)(real_dx * real_dx + real_dy * real_dy + real_dz * real_dz));
}

/* ---------------------------------------------------------------- distances */

/* pairs_distances_to_indexcoords (pairs_distances.pyx:874-916 -> 343-386 / 438-471).
   Entries [0,atomIndex) are left untouched when allAtoms is false. */
void ref_distances_to_indexcoords(int32_t atomIndex, const float *coords, int32_t n,
                                  const float *B, int isPBC, int allAtoms, float *out) {
    float ax = coords[3 * atomIndex], ay = coords[3 * atomIndex + 1], az = coords[3 * atomIndex + 2];
    int32_t start = allAtoms ? 0 : atomIndex;
    for (int32_t i = start; i < n; i++)
        out[i] = isPBC ? ref_pbc_dist(ax, ay, az, coords + 3 * i, B) : ref_ibc_dist(ax, ay, az, coords + 3 * i);
}

/* pairs_distances_to_point (pairs_distances.pyx:827-871 -> 288-340 / 410-433):
   diff = point - coords[i]. */
void ref_distances_to_point(const float *point, const float *coords, int32_t n,
                            const float *B, int isPBC, float *out) {
    for (int32_t i = 0; i < n; i++)
        out[i] = isPBC ? ref_pbc_dist(point[0], point[1], point[2], coords + 3 * i, B)
                       : ref_ibc_dist(point[0], point[1], point[2], coords + 3 * i);
}

/* point_to_point_distance (pairs_distances.pyx:786-824 -> 256-282): diff = p1 - p2 */
float ref_point_to_point_distance(const float *p1, const float *p2, const float *B, int isPBC) {
    return isPBC ? ref_pbc_dist(p1[0], p1[1], p1[2], p2, B) : ref_ibc_dist(p1[0], p1[1], p1[2], p2);
}

static inline void ref_pbc_diff(float ax, float ay, float az, const float *c, const float *B, float *o) {
    float diff_x = ax - c[0];
    float diff_y = ay - c[1];
    float diff_z = az - c[2];
    float box_dx = diff_x - ref_round(diff_x);
    float box_dy = diff_y - ref_round(diff_y);
    float box_dz = diff_z - ref_round(diff_z);
    o[0] = box_dx * B[0] + box_dy * B[3] + box_dz * B[6];
    o[1] = box_dx * B[1] + box_dy * B[4] + box_dz * B[7];
    o[2] = box_dx * B[2] + box_dy * B[5] + box_dz * B[8];
}

/* pairs_differences_to_indexcoords (pairs_distances.pyx:626-672 -> 195-253):
   out is (n,3); rows [0,atomIndex) untouched when allAtoms is false. */
void ref_differences_to_indexcoords(int32_t atomIndex, const float *coords, int32_t n,
                                    const float *B, int isPBC, int allAtoms, float *out) {
    float ax = coords[3 * atomIndex], ay = coords[3 * atomIndex + 1], az = coords[3 * atomIndex + 2];
    int32_t start = allAtoms ? 0 : atomIndex;
    for (int32_t i = start; i < n; i++) {
        if (isPBC) {
            ref_pbc_diff(ax, ay, az, coords + 3 * i, B, out + 3 * i);
        } else {
            out[3 * i] = ax - coords[3 * i];
            out[3 * i + 1] = ay - coords[3 * i + 1];
            out[3 * i + 2] = az - coords[3 * i + 2];
        }
    }
}

/* pairs_differences_to_point (pairs_distances.pyx:580-623 -> 135-188): diff = point - coords[i] */
void ref_differences_to_point(const float *point, const float *coords, int32_t n,
                              const float *B, int isPBC, float *out) {
    for (int32_t i = 0; i < n; i++) {
        if (isPBC) {
            ref_pbc_diff(point[0], point[1], point[2], coords + 3 * i, B, out + 3 * i);
        } else {
            out[3 * i] = point[0] - coords[3 * i];
            out[3 * i + 1] = point[1] - coords[3 * i + 1];
            out[3 * i + 2] = point[2] - coords[3 * i + 2];
        }
    }
}

/* from_to_points_differences (pairs_distances.pyx:483-527 -> 73-132): diff = to[i] - from[i] */
void ref_from_to_points_differences(const float *from, const float *to, int32_t n,
                                    const float *B, int isPBC, float *out) {
    for (int32_t i = 0; i < n; i++) {
        if (isPBC) {
            ref_pbc_diff(to[3 * i], to[3 * i + 1], to[3 * i + 2], from + 3 * i, B, out + 3 * i);
        } else {
            out[3 * i] = to[3 * i] - from[3 * i];
            out[3 * i + 1] = to[3 * i + 1] - from[3 * i + 1];
            out[3 * i + 2] = to[3 * i + 2] - from[3 * i + 2];
        }
    }
}

/* pair_difference_to_point (pairs_distances.pyx:534-577 -> 45-63 / 108-114):
   PBC diff = point2 - point1 (box-rounded); IBC diff = point2 - point1. */
void ref_pair_difference(const float *p1, const float *p2, const float *B, int isPBC, float *out) {
    if (isPBC) {
        ref_pbc_diff(p2[0], p2[1], p2[2], p1, B, out);
    } else {
        out[0] = p2[0] - p1[0];
        out[1] = p2[1] - p1[1];
        out[2] = p2[2] - p1[2];
    }
}

/* --------------------------------------------------------------- histograms */

static inline void ref_bin_one(float distance, int32_t atomSym, int32_t atomMol, int32_t molI, int32_t elemI,
                               int32_t nel, int32_t H, float minD, float maxD, float bin,
                               int32_t *cintra, int32_t *cinter, int64_t *overflow) {
    /* pairs_histograms.pyx:57-68 */
    if (distance < minD) return;
    if (distance >= maxD) return;
    if (distance != distance) { (*overflow)++; return; }
    int32_t binIndex = (int32_t)((distance - minD) / bin);
    int64_t flat = ((int64_t)atomSym * nel + elemI) * H + binIndex;
    if (flat < 0 || flat >= (int64_t)nel * nel * H) { (*overflow)++; return; }
    if (molI == atomMol) cintra[flat] += 1;
    else                 cinter[flat] += 1;
}

/* single_pairs_histograms (pairs_histograms.pyx:77-141): accumulates into
   float32 hintra/hinter in place. */
int64_t ref_single_pairs_histograms(int32_t atomIndex, const float *distances, int32_t n,
                                    const int32_t *mol, const int32_t *elem, int32_t nel, int32_t H,
                                    float *hintra, float *hinter, float minD, float maxD, float bin,
                                    int allAtoms) {
    int64_t size = (int64_t)nel * nel * H, overflow = 0;
    int32_t *ci = (int32_t *)calloc(size, sizeof(int32_t));
    int32_t *ce = (int32_t *)calloc(size, sizeof(int32_t));
    int32_t atomMol = mol[atomIndex], atomSym = elem[atomIndex];
    int32_t start = allAtoms ? 0 : atomIndex;
    for (int32_t i = start; i < n; i++) {
        if (i == atomIndex) continue;
        ref_bin_one(distances[i], atomSym, atomMol, mol[i], elem[i], nel, H, minD, maxD, bin, ci, ce, &overflow);
    }
    for (int64_t k = 0; k < size; k++) {
        hintra[k] += (float)ci[k];
        hinter[k] += (float)ce[k];
    }
    free(ci);
    free(ce);
    return overflow;
}

/* multiple_pairs_histograms_coords (pairs_histograms.pyx:150-217): outputs are
   zeroed (nel,nel,H) float32 arrays; for each i in indexes, in order, the
   distances of a3 then the binning of a4.  full_pairs_histograms_coords
   (:289-335) is this with indexes = arange(n) and allAtoms false.
   nthreads > 1 splits the index list over OpenMP threads with private
   histograms (the reference's shared '+=' in prange is racy). */
int64_t ref_multiple_pairs_histograms_coords(const int32_t *indexes, int32_t nidx, const float *coords,
                                             int32_t n, const float *B, int isPBC, const int32_t *mol,
                                             const int32_t *elem, int32_t nel, float minD, float maxD,
                                             float bin, int32_t H, int allAtoms, float *hintra,
                                             float *hinter, int nthreads) {
    int64_t size = (int64_t)nel * nel * H, overflow = 0;
    int32_t *ci = (int32_t *)calloc(size, sizeof(int32_t));
    int32_t *ce = (int32_t *)calloc(size, sizeof(int32_t));
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(+ : overflow)
    {
        float *dist = (float *)malloc(sizeof(float) * (n > 0 ? n : 1));
        int32_t *pi = ci, *pe = ce;
#ifdef _OPENMP
        if (omp_get_num_threads() > 1) {
            pi = (int32_t *)calloc(size, sizeof(int32_t));
            pe = (int32_t *)calloc(size, sizeof(int32_t));
        }
#pragma omp for schedule(dynamic, 16)
#endif
        for (int32_t k = 0; k < nidx; k++) {
            int32_t a = indexes[k];
            ref_distances_to_indexcoords(a, coords, n, B, isPBC, allAtoms, dist);
            int32_t atomMol = mol[a], atomSym = elem[a];
            int32_t start = allAtoms ? 0 : a;
            for (int32_t i = start; i < n; i++) {
                if (i == a) continue;
                ref_bin_one(dist[i], atomSym, atomMol, mol[i], elem[i], nel, H, minD, maxD, bin, pi, pe, &overflow);
            }
        }
#ifdef _OPENMP
        if (pi != ci) {
#pragma omp critical
            {
                for (int64_t s = 0; s < size; s++) {
                    ci[s] += pi[s];
                    ce[s] += pe[s];
                }
            }
            free(pi);
            free(pe);
        }
#endif
        free(dist);
    }
    for (int64_t s = 0; s < size; s++) {
        hintra[s] = (float)ci[s];
        hinter[s] = (float)ce[s];
    }
    free(ci);
    free(ce);
    return overflow;
}

/* multiple_pairs_histograms_dists (pairs_histograms.pyx:225-281): distances is
   (n,k) C-ordered; column c is the distance array of atom indexes[c]. */
int64_t ref_multiple_pairs_histograms_dists(const int32_t *indexes, int32_t nidx, const float *distances,
                                            int32_t n, int32_t k, const int32_t *mol, const int32_t *elem,
                                            int32_t nel, float minD, float maxD, float bin, int32_t H,
                                            int allAtoms, float *hintra, float *hinter) {
    int64_t size = (int64_t)nel * nel * H, overflow = 0;
    int32_t *ci = (int32_t *)calloc(size, sizeof(int32_t));
    int32_t *ce = (int32_t *)calloc(size, sizeof(int32_t));
    for (int32_t c = 0; c < nidx; c++) {
        int32_t a = indexes[c];
        int32_t atomMol = mol[a], atomSym = elem[a];
        int32_t start = allAtoms ? 0 : a;
        for (int32_t i = start; i < n; i++) {
            if (i == a) continue;
            ref_bin_one(distances[(int64_t)i * k + c], atomSym, atomMol, mol[i], elem[i], nel, H, minD, maxD,
                        bin, ci, ce, &overflow);
        }
    }
    for (int64_t s = 0; s < size; s++) {
        hintra[s] = (float)ci[s];
        hinter[s] = (float)ce[s];
    }
    free(ci);
    free(ce);
    return overflow;
}

/* Float64 recompute of the full histogram: every unordered pair i<j once,
   slot [e(i)][e(j)], with the distance evaluated in binary64 from the float32
   box coordinates and basis (differences, minimum-image rounding
   x>0 ? floor(x+0.5) : ceil(x-0.5), basis product ((bx*B0k + by*B1k) + bz*B2k),
   ((rx*rx + ry*ry) + rz*rz), correctly rounded sqrt, no FMA), range
   minD <= d < maxD and bin (int)((d - minD)/bin) in binary64, the flat
   (nel,nel,H) index aliasing like the reference's.  counts: int32
   [2][nel][nel][H] (zeroed here); returns the number of past-the-end bins.
   This is the second oracle: a GPU edge re-check mode must equal it exactly. */
static inline double ref_round64(double x) { return x > 0.0 ? floor(x + 0.5) : ceil(x - 0.5); }

int64_t ref_full_pairs_counts_f64(const float *coords, int32_t n, const float *B, int isPBC, const int32_t *mol,
                                  const int32_t *elem, int32_t nel, float minD, float maxD, float bin, int32_t H,
                                  int32_t *counts, int nthreads) {
    const int64_t size = (int64_t)nel * nel * H;
    int64_t overflow = 0;
    double Bd[9];
    for (int k = 0; k < 9; k++) Bd[k] = (double)B[k];
    const double mn = minD, mx = maxD, bn = bin;
    memset(counts, 0, sizeof(int32_t) * 2 * size);
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(+ : overflow)
    {
        int32_t *pc = counts;
#ifdef _OPENMP
        if (omp_get_num_threads() > 1) pc = (int32_t *)calloc(2 * size, sizeof(int32_t));
#pragma omp for schedule(dynamic, 16)
#endif
        for (int32_t i = 0; i < n; i++) {
            for (int32_t j = i + 1; j < n; j++) {
                double dx = (double)coords[3 * (int64_t)i] - (double)coords[3 * (int64_t)j];
                double dy = (double)coords[3 * (int64_t)i + 1] - (double)coords[3 * (int64_t)j + 1];
                double dz = (double)coords[3 * (int64_t)i + 2] - (double)coords[3 * (int64_t)j + 2];
                double rx, ry, rz;
                if (isPBC) {
                    const double bx = dx - ref_round64(dx), by = dy - ref_round64(dy), bz = dz - ref_round64(dz);
                    rx = (bx * Bd[0] + by * Bd[3]) + bz * Bd[6];
                    ry = (bx * Bd[1] + by * Bd[4]) + bz * Bd[7];
                    rz = (bx * Bd[2] + by * Bd[5]) + bz * Bd[8];
                } else {
                    rx = dx; ry = dy; rz = dz;
                }
                const double d = sqrt((rx * rx + ry * ry) + rz * rz);
                if (d < mn || d >= mx) continue;
                const int64_t b = (int64_t)((d - mn) / bn);
                const int64_t flat = ((int64_t)elem[i] * nel + elem[j]) * H + b;
                if (flat >= size) {
                    overflow++;
                    continue;
                }
                pc[(mol[i] == mol[j] ? 0 : size) + flat]++;
            }
        }
#ifdef _OPENMP
        if (pc != counts) {
#pragma omp critical
            for (int64_t s = 0; s < 2 * size; s++) counts[s] += pc[s];
            free(pc);
        }
#endif
    }
    return overflow;
}

/* Cell-list form of the periodic full histogram, for large boxes: the same
   per-pair arithmetic on the same unordered pairs (row atom i = the earlier
   index, column j > i) as full_pairs_histograms_coords (f64 == 0, float32
   reference binning with ref_bin_one) or ref_full_pairs_counts_f64
   (f64 != 0), but a pair is evaluated only if the fractional cells of i and
   j are neighbours (index offsets -1, 0, +1 along every axis, periodic).
   Axis a gets m_a = floor(h_a / w) cells, h_a = V / |b x c| the spacing of
   the cell's lattice planes and w = maxD (1 + 1e-5) + 0.01 A.  Two atoms
   whose cells are not neighbours along axis a have a periodic fractional
   difference of at least 1/m_a there; any image vector r = b.B then has a
   component of at least h_a / m_a >= w along the plane normal, so its
   length is above maxD by more than the float32 rounding of the reference's
   difference, image shift and norm (about 1e-5 A at these sizes): such a
   pair is out of range in both arithmetics and counts nowhere.  counts:
   int32 [2][nel][nel][H] (zeroed here).  Returns the number of past-the-end
   bins, or -1 when an axis has fewer than 3 cells (the all-pairs loops must
   be used then). */
int64_t ref_full_pairs_counts_cells(const float *coords, int32_t n, const float *B, const int32_t *mol,
                                    const int32_t *elem, int32_t nel, float minD, float maxD, float bin, int32_t H,
                                    int f64, int32_t *counts, int nthreads) {
    const int64_t size = (int64_t)nel * nel * H;
    double Bd[9];
    for (int k = 0; k < 9; k++) Bd[k] = (double)B[k];
    /* plane spacings h_a = V / |row_b x row_c| */
    int32_t m[3];
    const double w = (double)maxD * (1.0 + 1e-5) + 0.01;
    const double vol = fabs(Bd[0] * (Bd[4] * Bd[8] - Bd[5] * Bd[7]) - Bd[1] * (Bd[3] * Bd[8] - Bd[5] * Bd[6]) +
                            Bd[2] * (Bd[3] * Bd[7] - Bd[4] * Bd[6]));
    for (int a = 0; a < 3; a++) {
        const double *u = Bd + 3 * ((a + 1) % 3), *v = Bd + 3 * ((a + 2) % 3);
        const double cx = u[1] * v[2] - u[2] * v[1], cy = u[2] * v[0] - u[0] * v[2], cz = u[0] * v[1] - u[1] * v[0];
        const double h = vol / sqrt(cx * cx + cy * cy + cz * cz);
        m[a] = (int32_t)floor(h / w);
        if (m[a] < 3) return -1;
    }
    const int64_t ncell = (int64_t)m[0] * m[1] * m[2];
    int64_t *start = (int64_t *)calloc(ncell + 1, sizeof(int64_t));
    int32_t *cell_of = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
    for (int32_t i = 0; i < n; i++) {
        int32_t c3[3];
        for (int a = 0; a < 3; a++) {
            const double x = (double)coords[3 * (int64_t)i + a];
            int32_t c = (int32_t)((x - floor(x)) * m[a]);
            c3[a] = c < 0 ? 0 : (c >= m[a] ? m[a] - 1 : c);
        }
        cell_of[i] = (c3[0] * m[1] + c3[1]) * m[2] + c3[2];
        start[cell_of[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; c++) start[c + 1] += start[c];
    {
        int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * ncell);
        memcpy(fill, start, sizeof(int64_t) * ncell);
        for (int32_t i = 0; i < n; i++) order[fill[cell_of[i]]++] = i;
        free(fill);
    }
    memset(counts, 0, sizeof(int32_t) * 2 * size);
    const double mn = minD, mx = maxD, bn = bin;
    int64_t overflow = 0;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(+ : overflow)
    {
        int32_t *pc = counts;
#ifdef _OPENMP
        if (omp_get_num_threads() > 1) pc = (int32_t *)calloc(2 * size, sizeof(int32_t));
#pragma omp for schedule(dynamic, 1)
#endif
        for (int64_t c = 0; c < ncell; c++) {
            const int32_t cz0 = (int32_t)(c % m[2]), cy0 = (int32_t)((c / m[2]) % m[1]), cx0 = (int32_t)(c / ((int64_t)m[1] * m[2]));
            for (int o = 0; o < 27; o++) {
                const int32_t nx = (cx0 + o / 9 - 1 + m[0]) % m[0];
                const int32_t ny = (cy0 + (o / 3) % 3 - 1 + m[1]) % m[1];
                const int32_t nz = (cz0 + o % 3 - 1 + m[2]) % m[2];
                const int64_t c2 = ((int64_t)nx * m[1] + ny) * m[2] + nz;
                for (int64_t p = start[c]; p < start[c + 1]; p++) {
                    const int32_t i = order[p];
                    const float *a = coords + 3 * (int64_t)i;
                    for (int64_t q = start[c2]; q < start[c2 + 1]; q++) {
                        const int32_t j = order[q];
                        if (j <= i) continue;
                        if (!f64) {
                            const float d = ref_pbc_dist(a[0], a[1], a[2], coords + 3 * (int64_t)j, B);
                            ref_bin_one(d, elem[i], mol[i], mol[j], elem[j], nel, H, minD, maxD, bin, pc, pc + size,
                                        &overflow);
                            continue;
                        }
                        const double dx = (double)a[0] - (double)coords[3 * (int64_t)j];
                        const double dy = (double)a[1] - (double)coords[3 * (int64_t)j + 1];
                        const double dz = (double)a[2] - (double)coords[3 * (int64_t)j + 2];
                        const double bx = dx - ref_round64(dx), by = dy - ref_round64(dy), bz = dz - ref_round64(dz);
                        const double rx = (bx * Bd[0] + by * Bd[3]) + bz * Bd[6];
                        const double ry = (bx * Bd[1] + by * Bd[4]) + bz * Bd[7];
                        const double rz = (bx * Bd[2] + by * Bd[5]) + bz * Bd[8];
                        const double d = sqrt((rx * rx + ry * ry) + rz * rz);
                        if (d < mn || d >= mx) continue;
                        const int64_t b = (int64_t)((d - mn) / bn);
                        const int64_t flat = ((int64_t)elem[i] * nel + elem[j]) * H + b;
                        if (flat >= size) {
                            overflow++;
                            continue;
                        }
                        pc[(mol[i] == mol[j] ? 0 : size) + flat]++;
                    }
                }
            }
        }
#ifdef _OPENMP
        if (pc != counts) {
#pragma omp critical
            for (int64_t s = 0; s < 2 * size; s++) counts[s] += pc[s];
            free(pc);
        }
#endif
    }
    free(start);
    free(cell_of);
    free(order);
    return overflow;
}

/* ---------------------------------------------------- boundary conditions */

/* get_reciprocal_basis (boundary_conditions_collection.pyx:45-80).  The
   cofactor for rbasis[0,1] reuses basis[0,1] where basis[0,2] belongs (:59);
   it is kept.  vol is a C double holding the float32 triple product (Cython
   infers double for the untyped float expression); the normalising divisions
   run in double and round to float32 on store. */
/* --------------------------------------------------------- atomic_distances */

#define AD_INTER 1
#define AD_INTRA 2
#define AD_WITHIN 4
#define AD_RED_UPPER 8
#define AD_RED_LOWER 16
#define AD_RED 32

/* _single_atomic_distances_dists (atomic_distances.pyx:42-118) with
   ncores = 1: i ascending, float32 sums in that order.  Arrays (nel,nel). */
void ref_single_atomic_distances_dists(int32_t atomIndex, const float *distances, int32_t n, const int32_t *mol,
                                       const int32_t *elem, int32_t nel, float *dintra, float *dinter,
                                       int32_t *nintra, int32_t *ninter, const float *lowerLimit,
                                       const float *upperLimit, int flags, int allAtoms) {
    int32_t atomMol = mol[atomIndex], atomEl = elem[atomIndex];
    int32_t start = allAtoms ? 0 : atomIndex;
    for (int32_t i = start; i < n; i++) {
        if (i == atomIndex) continue;
        int32_t molI = mol[i];
        if (!(flags & AD_INTRA) && molI == atomMol) continue;
        if (!(flags & AD_INTER) && !(molI == atomMol)) continue;
        float distance = distances[i];
        int32_t elI = elem[i];
        float lower = lowerLimit[elI * nel + atomEl];
        float upper = upperLimit[elI * nel + atomEl];
        if (flags & AD_WITHIN) {
            if (distance < lower) continue;
            if (distance >= upper) continue;
        } else if ((distance >= lower) && (distance < upper)) {
            continue;
        }
        if (flags & AD_RED_UPPER) {
            distance = (float)fabs((double)(upper - distance));
        } else if (flags & AD_RED_LOWER) {
            distance = (float)fabs((double)(lower - distance));
        } else if (flags & AD_RED) {
            if (distance > (lower + upper) / 2.0f)
                distance = (float)fabs((double)(upper - distance));
            else
                distance = (float)fabs((double)(lower - distance));
        }
        if (molI == atomMol) {
            dintra[atomEl * nel + elI] += distance;
            nintra[atomEl * nel + elI] += 1;
        } else {
            dinter[atomEl * nel + elI] += distance;
            ninter[atomEl * nel + elI] += 1;
        }
    }
}

/* multiple_atomic_distances_coords (atomic_distances.pyx:326-410): each atom
   of indexes in order, distances by pairs_distances_to_indexcoords.  Adds
   into the outputs. */
void ref_multiple_atomic_distances_coords(const int32_t *indexes, int32_t nidx, const float *coords, int32_t n,
                                          const float *B, int isPBC, const int32_t *mol, const int32_t *elem,
                                          int32_t nel, const float *lowerLimit, const float *upperLimit, int flags,
                                          int allAtoms, int32_t *nintra, float *dintra, int32_t *ninter,
                                          float *dinter) {
    float *d = (float *)malloc(sizeof(float) * (n > 0 ? n : 1));
    for (int32_t c = 0; c < nidx; c++) {
        ref_distances_to_indexcoords(indexes[c], coords, n, B, isPBC, allAtoms, d);
        ref_single_atomic_distances_dists(indexes[c], d, n, mol, elem, nel, dintra, dinter, nintra, ninter,
                                          lowerLimit, upperLimit, flags, allAtoms);
    }
    free(d);
}

/* multiple_atomic_distances_dists (atomic_distances.pyx:419-491): column c of
   the (n,k) distances for atom indexes[c].  Adds into the outputs. */
void ref_multiple_atomic_distances_dists(const int32_t *indexes, int32_t nidx, const float *distances, int32_t n,
                                         int32_t k, const int32_t *mol, const int32_t *elem, int32_t nel,
                                         const float *lowerLimit, const float *upperLimit, int flags, int allAtoms,
                                         int32_t *nintra, float *dintra, int32_t *ninter, float *dinter) {
    float *d = (float *)malloc(sizeof(float) * (n > 0 ? n : 1));
    for (int32_t c = 0; c < nidx; c++) {
        for (int32_t i = 0; i < n; i++) d[i] = distances[(int64_t)i * k + c];
        ref_single_atomic_distances_dists(indexes[c], d, n, mol, elem, nel, dintra, dinter, nintra, ninter,
                                          lowerLimit, upperLimit, flags, allAtoms);
    }
    free(d);
}

/* _pair_elements_stats (atomic_distances.pyx:632-656).  Adds. */
void ref_pair_elements_stats(const int32_t *elem, const int32_t *mol, int32_t n, int32_t nel, int32_t *nintra,
                             int32_t *ninter) {
    for (int32_t i = 0; i < n - 1; i++)
        for (int32_t j = i + 1; j < n; j++) {
            if (mol[j] == mol[i])
                nintra[elem[i] * nel + elem[j]] += 1;
            else
                ninter[elem[i] * nel + elem[j]] += 1;
        }
}

double ref_get_reciprocal_basis(const float *b, float *r) {
    r[0] = b[4] * b[8] - b[7] * b[5];
    r[1] = -(b[1] * b[8] - b[7] * b[1]);
    r[2] = b[1] * b[5] - b[4] * b[2];
    r[3] = -(b[3] * b[8] - b[6] * b[5]);
    r[4] = b[0] * b[8] - b[6] * b[2];
    r[5] = -(b[0] * b[5] - b[3] * b[2]);
    r[6] = b[3] * b[7] - b[6] * b[4];
    r[7] = -(b[0] * b[7] - b[6] * b[1]);
    r[8] = b[0] * b[4] - b[3] * b[1];
    double vol = (double)(b[0] * r[0] + b[3] * r[1] + b[6] * r[2]);
    for (int k = 0; k < 9; k++) r[k] = (float)((double)r[k] / vol);
    return vol;
}

/* transform_coordinates (boundary_conditions_collection.pyx:88-110) */
void ref_transform_coordinates(const float *M, const float *c, int32_t n, float *out) {
    for (int32_t i = 0; i < n; i++) {
        const float *p = c + 3 * i;
        out[3 * i] = p[0] * M[0] + p[1] * M[3] + p[2] * M[6];
        out[3 * i + 1] = p[0] * M[1] + p[1] * M[4] + p[2] * M[7];
        out[3 * i + 2] = p[0] * M[2] + p[1] * M[5] + p[2] * M[8];
    }
}

/* box_coordinates_real_distances (boundary_conditions_collection.pyx:119-165):
   diff = boxCoords[i] - boxCoords[atomIndex], then the same rounding. */
void ref_box_coordinates_real_distances(int32_t atomIndex, const int32_t *indexes, int32_t k,
                                        const float *coords, const float *B, float *out) {
    float ax = coords[3 * atomIndex], ay = coords[3 * atomIndex + 1], az = coords[3 * atomIndex + 2];
    for (int32_t ii = 0; ii < k; ii++) {
        const float *c = coords + 3 * (int64_t)indexes[ii];
        float box_dx = c[0] - ax;
        float box_dy = c[1] - ay;
        float box_dz = c[2] - az;
        box_dx -= ref_round(box_dx);
        box_dy -= ref_round(box_dy);
        box_dz -= ref_round(box_dz);
        float real_dx = box_dx * B[0] + box_dy * B[3] + box_dz * B[6];
        float real_dy = box_dx * B[1] + box_dy * B[4] + box_dz * B[7];
        float real_dz = box_dx * B[2] + box_dy * B[5] + box_dz * B[8];
        out[ii] = (float)sqrt((double)(real_dx * real_dx + real_dy * real_dy + real_dz * real_dz));
    }
}

/* ------------------------------------------------------------- MT19937 */
/* Textbook MT19937 (Matsumoto & Nishimura 1998) as used by CPython's random
   module and NumPy's legacy RandomState.  State import/export goes through
   the 624-word arrays those libraries expose, so no seeding code is needed. */
typedef struct { uint32_t mt[624]; int32_t pos; } ref_mt_t;

static void ref_mt_twist(ref_mt_t *s) {
    for (int i = 0; i < 624; i++) {
        uint32_t y = (s->mt[i] & 0x80000000u) | (s->mt[(i + 1) % 624] & 0x7fffffffu);
        s->mt[i] = s->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    s->pos = 0;
}

uint32_t ref_mt_next32(ref_mt_t *s) {
    if (s->pos >= 624) ref_mt_twist(s);
    uint32_t y = s->mt[s->pos++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

/* random.random() / RandomState.random_sample(): 53-bit double */
double ref_mt_double(ref_mt_t *s) {
    uint32_t a = ref_mt_next32(s) >> 5, b = ref_mt_next32(s) >> 6;
    return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
}

/* random.randrange(0, n) through _randbelow_with_getrandbits (n < 2^31) */
int32_t ref_mt_randbelow(ref_mt_t *s, int32_t n) {
    if (n <= 0) return 0;
    int k = 0;
    while ((1u << k) <= (uint32_t)n && k < 32) k++;   /* k = n.bit_length() */
    uint32_t r;
    do { r = ref_mt_next32(s) >> (32 - k); } while (r >= (uint32_t)n);
    return (int32_t)r;
}

int ref_mt_sizeof(void) { return (int)sizeof(ref_mt_t); }
