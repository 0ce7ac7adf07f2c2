/*
 * frmc.h -- C ABI of libfrmc_hip.so, the gfx950 implementation of fullrmc's
 * pair-distance / pair-histogram hot path.
 *
 * Conventions
 *  - Every array argument is a DEVICE pointer (HBM), C-ordered, float32 or
 *    int32 as in the reference (fullrmc Globals.py:29-30).  The caller owns
 *    all memory; nothing here keeps a pointer past the call except the
 *    frmc_engine_* handle, which owns its state.
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream).  All
 *    entry points are asynchronous on that stream unless stated otherwise.
 *  - Return value: 0 on success, FRMC_E* on failure; frmc_last_error()
 *    gives the message (thread-local).  Nothing throws across the ABI.
 *  - Histogram outputs are (nel, nel, histSize) float32 arrays holding
 *    integer counts, exactly as the reference returns them.
 *
 * Each entry point cites the reference interface it replaces
 * (paths relative to the fullrmc source tree).
 */
#ifndef FRMC_H
#define FRMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FRMC_OK 0
#define FRMC_EINVAL 1   /* bad argument (shape, size, index) */
#define FRMC_EHIP 2     /* HIP runtime error */
#define FRMC_ENOMEM 3   /* device allocation failed */
#define FRMC_ESTATE 4   /* engine handle misuse */

const char *frmc_last_error(void);
int frmc_version(void);
/* number of pair bins that fell outside the histogram array in the last
   histogram call on this thread.  The reference computes the bin as
   (int)((d-min)/bin) with no bounds check, so a distance just below
   maxDistance can give bin == histSize: inside the (nel,nel,H) array that
   aliases the next row's bin 0 (reproduced), past the array's end it would
   write out of bounds (counted here instead). */
int64_t frmc_last_overflow(void);
/* When enabled (per thread), the all-pairs kernel of the next
   full_pairs call is bracketed by HIP events on its stream and the call
   synchronises; frmc_last_kernel_ms() returns that kernel's duration. */
void frmc_set_timing(int enable);
float frmc_last_kernel_ms(void);

/* Full-histogram schedule autotuner.  No counterpart in the reference (its
   pairs_histograms.pyx:86-176 has one fixed OpenMP loop); every schedule
   gives the same counts.  A schedule is (threads per workgroup = column
   tile: 256, 512 or 1024; row atoms per thread: 1, 2 or 4, or -1 for the
   cell-list kernel with 256 or 512 threads, available when a periodic cell
   holds at least three cells of the range's width per axis and 125 in all;
   workgroups per CU: 1..16, more than fit at once allowed).  With tuning on (the
   default) the first all-pairs call of a shape class (device, floor(log2 n),
   bins, boundary kind, float64-recompute mode) times each candidate twice
   with HIP events on the call's stream, outside the frmc_set_timing window,
   and later calls of the class reuse the fastest.
   The tuning call synchronises its stream, so the first call of a class
   cannot be captured in a HIP graph (tune first, or force a schedule).
   frmc_set_autotune: 1 on, 0 off (fixed rule), -1 off and forget the
   tuned classes.  frmc_force_full_schedule pins one schedule for every call
   (rowAtoms == 0 releases it).  frmc_last_full_schedule reports the
   schedule of this thread's last all-pairs call (tuned = 1 when it came
   from the tuner). */
void frmc_set_autotune(int enable);
int frmc_force_full_schedule(int32_t threads, int32_t rowAtoms, int32_t workgroupsPerCU);
int frmc_last_full_schedule(int32_t *threads, int32_t *rowAtoms, int64_t *workgroups, int32_t *tuned);
/* Pairs the last all-pairs call of this thread evaluated: N(N-1)/2, or for
   the cell-list kernel the pairs in visited cell pairs (counted only while
   frmc_set_timing is on; -1 otherwise).  Measurement only. */
int frmc_last_full_evaluated(int64_t *pairs);

/* ------------------------------------------------------------------------
 * pairs_distances  (Extensions/pairs_distances.pyx)
 * ---------------------------------------------------------------------- */

/* pairs_distances_to_indexcoords (pairs_distances.pyx:874-916).
   out[i] = |coords[atomIndex]-coords[i]| (PBC: fractional minimum image,
   then basis).  With allAtoms==0 only out[atomIndex:n] is written. */
int frmc_pairs_distances_to_indexcoords(int32_t atomIndex, const float *coords, int32_t n,
                                        const float *basis, int isPBC, int allAtoms, float *out,
                                        void *stream);

/* pairs_distances_to_multi_indexcoords (pairs_distances.pyx:975-1021).
   out is (n, k): column c holds the distances of atom indexes[c]. */
int frmc_pairs_distances_to_multi_indexcoords(const int32_t *indexes, int32_t k, const float *coords,
                                              int32_t n, const float *basis, int isPBC, int allAtoms,
                                              float *out, void *stream);

/* pairs_distances_to_point (pairs_distances.pyx:827-871); point is (3,) */
int frmc_pairs_distances_to_point(const float *point, const float *coords, int32_t n, const float *basis,
                                  int isPBC, float *out, void *stream);

/* pairs_distances_to_multi_points (pairs_distances.pyx:924-972).  points is
   (3, k) -- the reference iterates points[:,i] -- and out is (n, k). */
int frmc_pairs_distances_to_multi_points(const float *points, int32_t k, const float *coords, int32_t n,
                                         const float *basis, int isPBC, float *out, void *stream);

/* point_to_point_distance (pairs_distances.pyx:786-824); out is (1,) */
int frmc_point_to_point_distance(const float *point1, const float *point2, const float *basis, int isPBC,
                                 float *out, void *stream);

/* pairs_differences_to_indexcoords (pairs_distances.pyx:626-672); out (n,3) */
int frmc_pairs_differences_to_indexcoords(int32_t atomIndex, const float *coords, int32_t n,
                                          const float *basis, int isPBC, int allAtoms, float *out,
                                          void *stream);

/* pairs_differences_to_multi_indexcoords (pairs_distances.pyx:727-783); out (n,3,k) */
int frmc_pairs_differences_to_multi_indexcoords(const int32_t *indexes, int32_t k, const float *coords,
                                                int32_t n, const float *basis, int isPBC, int allAtoms,
                                                float *out, void *stream);

/* pairs_differences_to_point (pairs_distances.pyx:580-623); out (n,3) */
int frmc_pairs_differences_to_point(const float *point, const float *coords, int32_t n, const float *basis,
                                    int isPBC, float *out, void *stream);

/* pairs_differences_to_multi_points (pairs_distances.pyx:675-724); points (3,k), out (n,3,k) */
int frmc_pairs_differences_to_multi_points(const float *points, int32_t k, const float *coords, int32_t n,
                                           const float *basis, int isPBC, float *out, void *stream);

/* from_to_points_differences (pairs_distances.pyx:483-527); out (n,3) = BC(to-from) */
int frmc_from_to_points_differences(const float *pointsFrom, const float *pointsTo, int32_t n,
                                    const float *basis, int isPBC, float *out, void *stream);

/* pair_difference_to_point (pairs_distances.pyx:534-577); out (3,) = BC(point2-point1) */
int frmc_pair_difference_to_point(const float *point1, const float *point2, const float *basis, int isPBC,
                                  float *out, void *stream);

/* ------------------------------------------------------------------------
 * pairs_histograms  (Extensions/pairs_histograms.pyx)
 * ---------------------------------------------------------------------- */

/* single_pairs_histograms (pairs_histograms.pyx:77-141).  Adds the counts of
   atom atomIndex vs distances[start:n] into hintra/hinter IN PLACE
   (start = 0 if allAtoms else atomIndex). */
int frmc_single_pairs_histograms(int32_t atomIndex, const float *distances, int32_t n,
                                 const int32_t *moleculeIndex, const int32_t *elementIndex,
                                 float *hintra, float *hinter, int32_t numberOfElements, int32_t histSize,
                                 float minDistance, float maxDistance, float bin, int allAtoms, void *stream);

/* multiple_pairs_histograms_coords (pairs_histograms.pyx:150-217).  Writes
   (overwrites) hintra/hinter with the counts of every atom in indexes vs all
   atoms (allAtoms) or vs the atoms after it (allAtoms==0). */
int frmc_multiple_pairs_histograms_coords(const int32_t *indexes, int32_t nidx, const float *boxCoords,
                                          int32_t n, const float *basis, int isPBC,
                                          const int32_t *moleculeIndex, const int32_t *elementIndex,
                                          int32_t numberOfElements, float minDistance, float maxDistance,
                                          float bin, int32_t histSize, int allAtoms, float *hintra,
                                          float *hinter, void *stream);

/* multiple_pairs_histograms_dists (pairs_histograms.pyx:225-281).
   distances is (n, k); column c belongs to atom indexes[c] (c < nidx <= k). */
int frmc_multiple_pairs_histograms_dists(const int32_t *indexes, int32_t nidx, const float *distances,
                                         int32_t n, int32_t k, const int32_t *moleculeIndex,
                                         const int32_t *elementIndex, int32_t numberOfElements,
                                         float minDistance, float maxDistance, float bin, int32_t histSize,
                                         int allAtoms, float *hintra, float *hinter, void *stream);

/* full_pairs_histograms_coords (pairs_histograms.pyx:289-335): every
   unordered pair once, in slot [elem(earlier), elem(later)].  Tiled
   all-pairs kernel with LDS-private bins. */
int frmc_full_pairs_histograms_coords(const float *boxCoords, int32_t n, const float *basis, int isPBC,
                                      const int32_t *moleculeIndex, const int32_t *elementIndex,
                                      int32_t numberOfElements, float minDistance, float maxDistance,
                                      float bin, int32_t histSize, float *hintra, float *hinter,
                                      void *stream);

/* full_pairs_histograms_dists (pairs_histograms.pyx:343-383): distances (n,k),
   indexes = arange(k), allAtoms false. */
int frmc_full_pairs_histograms_dists(const float *distances, int32_t n, int32_t k,
                                     const int32_t *moleculeIndex, const int32_t *elementIndex,
                                     int32_t numberOfElements, float minDistance, float maxDistance, float bin,
                                     int32_t histSize, float *hintra, float *hinter, void *stream);

/* Same as frmc_full_pairs_histograms_coords but leaves int32 counts in
   counts[2][nel][nel][H] (intra then inter) without the float conversion;
   used by the device engine and by the benchmark. */
int frmc_full_pairs_counts(const float *boxCoords, int32_t n, const float *basis, int isPBC,
                           const int32_t *moleculeIndex, const int32_t *elementIndex, int32_t numberOfElements,
                           float minDistance, float maxDistance, float bin, int32_t histSize, int32_t *counts,
                           void *stream);

/* Full histogram of the binary64 recompute (no reference counterpart: the
   second oracle of the full-histogram path, "bit-exact against a float64
   recompute").  Every unordered pair i<j once, slot [e(i)][e(j)], with the
   distance evaluated in double from the float32 box coordinates and basis
   (the reference's minimum-image rule and operand order, no FMA, correctly
   rounded sqrt), range minDistance <= d < maxDistance and bin
   (int)((d-minDistance)/bin) in double.  Distances run in float32; a pair
   whose float32 distance is within a rigorous error bound of a bin edge or
   range end (or whose fractional difference is within rounding of +-1/2) is
   recomputed in double.  counts: int32 [2][nel][nel][H], then the overflow
   counter, one pad word and a 64-bit count of re-checked pairs (2*size+4
   ints).  *rechecked (host, may be NULL; forces a synchronisation) receives
   that count. */
int frmc_full_pairs_counts_f64(const float *boxCoords, int32_t n, const float *basis, int isPBC,
                               const int32_t *moleculeIndex, const int32_t *elementIndex, int32_t numberOfElements,
                               float minDistance, float maxDistance, float bin, int32_t histSize, int32_t *counts,
                               int64_t *rechecked, void *stream);

/* ------------------------------------------------------------------------
 * boundary_conditions_collection (Extensions/boundary_conditions_collection.pyx)
 * ---------------------------------------------------------------------- */

/* get_reciprocal_basis (boundary_conditions_collection.pyx:45-80), including
   the basis[0,1] cofactor typo of :59.  rbasis (3,3) float32 device; vol is
   a host double (the reference returns a Python float).  Synchronous. */
int frmc_get_reciprocal_basis(const float *basis, float *rbasis, double *vol, void *stream);

/* transform_coordinates (boundary_conditions_collection.pyx:88-110) */
int frmc_transform_coordinates(const float *transMatrix, const float *coords, int32_t n, float *out,
                               void *stream);

/* box_coordinates_real_distances (boundary_conditions_collection.pyx:119-165) */
int frmc_box_coordinates_real_distances(int32_t atomIndex, const int32_t *indexes, int32_t k,
                                        const float *boxCoords, int32_t n, const float *basis, float *out,
                                        void *stream);

/* ------------------------------------------------------------------------
 * atomic_distances  (Extensions/atomic_distances.pyx)
 * ---------------------------------------------------------------------- */

/* flags of the atomic_distances entries (the reference's bint arguments) */
#define FRMC_AD_INTER 1       /* interMolecular */
#define FRMC_AD_INTRA 2       /* intraMolecular */
#define FRMC_AD_WITHIN 4      /* countWithinLimits */
#define FRMC_AD_RED_UPPER 8   /* reduceDistanceToUpper */
#define FRMC_AD_RED_LOWER 16  /* reduceDistanceToLower */
#define FRMC_AD_RED 32        /* reduceDistance */

/* single_atomic_distances_dists (atomic_distances.pyx:241-318, the loop of
   :42-118).  For i in [start,n) (start = 0 if allAtoms else atomIndex),
   i != atomIndex, pairs passing the molecule filter and the limits
   lowerLimit/upperLimit[elem[i]][elem[atomIndex]] add their (reduced)
   distance to d{intra,inter}[elem[atomIndex]][elem[i]] and 1 to
   n{intra,inter}[...] IN PLACE.  Float sums are added in the reference's
   order (i ascending), so they are bit-identical.  lower/upper/d/n arrays
   are (nel,nel) (the reference's (nel,nel,1)).  Synchronises once. */
int frmc_single_atomic_distances_dists(int32_t atomIndex, const float *distances, int32_t n,
                                       const int32_t *moleculeIndex, const int32_t *elementIndex, float *dintra,
                                       float *dinter, int32_t *nintra, int32_t *ninter, int32_t numberOfElements,
                                       const float *lowerLimit, const float *upperLimit, int flags, int allAtoms,
                                       void *stream);

/* multiple_atomic_distances_coords (atomic_distances.pyx:326-410): the same
   for every atom of indexes in order, distances from pairs_distances on
   the coordinates; full_atomic_distances_coords (:500-560) is this with
   indexes = arange(n) and allAtoms = 0.  ADDS into the four outputs (the
   reference returns them from zeros). */
int frmc_multiple_atomic_distances_coords(const int32_t *indexes, int32_t nidx, const float *boxCoords, int32_t n,
                                          const float *basis, int isPBC, const int32_t *moleculeIndex,
                                          const int32_t *elementIndex, int32_t numberOfElements,
                                          const float *lowerLimit, const float *upperLimit, int flags,
                                          int allAtoms, int32_t *nintra, float *dintra, int32_t *ninter,
                                          float *dinter, void *stream);

/* multiple_atomic_distances_dists (atomic_distances.pyx:419-491): distances
   (n,k), column c for atom indexes[c], c < nidx <= k;
   full_atomic_distances_dists (:569-623) is this with indexes = arange(k),
   allAtoms = 0.  ADDS into the four outputs. */
int frmc_multiple_atomic_distances_dists(const int32_t *indexes, int32_t nidx, const float *distances, int32_t n,
                                         int32_t k, const int32_t *moleculeIndex, const int32_t *elementIndex,
                                         int32_t numberOfElements, const float *lowerLimit,
                                         const float *upperLimit, int flags, int allAtoms, int32_t *nintra,
                                         float *dintra, int32_t *ninter, float *dinter, void *stream);

/* pair_elements_stats (atomic_distances.pyx:632-675): for every pair i<j,
   nintra/ninter[elem[i]][elem[j]] += 1 by molecule.  Overwrites. */
int frmc_pair_elements_stats(const int32_t *elementIndex, const int32_t *moleculeIndex, int32_t n,
                             int32_t numberOfElements, int32_t *nintra, int32_t *ninter, void *stream);

/* ------------------------------------------------------------------------
 * G(r) -> S(Q) sine transform and reciprocal_space
 * (Extensions/reciprocal_space.pyx)
 * ---------------------------------------------------------------------- */

#define FRMC_SQ_EXACT 0  /* float32 products summed sequentially over r: bit-identical */
#define FRMC_SQ_MFMA 1   /* matrix-core contraction, float32 accumulation (not bit-identical) */

/* The transform of the S(Q) constraints for k curves at once
   (_get_Sq_from_Gr, StructureFactorConstraints.py:772-773 with addOne=1,
   :1255-1256 with addOne=0): Sq[p][q] = sum_r Gr[p][r]*sqMatrix[r][q]
   (+1).  Gr (k,H), sqMatrix (H,nQ) as the constraint builds it
   (:302-312), Sq (k,nQ).  FRMC_SQ_EXACT reproduces np.sum(..., axis=0)
   bit for bit; FRMC_SQ_MFMA runs on the matrix cores. */
int frmc_sq_transform(const float *Gr, int32_t k, int32_t H, const float *sqMatrix, int32_t nQ, int addOne,
                      int mode, float *Sq, void *stream);

/* frmc_sq_transform without any allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, allocation-free serving loops).  Same arguments
   and results (bit-identical in both modes); the matrix-core mode keeps its
   split partials in a library-owned device array of at most 2^22 floats
   (16 MiB) per GPU, held until the process ends, and returns FRMC_EINVAL
   when a shape needs more.  All calls on one GPU share that array: calls
   that overlap on one GPU race. */
int frmc_sq_transform_noalloc(const float *Gr, int32_t k, int32_t H, const float *sqMatrix, int32_t nQ,
                              int addOne, int mode, float *Sq, void *stream);

/* Gr_to_sq (reciprocal_space.pyx:89-122): sq (m,) = 1 + sum_r
   dr*sin(q*r)/q*Gr[r], dr = distances[1]-distances[0], each term and each
   accumulation in double, rounded to float32 after each accumulation as
   the Cython loop does.  n >= 2. */
int frmc_Gr_to_sq(const float *distances, const float *Gr, int32_t n, const float *qrange, int32_t m, float *sq,
                  void *stream);

/* gr_to_sq (reciprocal_space.pyx:42-80): sq = 1 + sum_r
   4*PI*rho*dr*r*sin(q*r)/q*(gr[r]-1), same arithmetic rules.  n >= 2. */
int frmc_gr_to_sq(const float *distances, const float *gr, int32_t n, const float *qrange, int32_t m, float rho,
                  float *sq, void *stream);

/* sq_to_Gr (reciprocal_space.pyx:131-162): Gr[r] = (2/PI) * sum_q
   dq*Q*(S(Q)-1)*sin(Q*r), dq = qValues[1]-qValues[0], in double, rounded
   once.  The reference's loop passes an array to the C sin and raises
   TypeError for every input; this is the transform it documents.  m >= 2. */
int frmc_sq_to_Gr(const float *qValues, const float *rValues, int32_t nr, const float *sq, int32_t m, float *Gr,
                  void *stream);

/* ------------------------------------------------------------------------
 * Device-resident engine: the Engine.run step (Engine.py:3166-3344) with the
 * constraint chain of PairDistributionConstraint / StructureFactorConstraint
 * (PairDistributionConstraints.py:847-895, 1044-1166;
 * StructureFactorConstraints.py:772-834, 1011-1096), PairCorrelationConstraint
 * (PairCorrelationConstraints.py:126-169, 305-392) and
 * ReducedStructureFactorConstraint (StructureFactorConstraints.py:1237-1261)
 * on the GPU, evaluating
 * k proposals per window.  Descriptor arrays are HOST pointers, copied at
 * creation.  No reference counterpart as a single call: the reference
 * reaches the same results through per-move Python calls.
 * ---------------------------------------------------------------------- */

#define FRMC_PDC 0   /* model = G(r) = prefactor*(g(r)-1) on the histogram grid */
#define FRMC_SFC 1   /* model = S(Q) = sum_r G(r) M[r,q] + 1 */
#define FRMC_PCF 2   /* model = g(r) on the histogram grid */
#define FRMC_RSFC 3  /* model = reduced S(Q) = sum_r G(r) M[r,q] */

typedef struct {
    int32_t kind;               /* FRMC_PDC, FRMC_SFC, FRMC_PCF or FRMC_RSFC */
    int32_t histSize;           /* H */
    float minDistance, maxDistance, bin;
    int32_t nQ;                 /* SFC/RSFC: number of Q points */
    const float *pairWeights;   /* [T] w_ij, T = nel(nel+1)/2 pairs in sorted-element order */
    const float *pairDensity;   /* [T] D_ij = N_ij / V as float32 */
    const float *shellVolumes;  /* [H] */
    const float *prefactor;     /* [H] (4*PI*shellCenters)*rho0 as float32 */
    const float *experimental;  /* [H] (PDC, PCF) or [nQ] (SFC, RSFC) */
    const float *dataWeights;   /* same length or NULL */
    const float *sqMatrix;      /* SFC/RSFC: [H][nQ] dr*sin(r*Q)/Q as float32 */
    float scaleFactor;          /* constant scale factor (1 = none) */
} frmc_constraint_desc;

typedef struct {
    int32_t n, nel, isPBC;
    const float *box;            /* (n,3) box (fractional) coordinates */
    const float *real;           /* (n,3) real coordinates, or NULL (needed by the native generator) */
    const int32_t *mol, *elem;   /* (n,) molecule and element index */
    float basis[9], recip[9];    /* basis vectors (rows) and reciprocal basis */
    int32_t ngroups;
    const int32_t *group_offsets; /* [ngroups+1] CSR offsets into group_atoms */
    const int32_t *group_atoms;   /* atom indexes of every group, in group order */
    int32_t nconstraints;
    frmc_constraint_desc constraints[4];
    int32_t kmax;                /* largest window (proposals per evaluation) */
    void *stream;                /* NULL: the engine creates its own stream */
} frmc_engine_desc;

typedef struct frmc_engine frmc_engine;

/* MT19937 state in the layout CPython's random.getstate() and NumPy's
   RandomState.get_state() expose: 624 words and the read position. */
typedef struct {
    uint32_t mt[624];
    int32_t pos;
} frmc_mt19937;

typedef struct {
    int64_t generated, tried, accepted, tolerated;
    int64_t windows, evaluated;  /* windows launched, proposals evaluated */
} frmc_run_stats;

int frmc_engine_create(frmc_engine **out, const frmc_engine_desc *desc);
int frmc_engine_destroy(frmc_engine *e);
/* full O(N^2) histograms of every constraint, then chi-squares and total
   (Engine.initialize_used_constraints -> compute_data, Engine.py:3084-3119) */
int frmc_engine_compute_data(frmc_engine *e, float *chi2_out, float *tse_out);
/* chi-squares and total of the current histograms */
int frmc_engine_refresh(frmc_engine *e, float *chi2_out, float *tse_out);
/* evaluate k proposals: group_ids[k], new_box = moved box coordinates of
   each proposal's group atoms, concatenated in proposal order.  Writes the
   total chi-square after each move (tse_out[k]) and the per-constraint
   chi-squares (chi2_out[k][C], may be NULL).  Synchronous. */
int frmc_engine_eval(frmc_engine *e, int32_t k, const int32_t *group_ids, const float *new_box, float *tse_out,
                     float *chi2_out);
/* A window whose proposals name their moved atoms (no reference
   counterpart as one call: the moves of Engine.run whose atoms are not one
   group, the swap's group plus its drawn partners, Engine.py:3192-3201 and
   Core/MoveGenerator.py:502-529, evaluated as frmc_engine_eval evaluates
   group moves).  Proposal p moves counts[p] distinct atoms (at most twice
   the largest group, 64 at most), listed in order in atoms, to the box
   coordinates at the same place in new_box; frmc_engine_commit commits one
   of them. */
int frmc_engine_eval_atoms(frmc_engine *e, int32_t k, const int32_t *counts, const int32_t *atoms,
                           const float *new_box, float *tse_out, float *chi2_out);
/* accept proposal p of the last evaluation (accept_move, PDC:1131-1152) */
int frmc_engine_commit(frmc_engine *e, int32_t p);
int frmc_engine_sync(frmc_engine *e);
/* read back coordinates, slot-exact counts [2][nel][nel][H]+1 and pair-type
   sums [T][H] of one constraint, current chi-squares and total (host ptrs) */
int frmc_engine_get_state(frmc_engine *e, float *box_out, int32_t constraint, int32_t *slot_out, int32_t *base_out,
                          float *chi2_out, float *tse_out);
/* Native run loop for single-group translation moves
   (RandomSelector + TranslationGenerator, Selectors/RandomSelectors.py:87-94,
   Generators/Translations.py:171-187, Core/Collection.py:674-701): proposals
   come from the two MT19937 streams (py = random, np = numpy.random), k per
   window, evaluated speculatively; the first accepted one is committed and
   both streams are rewound to just after it, so the trajectory equals the
   sequential one.  Optional logs (length nsteps): the group index of each
   generated move, whether it was accepted, and the total after the step. */
int frmc_engine_run_translation(frmc_engine *e, int64_t nsteps, int32_t k, float minAmp, float maxAmp,
                                float tolerance, frmc_mt19937 *py_state, frmc_mt19937 *np_state,
                                frmc_run_stats *stats, int32_t *group_log, uint8_t *accept_log, float *tse_log);
/* Independent replicas (no reference counterpart: many separate runs of
   Engine.run, Engine.py:3377-3516, each with its own engine and its own pair
   of random streams): every engine advances by nsteps generated moves
   exactly as frmc_engine_run_translation would advance it alone.
   py_states, np_states and stats are arrays of nengines entries (stats
   accumulate).  When every engine has single-atom groups (and batching is
   on), all active replicas' windows go out in one launch per kernel per
   step, on the first engine's stream; otherwise each engine's windows run
   on its own stream, several in flight at once.  Engines of one call share
   one device.  After an error return the engines' states are unspecified
   (a staged move may be half applied): destroy and recreate them. */
int frmc_engines_run_translation(frmc_engine **engines, int32_t nengines, int64_t nsteps, int32_t k, float minAmp,
                                 float maxAmp, float tolerance, frmc_mt19937 *py_states, frmc_mt19937 *np_states,
                                 frmc_run_stats *stats);
/* 1 (default): frmc_engines_run_translation batches single-atom replicas'
   windows into one launch per kernel; 0: one stream per engine. */
void frmc_set_replica_batching(int enable);
/* Host threads that draw and walk the replicas' windows (0, the default:
   the CPUs available to the process less two, at most 16). */
void frmc_set_replica_threads(int nthreads);
/* Host time of the last batched replica run by phase, in ms (ms_out holds
   5): [0] waiting for each window's totals, walking them and drawing and
   staging the next window (one parallel pass per lane step), [1] the
   uploads and launches, [2] unused (0), [3] set-up, [4] the final commits;
   steps_out the lane steps. */
int frmc_replica_batch_times(double *ms_out, int64_t *steps_out);
/* Per-kernel timing with HIP events recorded on the engine stream around
   the launches of every `every`-th window (two event records per timed
   launch).  frmc_engine_profile(e, every >= 1) resets and enables, (e, 0)
   disables; frmc_engine_kernel_times fills FRMC_NKERNELS totals (ms) and
   the number of timed launches. */
#define FRMC_K_DELTA 0   /* k_delta: moved group vs all atoms, k proposals */
#define FRMC_K_EVAL 1    /* k_eval_a: G(r), PDC chi-square, S(Q) per proposal */
#define FRMC_K_COMMIT 2  /* k_commit: histograms and coordinates of an accepted move
                            (launched only when not folded into the next k_delta) */
#define FRMC_K_EVAL_B 3  /* k_eval_b: SFC chi-squares, totals, completion flags */
#define FRMC_NKERNELS 4
int frmc_engine_profile(frmc_engine *e, int every);
int frmc_engine_kernel_times(frmc_engine *e, double *ms_out, int64_t *launches_out);
/* k_eval_a of the engine's last window state launched reps times back to
   back on the engine stream between two HIP events; *ms_out = the average
   per launch (launch gaps amortised, unlike the per-launch event pairs of
   frmc_engine_profile).  It writes only the window's scratch (S(Q) and
   chi-square partials), so the engine state is unchanged.  No reference
   counterpart (measurement only). */
int frmc_engine_time_eval(frmc_engine *e, int32_t k, int32_t reps, float *ms_out);
/* Window-evaluation shape tuning.  No reference counterpart (the reference
   evaluates one move at a time).  A shape s is a multiple of 4 in 0..60,
   or 80, 84, 88, 92 (s & 64: the split chain, leaf layout only):
   bins per thread per G(r) pass 8 (s & 4, above 1,024 bins) or 4; work
   units aligned to the XCDs (s & 8: unit u on XCD u mod 8); S(Q) units of
   one pairwise-sum leaf whose sums the host combines (s & 16, no k_eval_b)
   or of 64 Q points; the S(Q) matrix read from its copy interleaved by
   four rows (s & 32) or its transposed copy; with s & 64, 1,024-thread
   workgroups in which two waves add the S(Q) products that fourteen others
   compute into LDS.  Every shape gives the same results.  An engine times
   every shape for each window-size bucket ceil(log2 k) on its first window
   and keeps the fastest (cached per device, constraint layout, pair types
   and largest window).  frmc_set_eval_shape forces one shape for every
   engine (-1: tuned); frmc_engine_eval_shapes reports the engine's shape
   for buckets 0..n-1 (tuning first if needed). */
int frmc_set_eval_shape(int32_t shape);
int frmc_engine_eval_shapes(frmc_engine *e, int32_t *shapes_out, int32_t n);
/* No counterpart in the reference (an engine-internal schedule).  In a
   periodic box whose groups are single atoms, the engine keeps a copy of the
   atoms sorted into compact chunks of 256 with each chunk's bounding box in
   wrapped fractional coordinates, and k_delta skips a (chunk, proposal)
   whose old and new positions are both farther than the largest binned
   distance (plus a rounding margin) from the chunk's box along some lattice
   axis: those pairs are out of range in the minimum image, so the deltas are
   the same.  frmc_set_delta_cells: -1 (default) for engines of at least
   4,096 atoms, 1 for every such engine, 0 never (engines created after the
   call).  frmc_engine_delta_cells reports the chunks (0: no chunk walk) and
   the cut-off distance used. */
int frmc_set_delta_cells(int32_t mode);
int frmc_engine_delta_cells(frmc_engine *e, int32_t *nchunk_out, float *wcut_out);
/* Device window loop of frmc_engine_run_translation (no reference
   counterpart: the per-move loop of Engine.run, Engine.py:3231-3344, with
   its draws, Core/Collection.py:674-701 and Selectors/RandomSelectors.py:
   87-94, and its accept rule, Engine.py:3308-3317, as one device step per
   window).  For single-atom groups and k <= 128 the walk of each window's
   totals, the rewind of both random streams and the next window's draws run
   on the GPU, queued behind the window's evaluation in batches of up to 256
   windows of one size; the host synchronises once per batch and runs the
   last windows (fewer steps left than a window holds) itself.  Trajectory,
   logs, statistics and stream states are those of the host loop.
   frmc_engines_run_translation's batched replicas (single-atom groups)
   take the device loop too: one step kernel per step walks and draws for
   every replica.  frmc_set_window_loop: 1 on, 0 off, -1 (default) on for
   batched replicas and off for a single engine (there the device step
   costs as much as the host's round trip between windows, about 7 us, and
   measured runs were 2-6 % slower with it).  frmc_engine_loop_windows:
   windows the device loop has run on the engine. */
int frmc_set_window_loop(int32_t mode);
int frmc_engine_loop_windows(frmc_engine *e, int64_t *windows_out);
/* window size of frmc_engine_run_translation: adaptive (default; k is the
   maximum, the size follows the acceptance rate) or always k */
int frmc_engine_set_adaptive(frmc_engine *e, int enable);
/* replace the histograms of one constraint with slot counts
   [2][nel][nel][H]+1 (host), e.g. a constraint's current data */
int frmc_engine_set_counts(frmc_engine *e, int32_t constraint, const int32_t *slot_in);
/* replace / read the host mirror of real coordinates (n,3) the native
   generator moves from */
int frmc_engine_set_real(frmc_engine *e, const float *real_in);
int frmc_engine_get_real(frmc_engine *e, float *real_out);

#ifdef __cplusplus
}
#endif

#endif /* FRMC_H */
