/*
 * bem3_hip.h -- C-ABI of the MI355X (gfx950) implementation of the
 * bem3_emf_b2 hot path: coefficient-matrix assembly, the dense complex
 * LU / solve, the multi-object coupling loop, field evaluation and the
 * Maxwell-stress force.  Plain C types only; complex numbers cross as
 * interleaved (re, im) doubles.  Tables indexed by element use the
 * reference's 1-based layout with an unused row 0.
 *
 * Every int-returning entry point returns 0 on success or a negative BH_E*
 * code; bh_last_error() gives the message of the last failure on the
 * calling thread.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference tree).
 */
#ifndef BEM3_HIP_H
#define BEM3_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define BH_EHIP   -1  /* HIP runtime error                    */
#define BH_ENODEV -2  /* no GPU                               */
#define BH_EARG   -3  /* invalid argument                     */
#define BH_EUNSUP -4  /* unsupported element type / rule      */
#define BH_ESTATE -5  /* call order (e.g. solve before build) */
#define BH_ESING  -6  /* singular matrix (zero pivot)         */
#define BH_EDOM   -7  /* point inside another object (overlap)*/
#define BH_ECONV  -8  /* iterative refinement did not converge */

typedef struct bh_object bh_object;

const char *bh_last_error(void);

/* select the device and build the quadrature tables (called implicitly) */
int bh_init(int device);
/* Releases the library's device state (rule tables, work areas, timing
 * events) after the caller destroyed every object; its device streams stay
 * until the process ends and are used again.  The next entry point sets
 * the library up again.  For programs that must not leave device frees to
 * process exit (no counterpart in the reference).                         */
int bh_finalize(void);
int bh_sync(void);
void *bh_stream(void);

/* One object: the paired boundary mesh as the reference holds it after
 * read_mesh_data (setup2.c:205-503): node table rn[Nn+1][3], element table
 * ed[Ne+1][4] (ed[i][3] == 0 for triangles), main / sub domain ids md, sd,
 * MN interior media with refractive indices n_med[MN+1] (n_med[0] = the
 * surrounding medium) and the vacuum wavelength.  Computes the element
 * constants and node data (setup2.c:666-749, 838-869) and uploads them.
 * Replaces initialize_domd2 (setup2.c:99-122) for the device path.       */
bh_object *bh_object_create(int Nn, const double *rn, int Ne, const int *ed, const int *md, const int *sd, int MN,
                            const double *n_med, double lambda0);
void bh_object_destroy(bh_object *o);
/* out[0..3] = Nn, Ne, NN, MN; out[4+d] = elements seen by domain d;
 * out[12+d] = node rows of domain d (d < 8)                              */
int bh_object_info(const bh_object *o, int *out);
int bh_object_geometry(const bh_object *o, double *cr, double *cw, double *ren, double *wen, int *eni);

/* Coefficient matrices of every domain (create_cmatrix_domain,
 * solve_bieq2.c:98-184, with dcoef_bd_node_t2, d3b1_elem.c:482-539).
 * type must be 2 (GLN rule with the GHN near-pair switch, PREC_DEF).     */
int bh_object_assemble(bh_object *o, int type);
/* Arithmetic of the object's later calls (no counterpart in the reference,
 * which is FP64 throughout): BH_PREC_FP64 (default) or BH_PREC_MIXED --
 * far-element quadrature of the assembly and the boundary field sums in
 * FP32 (near-singular and self-element integrals, the force accumulation
 * and every residual stay FP64), factor mode 0 picks the FP32 block
 * elimination refined in FP64 (mode 3).                                   */
#define BH_PREC_FP64  0
#define BH_PREC_MIXED 1
/* BH_PREC_MIXED_LM: the mixed form for objects whose FP64 matrices do not
 * fit the device (one interior medium only): G, H are assembled straight
 * into complex-FP32 storage that the factorisation overwrites in place, and
 * every FP64 product with them (right-hand side, refinement residual,
 * boundary E/H rows) assembles their rows again, a block of rows at a time.
 * Results are those of BH_PREC_MIXED; bh_object_matrix and the many-beam
 * solve are not available.                                                 */
#define BH_PREC_MIXED_LM 2
int bh_object_set_precision(bh_object *o, int prec);
/* rows per block of BH_PREC_MIXED_LM's on-demand assembly (0, the default:
 * as many as about 16 GB of FP64 rows hold); no counterpart in the
 * reference                                                               */
int bh_object_set_block_rows(bh_object *o, int rows);
int bh_object_gh_count(const bh_object *o, int did);
/* download: which 0 G, 1 H (nd x nd), 2 dG, 3 dH (nd x 2 x nd complex),
 * 4 dF (nd x 3 doubles: diagonal 1+F or F, dF/dt_zeta, dF/dt_eta).
 * Rows/columns follow the domain's element order, 3 node slots each.    */
int bh_object_matrix(const bh_object *o, int did, int which, double *out);
/* nrows selected rows (0-based) of the same matrices, one after another   */
int bh_object_matrix_rows(const bh_object *o, int did, int which, int nrows, const int *rows, double *out);

/* Incident plane waves: 12 doubles per beam (lambda0, n_surrounding,
 * power, re/im e0x, re/im e0y, fx, fy, fz, theta, phi) as read by
 * read_data_mfb_ipw (multi_fbeam.c:145-238).  Also fills the incident
 * boundary data of the open-domain node slots (init_boundary_data_u2,
 * mo_setup.c:704-757).                                                    */
int bh_object_set_beams(bh_object *o, int nb, const double *par);
/* A beam set of any multi_fbeam types, 16 doubles per beam: type (0 plane
 * wave, 1 focused plane wave, 2 focused with spiral phase modulation, 3
 * Bessel, 4 Bessel with spiral phase, 5 focused radially / azimuthally
 * polarised), lambda0, ni, NA (1, 2, 5) or deflection angle (3, 4), power,
 * re/im e0x (e0r for 5), re/im e0y (e0a for 5), fx, fy, fz, theta, phi,
 * spiral mode (2, 4), Gauss-Legendre points (1, 2, 5).  Replaces
 * read_data_mfb / setup_mfb (multi_fbeam.c:20-36, 100-120) and the beam
 * set-up functions of fpw.c:115, lgb.c:116, bsb.c:112, bslgb.c:115,
 * rab.c:111 for the incident fields of calc_mfb_EH / calc_mfb_EH_dv
 * (multi_fbeam.c:764-876); computes the incident boundary data like
 * bh_object_set_beams.  BH_ECONV when a phi integral does not converge
 * ("phi integration limit over" in the reference). */
int bh_object_set_mfb(bh_object *o, int nb, const double *par);
/* Factorise the boundary-integral system (create_Amatrix_csr + lu_dec_A,
 * solve_bieq2.c:186-408).  mode 1: the reference's dense 8NN matrix;
 * mode 2: exact block elimination for one interior medium (LU of G0 and
 * of M; the Schur complement S of the potential is factorised when a
 * coupling loop of several objects first needs it, a single solve takes
 * GMRES on it to a 1e-13 relative residual); mode 0 picks 2 when
 * possible (3 in the mixed precision of bh_object_set_precision); mode 3: the block
 * elimination in complex FP32 (mixed-precision form), every later solve
 * refined in FP64 against the FP64 matrices to a relative residual of
 * 1e-10 (BH_ECONV if 30 steps do not reach it).                          */
int bh_object_factor(bh_object *o, int mode);
/* refinement steps and final relative residual ||b - A x|| / ||b|| of the
 * last bh_object_solve (0, 0 outside the mixed-precision form), and the
 * GMRES steps it spent on the Schur complement (0 once S is formed: the
 * coupling loop of several objects forms it, a single solve does not)     */
int bh_object_solve_stats(const bh_object *o, int *iters, double *resid, int *gmres_steps);
/* boundary potentials from the current incident data (create_Bmatrix +
 * solve_ABmatrix + tmatrix_bd_store, mo_solve.c:137-274, 457-545)         */
int bh_object_solve(bh_object *o);
/* boundary E, H and their normal derivatives (solve_eh_bv + solve_deh_bv,
 * mo_solve.c:547-789)                                                     */
int bh_object_solve_EH(bh_object *o);
/* boundary values per domain: which 0 U, 1 dU ([dNe+1][4][4] complex),
 * 2 E, 3 dE, 4 H, 5 dH ([dNe+1][4][3]); set != 0 uploads                 */
int bh_object_state(bh_object *o, int did, int which, double *buf, int set);
/* incident data: which 0 Ui, 1 dUi, 2 Ui0, 3 dUi0 ([Ne+1][4][4])         */
int bh_object_incident(bh_object *o, int which, double *buf, int set);

/* Assembly (type as bh_object_assemble) and factorisation (mode as
 * bh_object_factor) of N objects -- create_matrix of each object
 * (solve_bieq2.c:10-96).  The assembly of every object and domain runs as
 * one batch (one launch per assembly pass over all of them; tuning entry
 * BH_TUNE_ASM_BATCH = 0: each object on its stream before its
 * factorisation), the factorisations -- with N > 1 also the Schur
 * complements the coupling loop solves with -- spread over several device
 * streams (tuning entry BH_TUNE_LANES), so that the latency-bound pieces
 * of different objects overlap.                                           */
int bh_mo_prepare(bh_object **objs, int N, int type, int mode);
/* Coupling loop over N objects (mo_solve_bv, mo_solve.c:10-88): every
 * object factorised and given the same beams.  Convergence when the
 * change measure (ccd_f) < ccd after more than iter_min iterations.
 * trace[iter] (optional) receives the change per iteration.  With the
 * environment variable BEM3_HIP_PROGRESS=1 each iteration also prints
 * its change on stderr.                                                  */
int bh_mo_solve(bh_object **objs, int N, double ccd, int iter_min, int iter_max, double *trace, int *iters);
/* device streams (lanes) over which the per-object work of bh_mo_prepare,
 * bh_mo_solve and bh_mo_solve_beams is spread, one host worker each: 1..8,
 * 0 for the tuning table's entry BH_TUNE_LANES.  With 1 every kernel runs in
 * order on the library's stream.  No counterpart in the reference.      */
int bh_mo_set_lanes(int n);
/* How the coupling loop forms the field of one object at another's node
 * slots (no counterpart in the reference, which always sums every element
 * pair): 0 (default) a multipole series of the source object where both
 * objects are in the mixed-precision form, the source has triangles only,
 * the surrounding medium is lossless and the slots keep a distance (source
 * sphere radius / slot distance <= 0.6, truncation about 1e-10 relative),
 * decided for all slots of a destination object or, in the one-beam loop,
 * per tile of 64 of its slots; the reference's element sums elsewhere; 1
 * the element sums everywhere.                                            */
int bh_mo_set_coupling(int mode);
/* ordered object pairs of the last coupling loop of several objects summed
 * by the series and by element sums, and the largest series order used   */
int bh_mo_plan_stats(int *series, int *direct, int *lmax);
/* ordered pairs of the last one-beam coupling loop that were split: the
 * series takes the 64-slot tiles of the destination whose slots all meet
 * its conditions, the element sums the other tiles (neither counted by
 * bh_mo_plan_stats); no counterpart in the reference                     */
int bh_mo_plan_split(int *split);
/* The coupling sums of the loop (add_scattered_mEMP / mEMP2_s,
 * mo_solve.c:296-396, summed as bh_mo_solve would sum them) from the
 * objects' current boundary values, at object dst's node slots: out
 * receives 2 x [Ne+1][4][4] complex (the scattered U, then dU/dn, in the
 * incident-data layout); the objects' incident data are left unchanged.  */
int bh_mo_coupling_sums(bh_object **objs, int N, int dst, double *out);
/* Many plane waves as separate right-hand sides (no counterpart in the
 * reference, which sums the beams of a set into one incident field,
 * multi_fbeam.c:764-815; per-beam results there take one run per beam):
 * nb beams of 12 doubles each (as bh_object_set_beams) are given to every
 * object, the coupling loop runs for all of them at once until every
 * beam's change is below ccd after more than iter_min iterations, and the
 * boundary E, H of every beam follow.  Every object must be factorised in
 * the mixed-precision form.  *iters receives the iteration count.       */
int bh_mo_solve_beams(bh_object **objs, int N, int nb, const double *par, double ccd, int iter_min, int iter_max,
                      int *iters);
/* one beam's values after bh_mo_solve_beams: which 0..5 as
 * bh_object_state (domain did), 6..9 the incident data Ui, dUi, Ui0, dUi0 */
int bh_object_beam_state(bh_object *o, int b, int did, int which, double *buf);
/* beam b of the many-beam solve made the object's one-beam state: its
 * boundary values and incident data copied into the object's arrays and its
 * plane wave made the object's beam set, so that bh_mo_field, bh_mo_force,
 * bh_mo_absorb and bh_mo_EH_bd report beam b (what one reference run with
 * that beam alone holds after mo_solve_iter, mo_solve.c:10-88)          */
int bh_object_select_beam(bh_object *o, int b);
/* force and torque of every object under every beam of the many-beam solve
 * (mo_force_FN, mo_force.c:10-54, of each object in one reference run per
 * beam): rc[3 m] the torque centre of object m (null: origin); F, Nq
 * [nb][N][3]; flag[b N + m] as bh_mo_force's return (-1: no open-domain
 * faces), may be null.  The objects keep the last beam's state.         */
int bh_mo_force_beams(bh_object **objs, int N, int type, const double *rc, double *F, double *Nq, int *flag);
/* scattered E, H of every beam of the many-beam solve at np points of the
 * open region far from the objects (mo_EH_mEMP_s, mo_field.c:154-181, as
 * example4.c:51-177 calls it per point), by per-object series of rule type
 * (0..3): every point at least 1/0.6 bounding radii from each object's
 * centre.  what 0: E, H ([nb][np][6] complex), 1: |E|^2, |H|^2
 * ([nb][np][2] doubles).                                                */
int bh_mo_far_field_beams(bh_object **objs, int N, int type, int np, const double *pts, int what, double *out);
/* fields at np points (mo_EH_s/t/i, mo_mEMP_s/t/i, mo_EH_mEMP_s/t/i;
 * mo_field.c:11-282): what 0: E,H (6 complex per point), 1: U (4 complex),
 * 2: E,H from the potentials' gradients (EH_mEMP_s, d3b1_field.c:68-138,
 * the far-field form); mode 0 scattered + internal, 1 total, 2 incident;
 * type 0 4-point, 1 7-point, 2 GLN, 3 GHN.  oid[p] receives the object
 * containing the point or -1.                                             */
int bh_mo_field(bh_object **objs, int N, int what, int mode, int type, int np, const double *pts, double *out, int *oid);
/* domain ids of np points w.r.t. one object (domain_id, setup2.c:944-978) */
int bh_object_domain(bh_object *o, int np, const double *pts, int *did);
/* fields at np points on object oid's surface: element t[p] (1-based, in
 * domain did's element list), local coordinates ze[2p], ze[2p+1];
 * mode 0 s, 1 t, 2 i, 3 a (mo_EH_*_bd, mo_field.c:632-718); type 2 only. */
int bh_mo_EH_bd(bh_object **objs, int N, int oid, int did, int np, const int *t, const double *ze, int type, int mode,
                double *out);
/* radiation force and torque about rc (mo_force_FN, mo_force.c:10-240)   */
int bh_mo_force(bh_object **objs, int N, int oid, int type, const double *rc, double *F, double *Nq);

/* element coefficients of the reference's element layer for n requests on
 * object o's elements (coef_rt, coef_bd, coef_bd_node, dcoef_rt,
 * dcoef_bd_tz, dcoef_bd_te, dcoef_bd_t2, dcoef_bd_node_t2;
 * d3b1_elem.h:112-124, d3b1_elem.c:148-539): nd 0 CC only, 1 and dC along
 * v0, 2 along v0 and v1; rule type 0..3 (4-point, 7/9-point,
 * Gauss-Legendre, the 64-point rule; with derivatives and type < 3 the
 * 64-point rule again where |dF| along a direction exceeds 0.5); k = kr +
 * i ki.  Request i: signed source element s[i]; t (may be null) the signed
 * target element: t[i] == s[i] is the self element at local coordinates
 * ze[2i..2i+1] (polar rules; derivatives are the principal values along
 * the element's own tangents), otherwise the point rt[3i..3i+2].  out: 27
 * complex values per request: CC, dC/dv0, dC/dv1, each G 0-3, H 4-7, F 8. */
int bh_elem_coef(bh_object *o, int n, int nd, int type, double kr, double ki, const int *s, const int *t,
                 const double *ze, const double *rt, const double *v0, const double *v1, double *out);
/* absorbed power (mo_absorb_P, mo_absorb.c:9-44); returns -2 when the
 * flux came out negative (its magnitude is returned)                      */
int bh_mo_absorb(bh_object **objs, int N, int oid, int type, double *P);

/* dense complex LU solve A X = B (row-major, n x n and n x nrhs), the
 * device replacement of LAPACKE_zgetrf/zgetri/zgesv (solve_bieq2.c:389-399,
 * mo_solve.c:751); A is overwritten by its LU factors, B by X            */
int bh_zgesv(int n, int nrhs, double *A, double *B);
/* The same solve with no device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops); no
 * counterpart in the reference.  A (n x n, row-major, overwritten by its LU
 * factors), B (n x nrhs, overwritten by X), piv (n ints: 0-based pivot rows)
 * and info (one int: 0, or the 1-based index of the first zero pivot) are
 * device buffers of the caller; stream is the hipStream_t to enqueue on
 * (NULL: the library's stream).  Enqueues kernels only: no allocation, no
 * memset, no synchronisation.  The factorisation's work area is a fixed
 * device array of the library (8.5 MB per GPU, held until the process
 * ends), set up by bh_init (call it first: BH_ESTATE otherwise); n is at
 * most 8192 (BH_EARG beyond).  Every call on a GPU shares that array:
 * calls that overlap on one GPU race.                                     */
int bh_zgesv_noalloc(int n, int nrhs, void *A, void *B, int *piv, int *info, void *stream);
/* C = alpha A B + beta C (row-major, complex): the MFMA ZGEMM kernel     */
int bh_zgemm_host(int m, int n, int k, const double *alpha, const double *A, const double *B, const double *beta,
                  double *C);
/* the same in complex FP32 (interleaved float pairs): the factorisation
 * and product kernels of the mixed-precision form (no counterpart in the
 * reference, whose arithmetic is FP64 throughout)                        */
int bh_cgesv(int n, int nrhs, float *A, float *B);
int bh_cgemm_host(int m, int n, int k, const float *alpha, const float *A, const float *B, const float *beta, float *C);

/* Kernel timing (no counterpart in the reference): while enabled, HIP events
 * on the library's stream bracket every launch of the tracked kernels.
 * bh_prof_read returns the summed kernel time (ms), the summed algorithmic
 * work and the launch count of one kernel since the last enable.  Work units:
 * BH_K_ASM_GL, BH_K_ASM_GH, BH_K_ASM_SELF: (row node, source element) pairs;
 * BH_K_ZGEMM (the 128x128-tile kernel) and BH_K_ZGEMM_SMALL (64x64-tile
 * kernel): real flops (8 m n k); BH_K_LU_PANEL: real
 * flops of the unblocked panel updates; BH_K_TRSM: real flops; BH_K_FIELD,
 * BH_K_SCATTER, BH_K_FIELD_BD: (point, source element) pairs.
 * BH_K_CGEMM, BH_K_CGEMM_SMALL, BH_K_LU_PANEL_C, BH_K_TRSM_C,
 * BH_K_FIELD_BD32, BH_K_SCATTER32: the FP32 kernels of the mixed-precision
 * form (same work units); BH_K_ASM_GLR: its assembly pass (FP64, reduced
 * far-pair rules).  BH_K_SCATTER_FAR, BH_K_SCATTER_FAR32: the far-triangle
 * pass of the direct coupling sums ((slot, source element) pairs; the
 * second pass counts under BH_K_SCATTER, BH_K_SCATTER32 with zero work).
 * BH_K_MP_BASIS: the coupling series' basis at the source points
 * ((point, coefficient) pairs); BH_K_MP_MOM: its moments (real flops);
 * BH_K_MP_EVAL: its evaluation at the node slots ((slot, coefficient)
 * pairs).  BH_K_FIELD_ROWS: the E, H integral rows of the many-beam force
 * ((point, source element) pairs).  BH_K_ZGEMV, BH_K_CGEMV: the few-right-hand-side
 * (n <= 8, k >= 256) product kernel, real flops (8 m n k).               */
#define BH_K_ASM_GL   0
#define BH_K_ASM_GH   1
#define BH_K_ASM_SELF 2
#define BH_K_ZGEMM    3
#define BH_K_LU_PANEL 4
#define BH_K_TRSM     5
#define BH_K_FIELD    6
#define BH_K_SCATTER  7
#define BH_K_ZGEMM_SMALL 8
#define BH_K_FIELD_BD 9
#define BH_K_CGEMM    10
#define BH_K_CGEMM_SMALL 11
#define BH_K_LU_PANEL_C 12
#define BH_K_TRSM_C   13
#define BH_K_ASM_GLR  14
#define BH_K_FIELD_BD32 15
#define BH_K_ZGEMV    16
#define BH_K_CGEMV    17
#define BH_K_SCATTER32 18
#define BH_K_SCATTER_FAR 19
#define BH_K_SCATTER_FAR32 20
#define BH_K_MP_BASIS 21
#define BH_K_MP_MOM   22
#define BH_K_MP_EVAL  23
#define BH_K_FIELD_ROWS 24
#define BH_K_COUNT    25
int bh_prof_enable(int on);
int bh_prof_read(int kid, double *ms, double *work, int *launches);
/* as bh_prof_read, with the summed algorithmic bytes of the launches (the
 * operands read once and the results written once; stated for the GEMM
 * kernels, 0 for the others)                                            */
int bh_prof_read_bytes(int kid, double *ms, double *work, double *bytes, int *launches);

/* Tuning table (no counterpart in the reference: the reference's loops
 * have no schedule to choose).  Entries are schedule and tile choices that
 * leave results unchanged up to the FP32 rounding of the mixed form's
 * refined products.  bh_init loads the table named by the environment
 * variable BEM3_HIP_TUNE, else the file tune_gfx950.txt beside the library
 * when it exists (lines "name value", '#' comments; written by the
 * autotuner bem3_emf_b2_amd/autotune.py from measurements on the device),
 * unless bh_tune_set or bh_tune_load ran first; BEM3_HIP_LANES,
 * BEM3_HIP_ASM_BATCH and BEM3_HIP_GEMM_BF6 override single entries.     */
#define BH_TUNE_LANES        0  /* device streams of the per-object work, 1..8                     */
#define BH_TUNE_ASM_BATCH    1  /* bh_mo_prepare: 1 every object's assembly as one batch, 0 each
                                   object assembled on its stream before its factorisation        */
#define BH_TUNE_GEMM_T_D     2  /* FP64 products: tile count (128 x 128) from which the 128x128-tile
                                   kernel takes a product with k >= 32 (the 64x64 kernel below)   */
#define BH_TUNE_GEMM_T_F     3  /* the same for complex-FP32 products                             */
#define BH_TUNE_GEMM_LONGK_T 4  /* ... and from this tile count when k >= BH_TUNE_GEMM_LONGK_K    */
#define BH_TUNE_GEMM_LONGK_K 5
#define BH_TUNE_GEMM_BF6     6  /* complex-FP32 128x128 tiles: 1 the bf16-split MFMA kernel, 0 the
                                   FP32-MFMA kernel                                               */
#define BH_TUNE_COUNT        7
int bh_tune_set(int key, int value);
int bh_tune_get(int key, int *value);
/* replace the table by a file's entries (entries it does not name take
 * their defaults); BH_EARG on an unknown name or a value out of range   */
int bh_tune_load(const char *path);
const char *bh_tune_name(int key);
/* the file the table came from ("defaults" when none)                   */
const char *bh_tune_source(void);
/* Time one product kernel on device buffers of its own (the autotuner's
 * probe): C = A B with m x k and k x n operands, prec 0 complex FP64, 1
 * complex FP32; kernel 0 the 64x64-tile kernel, 1 the 128x128-tile MFMA
 * kernel, 2 the bf16-split 128x128 kernel (prec 1); *ms is the mean time
 * of reps launches after one warm-up launch.                            */
int bh_gemm_time(int prec, int kernel, int m, int n, int k, int reps, double *ms);

#ifdef __cplusplus
}
#endif
#endif
