/*
 * meca_oracle.h -- CPU restatement of Cytosim's implicit mechanics step (Meca).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path and the single-threaded CPU baseline timed beside it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library never links or calls it.
 *
 * Semantics follow the 2017 reference (FP64, `real` = double, DIM = 2 or 3 at
 * run time).  Every function cites the reference file:line it restates.
 */
#ifndef MECA_ORACLE_H
#define MECA_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_meca orc_meca;

typedef struct {
    double time_step;     /* SimulProp::time_step */
    double kT;            /* SimulProp::kT */
    double viscosity;     /* SimulProp::viscosity */
    double hydro_radius0; /* FiberProp::hydrodynamic_radius[0] (default 0.0125) */
    double hydro_radius1; /* FiberProp::hydrodynamic_radius[1] (default 10) */
} orc_params;

typedef struct {
    double tolerance;     /* SimulProp::tolerance */
    int    tol_mode;      /* 0: |r|_inf < tolerance*noiseLevel (meca.cc:731); 1: |r|_inf < tolerance*|b|_inf */
    int    precondition;  /* SimulProp::precondition */
    int    max_iter;      /* <=0: DIM*nbPts (meca.cc:731) */
} orc_solve_params;

typedef struct {
    int    iterations;    /* Monitor::iterations of the final attempt */
    int    flag;          /* Monitor::flag */
    int    converged;
    int    fallback;      /* 0 none, 1 rescued by seed vRHS, 2 rescued by method switch, 3 dead */
    int    precond_used;  /* 1 if the final attempt used BCGSP */
    double residual;      /* Monitor::residual */
    double threshold;     /* residual threshold of the monitor */
    double noise_level;   /* noiseLevel (meca.cc:633-643) */
    int    total_iterations; /* sum over all attempts */
} orc_report;

orc_meca* orc_create(int dim);
void      orc_destroy(orc_meca*);
void      orc_set_params(orc_meca*, const orc_params*);

/* Meca::clear / Meca::add (meca.cc:55-79) for a RigidFiber Mecable */
void      orc_clear(orc_meca*);
int       orc_add_fiber(orc_meca*, uint32_t nb_points, double segmentation, double rigidity);
/* the same with the fiber's own FiberProp::hydrodynamic_radius (0: the
   parameters') and, height > 0, the surface drag model (fiber.cc:401-447) */
int       orc_add_fiber_ex(orc_meca*, uint32_t nb_points, double segmentation, double rigidity,
                           double hydro_radius0, double hydro_radius1, double height);
int       orc_add_solid(orc_meca*, uint32_t nb_points, const double* radius, const double* points,
                        int reshape_timer);   /* Solid after fixShape (solid.cc:352-400) */
uint32_t  orc_nb_points(const orc_meca*);
uint32_t  orc_mat_index(const orc_meca*, int mecable);

/* FiberNaked::bestNbPoints (fiber_naked.cc:22-30) */
uint32_t  orc_best_nb_points(double ratio);

/* Meca::prepare (meca.cc:488-537): points of all Mecables, DIM*nbPts, in matIndex order */
int       orc_prepare(orc_meca*, const double* points);

/* FiberNaked::interpolate (fiber_naked.cc:1050-1065) with fnAbscissa = 0 */
void      orc_interpolate(const orc_meca*, int fiber, double abscissa,
                          uint32_t* mat1, uint32_t* mat2, double* coef);

/* assembly (meca_inter.cc) -- indices are matrix indices (matIndex) */
void      orc_add_pure_force(orc_meca*, uint32_t a, const double* force);
void      orc_inter_link_ee(orc_meca*, uint32_t a, uint32_t b, double w);
void      orc_inter_link_ie(orc_meca*, uint32_t a1, uint32_t a2, double ca, uint32_t b, double w);
void      orc_inter_link_ii(orc_meca*, uint32_t a1, uint32_t a2, double ca,
                            uint32_t b1, uint32_t b2, double cb, double w);
/* Couple::setInteractions (couple.cc:87-92) for n couples (hands: fiber,
   abscissa from the minus end); state: bit 1|2 both attached, NULL: all */
void      orc_inter_couples(orc_meca*, uint32_t n, const uint32_t* fa, const double* aa, const uint32_t* fb,
                            const double* ab, const double* w, const uint8_t* state);
void      orc_inter_long_link_ee(orc_meca*, uint32_t a, uint32_t b, double len, double w);
void      orc_inter_clamp_e(orc_meca*, uint32_t a, const double* g, double w);
void      orc_inter_clamp_i(orc_meca*, uint32_t a1, uint32_t a2, double ca, const double* g, double w);
void      orc_inter_long_clamp(orc_meca*, uint32_t a, const double* center, double len, double w);
void      orc_inter_plane_e(orc_meca*, uint32_t a, const double* dir, const double* g, double w);
/* position-dependent emitters on interpolated points (meca_inter.cc) */
void      orc_add_pure_force_i(orc_meca*, uint32_t a1, uint32_t a2, double ca, const double* force);
void      orc_inter_long_link_ie(orc_meca*, uint32_t a1, uint32_t a2, double ca, uint32_t b, double len, double w);
void      orc_inter_long_link_ii(orc_meca*, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2,
                                 double cb, double len, double w);
void      orc_inter_side_link_ie(orc_meca*, uint32_t a1, uint32_t a2, double ca, uint32_t b, double len, double w);
void      orc_inter_side_link_ii(orc_meca*, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2,
                                 double cb, double len, double w);
int       orc_inter_side_side_link(orc_meca*, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2,
                                   double cb, double len, double w);
void      orc_inter_sliding_link_ie(orc_meca*, uint32_t a1, uint32_t a2, double ca, uint32_t b, double w);
void      orc_inter_sliding_link_ii(orc_meca*, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2,
                                    double cb, double w);
void      orc_inter_side_sliding_link_ie(orc_meca*, uint32_t a1, uint32_t a2, double ca, uint32_t b, double len,
                                         double w);
void      orc_inter_side_sliding_link_ii(orc_meca*, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2,
                                         double cb, double len, double w);
int       orc_inter_torque(orc_meca*, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2, double cb,
                           double cosinus, double sinus, double w);
void      orc_inter_side_clamp(orc_meca*, uint32_t a1, uint32_t a2, double ca, const double* g, double len,
                               double w);
void      orc_inter_sliding_clamp(orc_meca*, uint32_t a1, uint32_t a2, double ca, const double* g, double w);
void      orc_inter_plane_i(orc_meca*, uint32_t a1, uint32_t a2, double ca, const double* dir, const double* g,
                            double w);
/* steric interactions (simul_solve.cc:12-188, point_grid.cc); returns 0, -1 (no range),
   -2 (negative extension) or -3 (steric_max_range too short for a segment or a point);
   brute_force: every pair once instead of the grid's neighbour cells; the calls
   made are logged (kind 0 interLongLink PE-PE, 1 interSideSlidingLink PI-PE, 2 PI-PI) */
typedef struct {
    double extension[3];
    double max_range;
    double stiff_push, stiff_pull;
} orc_steric_params;
typedef struct {
    int      kind;
    uint32_t p[4];
    double   ca, cb, len, w;
} orc_steric_call;
int       orc_steric(orc_meca*, const orc_steric_params*, const int* steric, const double* radius,
                     const double* range, int brute_force, orc_steric_call* log, uint32_t caplog,
                     uint32_t* ncalls);
/* Fiber::setInteractions CONFINE_INSIDE with SpaceSphere (fiber.cc:471-481; space_sphere.cc:148-151) */
void      orc_confine_sphere(orc_meca*, double radius, double stiffness);

/* Meca::solve (meca.cc:598-845), split in phases:
   setup   = matrices ready, forces, Brownian, RHS, P' (meca.cc:605-693)
   iterate = preconditioner + BCGSP/BCGS + fallback chain (meca.cc:704-799)
   apply   = vPTS += vSOL, force recompute, export (meca.cc:802-835)
   gauss: N(0,1) draws, one per degree of freedom, in object order.        */
int       orc_setup(orc_meca*, const double* gauss);
int       orc_iterate(orc_meca*, const orc_solve_params*, const double* guess, orc_report*);
int       orc_apply(orc_meca*);
int       orc_solve(orc_meca*, const orc_solve_params*, const double* gauss, orc_report*);

/* operator-level entry points, valid after orc_setup */
void      orc_multiply(const orc_meca*, const double* X, double* Y);      /* meca.cc:157-181 */
int       orc_compute_preconditioner(orc_meca*);                          /* meca.cc:411-444 */
void      orc_precondition(const orc_meca*, const double* X, double* Y);  /* meca.cc:448-466 */
void      orc_get_block(const orc_meca*, int mecable, double* blk);       /* meca.cc:231-283 */
int       orc_get_inverse_block(const orc_meca*, int mecable, double* blk);
void      orc_project(const orc_meca*, int mecable, const double* X, double* Y, double sc); /* rigid_fiber_project.cc:137-149 */
void      orc_add_rigidity(const orc_meca*, int mecable, const double* X, double* Y);      /* rigid_fiber.cc:294-303 */
void      orc_add_projection_diff(const orc_meca*, int mecable, const double* X, double* Y);
void      orc_compute_forces(orc_meca*);                                  /* meca.cc:550-580 */
void      orc_forces_linear(const orc_meca*, const double* X, double* Y, int with_rigidity); /* meca.cc:143-147 */

/* state export */
enum { ORC_PTS = 0, ORC_SOL, ORC_BAS, ORC_RHS, ORC_FOR };
void      orc_get_vector(const orc_meca*, int which, double* out);
void      orc_get_fiber_points(const orc_meca*, double* out);   /* points held by the Mecables */
void      orc_get_tensions(const orc_meca*, double* out);       /* rfLag, one per segment */
double    orc_get_mobility(const orc_meca*, int mecable);
void      orc_get_solid(const orc_meca*, int mecable, double* out); /* drag, drag_rot, mom2d, center[3], mom[9] */
int       orc_solid_timer(const orc_meca*, int mecable);
void      orc_get_jjt(const orc_meca*, int mecable, double* diag, double* offd); /* factored pttrf */
/* mB (which=0) or mC (which=1) in the Numerical-Recipes layout of prepareForMultiply
   (matsparsesym1.cc:472-523); returns nbe; arrays may be NULL to query the size */
uint32_t  orc_get_matrix_nr(orc_meca*, int which, uint32_t* ija, double* sa);
/* dense (size x size) symmetric matrix of mB or mC, column-major */
void      orc_get_matrix_dense(const orc_meca*, int which, double* out);

/* FiberGrid::nearbySegments / tryToAttach's distance test by brute force
   (fiber_grid.cc:248-337, fiber_locus.cc:62-97); returns the number kept */
uint32_t  orc_nearby_segments(const orc_meca*, const double* place, double range, int inclusive, int exclude,
                              uint32_t cap, uint32_t* fiber, uint32_t* point, double* abscissa, double* dis);

/* the storage of mB/mC on its own (MatrixSparseSymmetric1, matsparsesym1.cc),
   for checks against that class: add = operator()(i,j) += v (137-224),
   nnz = nbNonZeroElements (355-362), prepare = prepareForMultiply (472-523),
   mul_add = vecMulAdd (dim 1) / vecMulAddIso2D / vecMulAddIso3D (526-631),
   block = addDiagonalBlock (full) / addTriangularBlock (262-310) */
typedef struct orc_sparse orc_sparse;
orc_sparse* orc_sparse_create(uint32_t n);
void      orc_sparse_destroy(orc_sparse*);
void      orc_sparse_zero(orc_sparse*);
void      orc_sparse_add(orc_sparse*, uint32_t i, uint32_t j, double v);
uint32_t  orc_sparse_nnz(const orc_sparse*);
int       orc_sparse_nonzero(const orc_sparse*);
void      orc_sparse_prepare(orc_sparse*);
void      orc_sparse_mul_add(const orc_sparse*, int dim, const double* X, double* Y);
void      orc_sparse_block(const orc_sparse*, int full, uint32_t x, uint32_t sx, double* M);
uint32_t  orc_sparse_nr(const orc_sparse*, uint32_t* ija, double* sa);

/* motor couples between solves: Couple::stepAA / stepAF / stepFA
   (couple.cc:125-150) with Motor::stepLoaded / stepUnloaded (motor.cc:19-57)
   and the end handling of Hand::handleOutOfRange (hand.cc:102-122) for
   fibers that neither grow nor shrink; the random unbinding tests are not
   modelled.  state[k]: 1 hand 1 attached, 2 hand 2 attached, 4 hand 1 is a
   motor, 8 hand 2 is a motor.  Uses the points held by the Mecables. */
typedef struct {
    double max_speed, stall_force;
    int    limit_speed, hold_growing_end;
} orc_motor;
void      orc_step_couples(const orc_meca*, uint32_t n, const uint32_t* fa, double* aa, const uint32_t* fb,
                           double* ab, const double* w, uint8_t* state, const orc_motor* mp);

/* trajectory frame: Simul::writeObjects (simul_file.cc:402-489) in binary mode, the
   fiber and solid sections from the points held by the Mecables.  real_bytes 4 writes
   reals as OutputWrapper::writeFloat (iowrapper.cc:535-547), 8 as writeDouble
   (iowrapper.cc:570-582; the lossless form: the time as %.17g and ", real 8" after
   the format).  Returns the frame's length; writes out[0..cap) only.  Per Mecable
   in registration order: */
typedef struct {
    uint32_t number;
    int32_t  mark;
    uint32_t prop;
    uint32_t signature;
    double   cut_wished, abscissa;
} orc_frame_record;
size_t    orc_write_frame(const orc_meca*, const orc_frame_record* records, double time, const char* date,
                          int real_bytes, const void* before, size_t nbefore, const void* after, size_t nafter,
                          unsigned char* out, size_t cap);

/* LAPACK restatements (documented algorithms), exported for pinning tests */
int       orc_dgetrf(int n, double* a, int lda, int* ipiv);
int       orc_dgetri(int n, double* a, int lda, const int* ipiv);
int       orc_dpttrf(int n, double* d, double* e);
void      orc_dptts2(int n, const double* d, const double* e, double* b);
int       orc_dgtsv(int n, double* dl, double* d, double* du, double* b);
int       orc_reshape(int dim, uint32_t ns, const double* src, double* dst, double cut);

#ifdef __cplusplus
}
#endif
#endif
