/*
 * cymeca.h -- C ABI of the MI355X implicit mechanics step (a drop-in for the
 * reference's `Meca`).
 *
 * One context per host thread; all functions return 0 (CYMG_OK) or an error
 * code and never throw; cymg_last_error() gives the message.  Host pointers
 * are borrowed for the duration of a call; device buffers belong to the
 * context.  Indices are matrix indices (`Mecable::matIndex() + point`, the
 * reference's `PointExact::matIndex`).  Vectors are interleaved x,y(,z) per
 * point, DIM*nbPts doubles, as in the reference (mecable.h:86-90).
 *
 * Each declaration names the reference interface it replaces
 * (paths relative to the reference tree).
 */
#ifndef CYMECA_H
#define CYMECA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CYMG_OK              0
#define CYMG_ERR_INVALID     1   /* bad argument or call out of sequence */
#define CYMG_ERR_HIP         2   /* HIP runtime error */
#define CYMG_ERR_PROJECTION  3   /* "could not build Fiber projection" (rigid_fiber_project.cc:84-85) */
#define CYMG_ERR_CONVERGENCE 4   /* "convergence failure in solver" (meca.cc:795) */
#define CYMG_ERR_UNSUPPORTED 5   /* size beyond the kernels' stated limits */

/* largest fiber handled by the per-fiber kernels (points).  Fibers of up to
   170 points run the two-launch iteration; a system with a longer fiber runs
   the four-launch one.  Preconditioner blocks beyond the workgroup-local
   kernels (more than 111 unknowns in 3D, 128 in 2D) are assembled column by
   column and inverted in device memory. */
#define CYMG_MAX_FIBER_POINTS 512
#define CYMG_MAX_SOLID_POINTS 65536

typedef struct cymg_context cymg_context;

/* SimulProp::time_step, kT, viscosity (simul_prop.h) and
   FiberProp::hydrodynamic_radius[2] (fiber_prop.cc:198-199) */
typedef struct {
    double time_step;
    double kT;
    double viscosity;
    double hydro_radius0;
    double hydro_radius1;
} cymg_sim_params;

enum {
    CYMG_TOL_REFERENCE = 0,   /* |r|_inf < tolerance*noiseLevel (meca.cc:731, bicgstab.cc:16,48) */
    CYMG_TOL_RELATIVE  = 1    /* |r|_inf < tolerance*|b|_inf */
};

enum {
    CYMG_PRECOND_NONE = 0,    /* Solver::BCGS (bicgstab.h:163-243) */
    CYMG_PRECOND_FP64 = 1,    /* Solver::BCGSP with FP64 getrf+getri blocks (meca.cc:323-466) */
    CYMG_PRECOND_FP32 = 2     /* Solver::BCGSP with blocks factorised and applied in FP32 */
};

enum {
    CYMG_GUESS_ZERO     = 0,  /* vSOL = 0 (meca.cc:704) */
    CYMG_GUESS_PREVIOUS = 1,  /* previous step's solution, kept per fiber */
    CYMG_GUESS_LEARNED  = 2   /* initial guess from the trained per-fiber network */
};

typedef struct {
    double tolerance;     /* SimulProp::tolerance */
    int    tol_mode;      /* CYMG_TOL_* */
    int    precondition;  /* CYMG_PRECOND_* */
    int    max_iter;      /* <= 0: DIM*nbPts (meca.cc:731) */
    int    guess;         /* CYMG_GUESS_* */
    int    fallback;      /* nonzero: the reference's rescue chain on failure (meca.cc:749-799) */
    int    gauss_on_device; /* nonzero: the gauss pointer given to setup/solve is device memory */
} cymg_solve_params;

typedef struct {
    int    iterations;        /* Monitor::iterations() of the final attempt */
    int    flag;              /* Monitor::flag(): 0, 2 (rho=0), 3 (omega=0), 4 (delta=0) */
    int    converged;
    int    fallback;          /* 0 none, 1 rescued by seed vRHS, 2 by switching method, 3 dead */
    int    precond_used;      /* CYMG_PRECOND_* of the final attempt */
    int    total_iterations;  /* over all attempts */
    double residual;          /* Monitor::residual() */
    double threshold;
    double noise_level;       /* noiseLevel (meca.cc:633-643) */
    int    guess_used;        /* CYMG_GUESS_* seeding the first attempt (LEARNED falls back to ZERO
                                 when no network is loaded) */
} cymg_solve_report;

/* context ---------------------------------------------------------------- */
int         cymg_create(int dim, int device, cymg_context** out);
int         cymg_destroy(cymg_context* ctx);
/* NULL: the message of this thread's last call without a context (null context, cymg_read_frame) */
const char* cymg_last_error(const cymg_context* ctx);
int         cymg_set_params(cymg_context* ctx, const cymg_sim_params* p);

/* registration: Meca::clear (meca.cc:55-60), Meca::add (meca.cc:67-79) of a
   RigidFiber (rigid_fiber.h) with its segmentation fnCut and FiberProp::rigidity */
int         cymg_clear(cymg_context* ctx);
int         cymg_add_fiber(cymg_context* ctx, uint32_t nb_points, double segmentation,
                           double rigidity, uint32_t* mat_index);
int         cymg_nb_points(const cymg_context* ctx, uint32_t* nb_points);
/* Meca::add of a Fiber with its own FiberProp drag settings (fiber.cc:315-447):
   hydrodynamic_radius[2] (0: the context's cymg_sim_params values) and
   surface_effect with cylinder_height (> 0: Fiber::dragCoefficientSurface,
   drag = 2 pi viscosity length / acosh(1 + height / radius0); 0: the drag in
   a volume, Fiber::dragCoefficientVolume) */
typedef struct {
    uint32_t nb_points;
    double   segmentation;     /* fnCut */
    double   rigidity;         /* FiberProp::rigidity */
    double   hydro_radius0;    /* FiberProp::hydrodynamic_radius[0], 0: context value */
    double   hydro_radius1;    /* FiberProp::hydrodynamic_radius[1], 0: context value */
    int      surface_effect;   /* FiberProp::surface_effect */
    double   cylinder_height;  /* FiberProp::cylinder_height (> 0 when surface_effect) */
} cymg_fiber_desc;
int         cymg_add_fiber_ex(cymg_context* ctx, const cymg_fiber_desc* desc, uint32_t* mat_index);
/* Inventoried::number() (inventoried.h:43) of every Mecable registered since cymg_clear (fibers in
   registration order, then Solids), unique.  When the previous registration
   gave identities too, the device keeps per Mecable what it knew of the same
   object (same identity and number of points) across the new order: its points
   (so cymg_prepare(NULL) can follow a reshuffle) and its share of the previous
   solution (CYMG_GUESS_PREVIOUS and the learned guess's inputs).  This is the
   per-object keying that a warm start needs when Simul::step reshuffles the
   Mecables (node_list.cc:244-280; meca.cc:695-703 explains why the reference
   seeds with zero instead).  Without identities these survive only an
   unchanged registration (same point counts in the same order). */
int         cymg_set_identities(cymg_context* ctx, uint32_t n, const uint32_t* ids);
/* Meca::add of a Solid (solid.h), after every fiber (simul_solve.cc:204-207):
   per-point radii (Solid::addSphere, solid.cc:352-362; 0 = no drag), the
   reference shape fixed from `points` (DIM per point, Solid::fixShape,
   solid.cc:370-400) and the initial soReshapeTimer (solid.cc:28, in [0, 7)).
   Its preconditioner block is kept as identity + rank 6 (3 in 2D), the exact
   inverse of the reference's dense block; a system with Solids runs the
   four-launch iteration and may not be split into independent systems. */
int         cymg_add_solid(cymg_context* ctx, uint32_t nb_points, const double* radius, const double* points,
                           int reshape_timer, uint32_t* mat_index);
/* soReshapeTimer of every Solid after the last getPoints (the caller's Solid
   objects keep it across Meca::clear/add, as the reference's Solids do) */
int         cymg_get_solid_timers(cymg_context* ctx, uint32_t nsol, int* timers);
/* independent systems solved together (no counterpart in the reference, which
   solves one system per Meca): fibers [first_fiber[s], first_fiber[s+1]) form
   system s, with its own Krylov scalars, threshold (noise level or |b|_inf) and
   stop test; interactions may not couple two systems.  Call after the add()s. */
int         cymg_set_systems(cymg_context* ctx, uint32_t nsys, const uint32_t* first_fiber);
/* per-system reports of the last solve (the solve's own report is the system
   with the most iterations, converged only if every system converged) */
int         cymg_get_system_reports(cymg_context* ctx, uint32_t nsys, cymg_solve_report* out);

/* Meca::prepare (meca.cc:488-537) with Fiber::prepareMecable (fiber.cc:450-460).
   points: host array of the Mecables' points (PointSet::putPoints), or NULL to
   reuse the points the device holds from the previous step's export. */
int         cymg_prepare(cymg_context* ctx, const double* points);

/* assembly, recorded in call order (meca_inter.cc); PointInterpolated is given
   as (mat1, mat2, coef) with pos = (1-coef)*P1 + coef*P2 (point_interpolated.h:26-106) */
int         cymg_add_pure_force(cymg_context* ctx, uint32_t a, const double* force);           /* meca_inter.cc:123-134 */
int         cymg_inter_link_ee(cymg_context* ctx, uint32_t a, uint32_t b, double weight);      /* meca_inter.cc:196-230 */
int         cymg_inter_link_ie(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca,
                               uint32_t b, double weight);                                     /* meca_inter.cc:240-275 */
int         cymg_inter_link_ii(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca,
                               uint32_t b1, uint32_t b2, double cb, double weight);           /* meca_inter.cc:285-321 */
/* evaluated on the device from the points, as the interactions below */
int         cymg_inter_long_link_ee(cymg_context* ctx, uint32_t a, uint32_t b, double len,
                                    double weight);                                            /* meca_inter.cc:334-415 */
int         cymg_inter_clamp_e(cymg_context* ctx, uint32_t a, const double* g, double weight); /* meca_inter.cc:1883-1907 */
int         cymg_inter_clamp_i(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca,
                               const double* g, double weight);                                /* meca_inter.cc:1918-1963 */
/* evaluated on the device from the points, as the interactions below */
int         cymg_inter_long_clamp(cymg_context* ctx, uint32_t a, const double* center,
                                  double len, double weight);                                 /* meca_inter.cc:1977-2022 */
int         cymg_inter_plane_e(cymg_context* ctx, uint32_t a, const double* dir,
                               const double* g, double weight);                               /* meca_inter.cc:2301-2326 */
int         cymg_add_pure_force_i(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca,
                                  const double* force);                                       /* meca_inter.cc:140-156 */
int         cymg_inter_plane_i(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, const double* dir,
                               const double* g, double weight);                               /* meca_inter.cc:2340-2378 */

/* Interactions that read the positions (meca_inter.cc).  Each call is recorded
   in call order and evaluated on the device from the points of cymg_prepare
   when the matrices are built (no host copy of the points); the entries it
   makes are those of the reference, summed in call order.  An interpolated
   point needs two distinct points of one Mecable.  Links between points that
   share a model point of one Mecable record nothing (point_interpolated.cc:21-33).
   Where the reference draws a random side for a target exactly on the fiber
   axis (Random::sign_exc, random.cc:73-82) the side is +1.  No periodic space. */
int         cymg_inter_long_link_ie(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b,
                                    double len, double weight);                               /* meca_inter.cc:425-505 */
int         cymg_inter_long_link_ii(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b1,
                                    uint32_t b2, double cb, double len, double weight);       /* meca_inter.cc:516-596 */
int         cymg_inter_side_link_ie(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b,
                                    double len, double weight);                               /* meca_inter.cc:617-759 */
int         cymg_inter_side_link_ii(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b1,
                                    uint32_t b2, double cb, double len, double weight);       /* meca_inter.cc:764-935 */
/* 2D only: CYMG_ERR_UNSUPPORTED in 3D, where the reference throws */
int         cymg_inter_side_side_link(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b1,
                                      uint32_t b2, double cb, double len, double weight);     /* meca_inter.cc:943-1059 */
int         cymg_inter_sliding_link_ie(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b,
                                       double weight);                                        /* meca_inter.cc:1075-1140 */
int         cymg_inter_sliding_link_ii(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b1,
                                       uint32_t b2, double cb, double weight);                /* meca_inter.cc:1153-1258 */
int         cymg_inter_side_sliding_link_ie(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b,
                                            double len, double weight);                       /* meca_inter.cc:1266-1499 */
int         cymg_inter_side_sliding_link_ii(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b1,
                                            uint32_t b2, double cb, double len, double weight); /* meca_inter.cc:1508-1751 */
/* 2D only (interTorque2D); the interpolation coefficients play no part */
int         cymg_inter_torque(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2,
                              double cb, double cosinus, double sinus, double weight);        /* meca_inter.cc:1772-1869 */
int         cymg_inter_side_clamp(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, const double* g,
                                  double len, double weight);                                 /* meca_inter.cc:2032-2226 */
int         cymg_inter_sliding_clamp(cymg_context* ctx, uint32_t a1, uint32_t a2, double ca, const double* g,
                                     double weight);                                          /* meca_inter.cc:2241-2283 */

/* steric interactions: Simul::setStericInteractions (simul_solve.cc:139-188) with
   PointGrid::setInteractions (point_grid.cc:361-413), on the device, placed in the
   call order like the calls above.  Items: the segments of the steric fibers
   (FiberProp::steric_radius, steric_range) and the points of the steric Solids whose
   radius exceeds REAL_EPSILON, in registration order; the grid covers
   Space::extension() with cells of steric_max_range (0: estimated as
   Simul::estimateStericRange does, with the fibers' segmentation).  The checks make
   the reference's interLongLink / interSideSlidingLink calls, evaluated from the
   device points.  Where the reference orders two segments by the address of their
   records (point_grid.cc:205,295) the cell-major order of the items is used.  One
   device-to-host copy per call (the number of interactions). */
typedef struct {
    double extension[3];   /* Space::extension() (the third ignored in 2D) */
    double max_range;      /* SimulProp::steric_max_range; 0: estimated */
    double stiff_push;     /* SimulProp::steric_stiffness_push[0] */
    double stiff_pull;     /* SimulProp::steric_stiffness_pull[0] */
} cymg_steric_params;
/* per Mecable in registration order: steric flag (FiberProp/SolidProp::steric), and
   for fibers steric_radius and steric_range (ignored for Solids) */
int         cymg_set_steric(cymg_context* ctx, const cymg_steric_params* params, uint32_t nmec,
                            const int* steric, const double* radius, const double* range);
int         cymg_steric_interactions(cymg_context* ctx);

/* device-side emitters, placed in the call order like the calls above.
   Fiber::setInteractions with CONFINE_INSIDE in a SpaceSphere centred at the
   origin, for every registered fiber (fiber.cc:464-481, space_sphere.cc:113-116,148-151) */
int         cymg_confine_sphere(cymg_context* ctx, double radius, double stiffness);
/* Couple::setInteractions -> interLink(PI, PI) (couple.cc:87-92) for a list of
   doubly-attached couples; hands are (fiber index in registration order,
   abscissa from the minus end), interpolated as FiberNaked::interpolate
   (fiber_naked.cc:1050-1065).  The list is uploaded by set_couples and
   emitted on the device by couples_interactions. */
int         cymg_set_couples(cymg_context* ctx, uint32_t n, const uint32_t* fiber_a,
                             const double* abscissa_a, const uint32_t* fiber_b,
                             const double* abscissa_b, const double* stiffness);
int         cymg_couples_interactions(cymg_context* ctx);
/* motor couples (Couple::stepAA / stepAF / stepFA, couple.cc:125-150, with
   Motor::stepLoaded / stepUnloaded, motor.cc:19-57): per uploaded couple a
   state byte, bit 1 hand 1 attached, 2 hand 2 attached, 4 hand 1 is a motor,
   8 hand 2 is a motor (cymg_set_couples sets 3: attached crosslinkers).  Only
   couples with both hands attached interact (couple.cc:87-92).  Stepping
   uses the current points (after cymg_apply): f = k (pos2 - pos1), a loaded
   motor moves by max_speed*dt + (f . dir)*|max_speed|*dt/stall_force, clamped to
   [0, 2 max_speed dt] when limit_speed; a motor whose partner is free moves by
   max_speed*dt; past a fiber end it holds the end (hold_growing_end) or
   detaches (hand.cc:102-122, fibers that neither grow nor shrink).  The
   random unbinding tests are not modelled (unbinding off). */
typedef struct {
    double max_speed;        /* MotorProp::max_speed, um/s */
    double stall_force;      /* MotorProp::stall_force, pN, > 0 */
    int    limit_speed;      /* MotorProp::limit_speed */
    int    hold_growing_end; /* HandProp::hold_growing_end */
} cymg_motor_props;
int         cymg_set_couple_state(cymg_context* ctx, const uint8_t* state);
int         cymg_get_couples(cymg_context* ctx, double* aa, double* ab, uint8_t* state);
int         cymg_step_couples(cymg_context* ctx, const cymg_motor_props* props);

/* Segments near places: FiberGrid::nearbySegments (fiber_grid.cc:302-337) for
   nq places at once, with FiberLocus::projectPoint (fiber_locus.cc:62-97) on
   the current points; inclusive != 0 applies tryToAttach's range test instead
   (fiber_grid.cc:273-274: dis <= range^2 is kept, nearbySegments keeps
   dis < range^2).  places: host [nq][dim]; exclude: host [nq] fiber to skip per
   place (nearbySegments' `exclude`), or NULL.  Writes counts[nq] and *total;
   when *total <= capacity, also the kept segments place after place, each
   place's in fiber order then point order (the reference's list order):
   fiber, the segment's first point in the fiber, the abscissa from the minus
   end (segmentation * point + projection, the FiberBinder site of
   fiber_grid.cc:278) and the squared distance.  Fibers only (Solids have no
   segments); no periodic space.  At most 64 segments per place
   (CYMG_ERR_UNSUPPORTED beyond). */
int         cymg_nearby_segments(cymg_context* ctx, uint32_t nq, const double* places, const uint32_t* exclude,
                                 double range, int inclusive, uint32_t* counts, uint32_t capacity,
                                 uint32_t* fiber, uint32_t* point, double* abscissa, double* dist_sqr,
                                 uint32_t* total);
/* cymg_nearby_segments without allocating: no counterpart in the reference;
   for callers that cannot allocate device memory (code under HIP graph
   capture, serving loops, harnesses that forbid hipMalloc).  Same arguments,
   results and errors as cymg_nearby_segments; its device memory (input and
   output staging, cell buckets, counts, sort work space) is one fixed array of
   CYMG_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, so it
   refuses (CYMG_ERR_UNSUPPORTED) more than CYMG_NOALLOC_MAX_PLACES places,
   a layout of more than CYMG_NOALLOC_MAX_SEGMENTS fiber segments, or more than
   CYMG_NOALLOC_MAX_KEPT kept segments when they would be written out.  Every
   call on a GPU shares that GPU's array: calls that overlap on one GPU (from
   two contexts or threads) race. */
#define CYMG_NOALLOC_MAX_PLACES    65536
#define CYMG_NOALLOC_MAX_SEGMENTS  131072
#define CYMG_NOALLOC_MAX_KEPT      262144
#define CYMG_NOALLOC_SCRATCH_BYTES 25165824
int         cymg_nearby_segments_noalloc(cymg_context* ctx, uint32_t nq, const double* places,
                                         const uint32_t* exclude, double range, int inclusive, uint32_t* counts,
                                         uint32_t capacity, uint32_t* fiber, uint32_t* point, double* abscissa,
                                         double* dist_sqr, uint32_t* total);
/* Aster::setInteractions (aster.cc:53-89) for a list of aster links, uploaded
   by set_aster_links and emitted on the device by aster_interactions, each as
   the host interLink call it stands for:
     kind 0: interLink(PointExact(solid point s1), minus end of `fiber`)
     kind 1: interLink(`fiber` interpolated at abscissa ab from its minus end,
             PointExact(solid point s1))
     kind 2: interLink(PointInterpolated(solid points s1, s2, coefficient ab),
             plus end of `fiber`)  (a fiber shorter than the clamp distance)
   s1, s2 are matIndex values of Solid points; w the stiffness. */
int         cymg_set_aster_links(cymg_context* ctx, uint32_t n, const int* kind, const uint32_t* s1,
                                 const uint32_t* s2, const uint32_t* fiber, const double* ab, const double* w);
int         cymg_aster_interactions(cymg_context* ctx);

/* Meca::solve (meca.cc:598-845) and its phases:
   setup   (meca.cc:605-693): matrices, forces, Brownian terms, RHS, P'
   iterate (meca.cc:704-799): preconditioner, BCGSP/BCGS, rescue chain
   apply   (meca.cc:802-835): vPTS += vSOL, forces, getPoints (reshape)
   gauss: N(0,1) per degree of freedom in object order (the draws of
   RigidFiber::addBrownianForces, rigid_fiber.cc:90-98), host or device memory
   per params->gauss_on_device; NULL means no Brownian term. */
int         cymg_solve(cymg_context* ctx, const cymg_solve_params* params, const double* gauss,
                       cymg_solve_report* report);
int         cymg_setup(cymg_context* ctx, const double* gauss, int gauss_on_device);
int         cymg_iterate(cymg_context* ctx, const cymg_solve_params* params,
                         cymg_solve_report* report);
/* needs a successful cymg_iterate since the last prepare/setup; the timing
   calls (cymg_autotune, cymg_time_kernel) reuse the solver's vectors, so they
   also require a new iterate before apply */
int         cymg_apply(cymg_context* ctx);

/* operators, valid after setup (host arrays of DIM*nbPts) */
int         cymg_multiply(cymg_context* ctx, const double* X, double* Y);       /* Meca::multiply, meca.cc:157-181 */
int         cymg_compute_preconditioner(cymg_context* ctx, int precision);     /* Meca::computePreconditionner, meca.cc:411-444 */
int         cymg_precondition(cymg_context* ctx, const double* X, double* Y);   /* Meca::precondition, meca.cc:448-466 */
int         cymg_compute_forces(cymg_context* ctx);                             /* Meca::computeForces(), meca.cc:550-580 */
/* same on device pointers, on the context's stream */
int         cymg_multiply_device(cymg_context* ctx, const double* dX, double* dY);
int         cymg_precondition_device(cymg_context* ctx, const double* dX, double* dY);

/* export */
/* CYMG_VEC_PREV (no counterpart in the reference): the previous step's solution
   as the current registration sees it, the seed of CYMG_GUESS_PREVIOUS */
enum { CYMG_VEC_PTS = 0, CYMG_VEC_SOL, CYMG_VEC_BAS, CYMG_VEC_RHS, CYMG_VEC_FOR, CYMG_VEC_PREV };
int         cymg_get_vector(cymg_context* ctx, int which, double* out);  /* vPTS..vFOR (meca.h:101-106) */
/* Meca::dump (meca.cc:957-990): ddim.txt, diagonal.txt, drhs.bin, dsol.bin, dpts.bin and
   dmat.bin (the full operator, column by column) in `directory` (NULL: the working directory);
   after cymg_setup.  For debugging small systems: dmat.bin holds size^2 doubles. */
int         cymg_dump(cymg_context* ctx, const char* directory);
/* trajectory frames: Simul::writeObjects (simul_file.cc:402-489) in binary mode, with the
   fiber and solid sections (ObjectSet::write, object_set.cc:380-389; Object::writeReference,
   object.cc:76-97; FiberNaked::write, fiber_naked.cc:1195-1202; PointSet::write,
   point_set.cc:337-342; Solid::write, solid.cc:896-905) packed on the device from the
   points of the last solve (or prepare).  Per Mecable in registration order, what the
   records carry beside the points: */
typedef struct {
    uint32_t number;       /* Object::number() (> 0) */
    int32_t  mark;         /* Object::mark() */
    uint32_t prop;         /* property()->index() (< 65536) */
    uint32_t signature;    /* FiberNaked::fnSignature (ignored for Solids) */
    double   cut_wished;   /* FiberNaked::fnCutWished = FiberProp::segmentation (fibers) */
    double   abscissa;     /* FiberNaked::fnAbscissa, the minus end (fibers) */
} cymg_frame_record;
int         cymg_set_frame_records(cymg_context* ctx, uint32_t nmec, const cymg_frame_record* records);
/* CYMG_FRAME_FLOAT writes every real as the reference does (OutputWrapper::writeFloat,
   iowrapper.cc:535-547: 4-byte floats), byte for byte; CYMG_FRAME_DOUBLE is the lossless
   form of the same frame: every real as 8 bytes (OutputWrapper::writeDouble,
   iowrapper.cc:570-582), the time as %.17g, and ", real 8" after the format on the
   #time line */
#define CYMG_FRAME_FLOAT  4
#define CYMG_FRAME_DOUBLE 8
typedef struct {
    double      time;        /* Simul::simTime() */
    const char* date;        /* TicToc::date (24 characters); NULL: empty */
    int         real_bytes;  /* CYMG_FRAME_FLOAT or CYMG_FRAME_DOUBLE */
    const void* before;      /* the sections written before the fibers (space, field) as bytes */
    size_t      nbefore;
    const void* after;       /* the sections after the solids (bead, sphere, single, couple, organizer) */
    size_t      nafter;
} cymg_frame_params;
/* the whole frame into out[capacity]; *size = its length (capacity 0: size only) */
int         cymg_write_frame(cymg_context* ctx, const cymg_frame_params* params, void* out, size_t capacity,
                             size_t* size);
/* reads the first frame of data[size] written by cymg_write_frame (either form) or by
   Simul::writeObjects in binary mode: for each record of its fiber and solid sections the
   metadata, and its points converted to `dim` coordinates as InputWrapper::readFloatVector
   does (iowrapper.cc:233-283: missing coordinates are 0), with the Solid radii (0 for fiber
   points); other sections are skipped to their end.  *used = bytes up to the end of the
   frame.  *nobjects and *npoints count the whole frame even where they exceed max_objects /
   max_points (only that many are stored: call with 0 to learn the sizes).  Host code only,
   no context: errors are reported by cymg_last_error(NULL). */
typedef struct {
    char     tag;            /* 'f' fiber, 'd' solid */
    uint32_t prop, number;
    int32_t  mark;
    uint32_t signature;
    double   length, cut_wished, abscissa;
    uint32_t points;         /* number of points */
    uint64_t first;          /* its first point in points[] */
} cymg_frame_object;
int         cymg_read_frame(const void* data, size_t size, int dim, cymg_frame_object* objects,
                            uint32_t max_objects, uint32_t* nobjects, double* points, double* radii,
                            size_t max_points, size_t* npoints, double* time, size_t* used);
int         cymg_get_points(cymg_context* ctx, double* out);   /* Mecable points after getPoints (fiber_naked.cc:414-425) */
int         cymg_get_tensions(cymg_context* ctx, double* out); /* RigidFiber::tension, one per segment (rigid_fiber.h:90) */
int         cymg_get_block(cymg_context* ctx, uint32_t mecable, int inverse, double* out); /* column-major bs*bs */
/* mB (which=0) or mC (which=1) as sorted unique lower-triangle entries (row>=col),
   column-major order as MatrixSparseSymmetric1 stores them (matsparsesym1.cc:137-224).
   Returns the count in *nnz; arrays may be NULL to query it. */
int         cymg_get_matrix(cymg_context* ctx, int which, uint32_t* nnz, uint32_t* rows,
                            uint32_t* cols, double* vals);

/* implementation choices (no counterpart in the reference; results agree to
   the solver tolerance, see the parity tests):
   CYMG_OPT_KRYLOV_LAUNCHES: 2 (default, recurrences for the preconditioned
     vectors) or 4 launches per BiCGStab iteration;
   CYMG_OPT_BLOCK_INVERSE: 0 Gauss-Jordan with the block in registers
     (blocks beyond 96 unknowns in FP64, 128 in FP32, use 2),
     1 getrf + getri order in LDS, 2 Gauss-Jordan in LDS, 3 (default)
     blocked Gauss-Jordan in LDS with the panel updates on the matrix cores
     (blocks beyond 96 unknowns use 2);
   CYMG_OPT_TILE_POINTS: target points per tile of whole fibers in the
     two-launch iteration, 1..170 (1: one fiber per tile; above 256/DIM points
     each unknown gets one thread instead of two, which suits many tiles),
     0 (default): chosen from the layout (or by cymg_autotune);
   CYMG_OPT_GRAPHS: 1 (default) replays chunks of 2 and 8 iterations of the
     two-launch iteration as captured HIP graphs, 0 launches every kernel;
   CYMG_OPT_TILE_THREADS: 512 (default) or 256 threads per tile of the
     two-launch iteration; 256 (tiles of at most 256/DIM points, up to four
     tiles resident per compute unit, a Solid's tile four rows per thread)
     applies to layouts whose fibers all fit such a tile and whose Solids all
     have tiles of their own, others keep 512;
   CYMG_OPT_GUESS_GUARD: 1 (default) replaces a learned guess whose residual
     |b - A x0|_inf is not below |b|_inf (that of the zero guess) by zero, on
     the device, and reports guess_used = CYMG_GUESS_ZERO; 0 keeps every guess;
   CYMG_OPT_FP32_APPLY: 0 (default) applies the FP32 blocks of
     CYMG_PRECOND_FP32 to the FP64 operand with FP64 accumulation (a fixed
     linear operator, which the two-launch iteration's recurrences need);
     1 applies them in FP32 arithmetic (operand rounded to FP32, FP32
     accumulation), in the solve (which then runs the four-launch iteration,
     applying M^-1 directly) and in cymg_precondition;
   CYMG_OPT_BLOCK_STAGE: staged launches of the two-launch iteration's
     512-thread tiles (fiber-only layouts): four more wavefronts per tile copy
     the tile's block inverses into LDS (up to 128 KB per tile) from the start
     of each launch, for the block apply to read there; same results bit for
     bit.  1 (default) stages FP32 blocks (CYMG_PRECOND_FP32), 2 FP32 and
     FP64 blocks, 0 none (the blocks are read from memory) */
enum { CYMG_OPT_KRYLOV_LAUNCHES = 1, CYMG_OPT_BLOCK_INVERSE = 2, CYMG_OPT_TILE_POINTS = 3, CYMG_OPT_GRAPHS = 4,
       CYMG_OPT_TILE_THREADS = 5, CYMG_OPT_GUESS_GUARD = 6, CYMG_OPT_FP32_APPLY = 7, CYMG_OPT_BLOCK_STAGE = 8 };
int         cymg_set_option(cymg_context* ctx, int option, int value);
int         cymg_get_option(cymg_context* ctx, int option, int* value);   /* the value in use */
/* autotuner (no counterpart in the reference): on the current system (after
   setup), times the block-inverse methods (Gauss-Jordan in registers / in
   LDS / blocked on the matrix cores) and the tile shapes of the two-launch iteration (target points per
   tile 16..170 and the layout default, for 512- and 256-thread tiles, `reps` launches each) and keeps the
   fastest; with 512-thread tiles of a fiber-only layout it then times the current precision's launch
   staged and unstaged and sets CYMG_OPT_BLOCK_STAGE accordingly (FP64 blocks: 2 or 1; FP32
   blocks: 1 (or 2) or 0); reports the choices and the chosen launch time (ms) */
int         cymg_autotune(cymg_context* ctx, int reps, int* tile_points, int* block_inverse, double* launch_ms);

/* device timings of the last solve, milliseconds */
typedef struct {
    double assemble_ms;     /* stream sort/reduce, CSR build */
    double setup_ms;        /* forces, Brownian, RHS, P' */
    double precond_ms;      /* block build + factorisation */
    double iterate_ms;      /* BiCGStab iterations */
    double apply_ms;        /* update + reshape */
    double matvec_avg_ms;   /* average duration of one fused operator launch */
    double pcapply_avg_ms;  /* average duration of one preconditioner apply launch */
    int    matvec_count;
} cymg_timings;
int         cymg_get_timings(cymg_context* ctx, cymg_timings* out);
int         cymg_enable_timing(cymg_context* ctx, int enable);
/* average duration (ms) of one launch of a hot kernel, timed with HIP events
   around `reps` back-to-back launches on the context's stream, on the current
   system state: which = 0 fused operator (Meca::multiply), 1 block apply
   (Meca::precondition), 2 block build + factorisation (FP64), 3 the same in FP32,
   4 block build without factorisation, 5 build with getrf + getri instead of Gauss-Jordan,
   6 one launch of the two-launch BiCGStab iteration with the current preconditioner
     (`reps` launches after an init launch, zero threshold; no counterpart in the reference),
   7 the same without preconditioner */
int         cymg_time_kernel(cymg_context* ctx, int which, int reps, double* avg_ms);

/* learned initial guess (no counterpart in the reference, which seeds BCGS with
   vSOL = 0, meca.cc:704).  A per-fiber network (inputs: the fiber's points
   relative to their centroid / 5, rhs / s, previous solution / s and the
   block-Jacobi correction M^-1 (b - A M^-1 b) / s, which carries what the
   linked fibers add (the fiber's own block is solved exactly by M^-1 b),
   s = the fiber's |rhs|_inf, 31 points x 3 each, zero-padded to 384; three hidden
   layers of 256 with GELU; 96 outputs) predicts the correction c of the
   block-Jacobi solution: x0 = M^-1 b + s*c (M^-1 = 0 without preconditioner).
   Fibers of more than 31 points are seeded with M^-1 b.  GEMMs run on BF16
   MFMA with FP32 accumulation, weights are FP32 with BF16 copies.  The solver
   iterates from x0 to the same stop test as from any other guess.
   Weights are flat FP32, layer by layer: W1 [384][256], b1 [256], W2 [256][256],
   b2, W3 [256][256], b3, W4 [256][96], b4 [96]; W[in][out] row-major. */
#define CYMG_GUESS_INPUTS   384
#define CYMG_GUESS_OUTPUTS  96
#define CYMG_GUESS_POINTS   31
#define CYMG_GUESS_WEIGHTS  254816
int         cymg_guess_init(cymg_context* ctx, uint64_t seed);             /* uniform(+-1/sqrt(in)) weights, zero biases */
int         cymg_guess_weights(cymg_context* ctx, int set, float* w);      /* set != 0: load w; else copy the weights out */
/* record (features, targets) of every fiber of each converged solve into a
   ring of `capacity` samples, discarding earlier samples; 0 stops recording and
   keeps the samples for training.  Targets are (x - M^-1 b)/s. */
int         cymg_guess_record(cymg_context* ctx, uint32_t capacity);
int         cymg_guess_samples(cymg_context* ctx, uint32_t* count);
/* `steps` Adam steps (beta 0.9/0.999, eps 1e-8) on random minibatches of
   `batch` recorded samples (batch a multiple of 16); loss of the last step:
   sum of squared errors over the valid outputs / batch */
int         cymg_guess_train(cymg_context* ctx, int steps, int batch, double lr, uint64_t seed, double* loss);
/* the same loss for the current weights on `batch` recorded samples drawn with `seed`, without
   a training step (a held-still measure of progress; the training loss is one minibatch's) */
int         cymg_guess_eval(cymg_context* ctx, int batch, uint64_t seed, double* loss);
/* forward + backward on host inputs X [batch][384] (rounded to BF16) and
   targets T [batch][96] (T[.][95] = number of valid outputs); loss and the
   gradient of every weight (CYMG_GUESS_WEIGHTS floats) */
int         cymg_guess_check(cymg_context* ctx, const float* X, const float* T, int batch, double* loss,
                             float* grads);
/* features [nfib][384] (as FP32 values of the BF16 inputs) and scales [nfib] of
   the last iterate() that used or recorded the learned guess */
int         cymg_guess_features(cymg_context* ctx, float* X, float* scale);
/* the learned initial guess x0 = M^-1 b + s*c of the current system (after
   cymg_setup) without solving, and optionally its base M^-1 b; builds the
   preconditioner of the given precision first (0: M^-1 = 0) */
int         cymg_guess_predict(cymg_context* ctx, int precondition, double* x0, double* base);

#ifdef __cplusplus
}
#endif
#endif
