/*
 * meca_oracle.c -- CPU restatement of Cytosim's implicit mechanics step (Meca).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the
 * single-threaded CPU baseline ("port") timed beside it.  Nothing in the
 * product links or calls this file.
 *
 * Written from the reference's behaviour, function by function; each block
 * cites the reference file:line it restates (paths relative to the reference
 * tree).  FP64 throughout, as the reference (`real` = double, src/math/real.h).
 * The reference calls LAPACK/BLAS (Fortran ABI, not vendored); the routines it
 * uses are restated here from LAPACK's documented algorithms (dgetrf/dgetf2,
 * dgetri/dtrtri, dpttrf, dptts2, dgtsv) and pinned in tests against scipy's
 * LAPACK.
 *
 * Two Mecable kinds are modelled: RigidFiber and Solid (the aster cores of the
 * aster configuration), registered fibers first as Simul::setInteractions
 * does (simul_solve.cc:204-207).  Brownian draws are an input (N(0,1) per degree of freedom,
 * in object order), so the SFMT stream of the reference is not restated.
 */
#include "meca_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define REAL_EPSILON (128.0 * DBL_EPSILON) /* src/math/real.h:32 */

/* ------------------------------------------------------------------------ */
/* symmetric sparse matrix, lower storage per column (matsparsesym1.{h,cc})  */

typedef struct { double val; uint32_t line; } sp_elem;

typedef struct {
    uint32_t  size, alloc;
    sp_elem** col;
    uint32_t* colSize;
    uint32_t* colMax;
    /* Numerical-Recipes layout built by prepareForMultiply */
    uint32_t  nmax;
    uint32_t* ija;
    double*   sa;
    uint32_t* colF;
} spsym;

/* MatrixSparseSymmetric1::allocate (matsparsesym1.cc:35-74) */
static void sp_allocate(spsym* m, uint32_t sz)
{
    m->size = sz;
    if (sz > m->alloc) {
        m->col     = (sp_elem**)realloc(m->col, sizeof(sp_elem*) * sz);
        m->colSize = (uint32_t*)realloc(m->colSize, sizeof(uint32_t) * sz);
        m->colMax  = (uint32_t*)realloc(m->colMax, sizeof(uint32_t) * sz);
        for (uint32_t i = m->alloc; i < sz; ++i) {
            m->col[i] = NULL;
            m->colSize[i] = 0;
            m->colMax[i] = 0;
        }
        m->colF = (uint32_t*)realloc(m->colF, sizeof(uint32_t) * (sz + 1));
        m->alloc = sz;
    }
}

static void sp_free(spsym* m)
{
    for (uint32_t i = 0; i < m->alloc; ++i) free(m->col[i]);
    free(m->col); free(m->colSize); free(m->colMax);
    free(m->ija); free(m->sa); free(m->colF);
    memset(m, 0, sizeof(*m));
}

/* makeZero (matsparsesym1.cc:246-250) */
static void sp_zero(spsym* m)
{
    for (uint32_t i = 0; i < m->size; ++i) m->colSize[i] = 0;
}

/* allocateColumn: capacity grows to a multiple of 4 (matsparsesym1.cc:97-124) */
static sp_elem* sp_col_reserve(spsym* m, uint32_t j, uint32_t need)
{
    if (need > m->colMax[j]) {
        uint32_t cap = (need + 3u) & ~3u;
        sp_elem* c = (sp_elem*)malloc(sizeof(sp_elem) * cap);
        if (m->col[j]) {
            memcpy(c, m->col[j], sizeof(sp_elem) * m->colMax[j]);
            free(m->col[j]);
        }
        m->col[j] = c;
        m->colMax[j] = cap;
    }
    return m->col[j];
}

/* operator()(i,j): element address in the lower triangle, created on demand;
   the diagonal is the first element of a column, rows are kept ascending
   (matsparsesym1.cc:137-224) */
static double* sp_ref(spsym* m, uint32_t i, uint32_t j)
{
    if (j > i) { uint32_t t = i; i = j; j = t; }
    if (m->colSize[j] == 0) {
        sp_elem* c = sp_col_reserve(m, j, 2);
        c[0].line = j; c[0].val = 0.0;
        m->colSize[j] = 1;
        if (i != j) {
            c[1].line = i; c[1].val = 0.0;
            m->colSize[j] = 2;
            return &c[1].val;
        }
        return &c[0].val;
    }
    sp_elem* c = m->col[j];
    if (i == j) return &c[0].val;
    uint32_t n = m->colSize[j];
    uint32_t k = 1;
    while (k < n) {
        if (c[k].line == i) return &c[k].val;
        if (c[k].line > i) break;
        ++k;
    }
    c = sp_col_reserve(m, j, n + 1);
    memmove(c + k + 1, c + k, sizeof(sp_elem) * (n - k));
    c[k].line = i;
    c[k].val = 0.0;
    m->colSize[j] = n + 1;
    return &c[k].val;
}

/* nonZero (matsparsesym1.cc:342-352) */
static int sp_nonzero(const spsym* m)
{
    for (uint32_t j = 0; j < m->size; ++j)
        for (uint32_t k = 0; k < m->colSize[j]; ++k)
            if (m->col[j][k].val != 0.0) return 1;
    return 0;
}

/* setColF(false) + prepareForMultiply (matsparsesym1.cc:445-523) */
static void sp_prepare(spsym* m)
{
    uint32_t n = m->size;
    uint32_t inx = n;
    m->colF[n] = n;
    for (int64_t j = (int64_t)n - 1; j >= 0; --j) {
        if (m->colSize[j] > 0) inx = (uint32_t)j;
        m->colF[j] = inx;
    }
    uint32_t nbe = 1;
    for (uint32_t j = 0; j < n; ++j) nbe += m->colSize[j] > 0 ? m->colSize[j] : 1;
    if (nbe > m->nmax) {
        free(m->ija); free(m->sa);
        m->nmax = nbe + n;
        m->ija = (uint32_t*)malloc(sizeof(uint32_t) * m->nmax);
        m->sa = (double*)malloc(sizeof(double) * m->nmax);
    }
    m->ija[0] = n + 1;
    uint32_t kk = n;
    for (uint32_t j = 0; j < n; ++j) {
        if (m->colSize[j] > 0) {
            m->sa[j] = m->col[j][0].val;
            for (uint32_t c = 1; c < m->colSize[j]; ++c) {
                ++kk;
                m->sa[kk] = m->col[j][c].val;
                m->ija[kk] = m->col[j][c].line;
            }
        } else {
            m->sa[j] = 0.0;
        }
        m->ija[j + 1] = kk + 1;
    }
}

/* vecMulAdd: Y += M X from the lower storage (matsparsesym1.cc:526-542) */
static void sp_mul_add(const spsym* m, const double* X, double* Y)
{
    const uint32_t n = m->size;
    for (uint32_t j = m->colF[0]; j < n; j = m->colF[j + 1]) {
        double x0 = X[j];
        double y0 = Y[j] + m->sa[j] * x0;
        for (uint32_t k = m->ija[j]; k < m->ija[j + 1]; ++k) {
            double a = m->sa[k];
            uint32_t i = m->ija[k];
            Y[i] += a * x0;
            y0 += a * X[i];
        }
        Y[j] = y0;
    }
}

/* vecMulAddIso2D / vecMulAddIso3D (matsparsesym1.cc:575-631) */
static void sp_mul_add_iso(const spsym* m, int dim, const double* X, double* Y)
{
    const uint32_t n = m->size;
    for (uint32_t j = m->colF[0]; j < n; j = m->colF[j + 1]) {
        const uint32_t Dj = dim * j;
        double xj[3], yj[3];
        for (int d = 0; d < dim; ++d) {
            xj[d] = X[Dj + d];
            yj[d] = Y[Dj + d] + m->sa[j] * xj[d];
        }
        for (uint32_t k = m->ija[j]; k < m->ija[j + 1]; ++k) {
            const uint32_t Di = dim * m->ija[k];
            const double a = m->sa[k];
            for (int d = 0; d < dim; ++d) {
                yj[d] += a * X[Di + d];
                Y[Di + d] += a * xj[d];
            }
        }
        for (int d = 0; d < dim; ++d) Y[Dj + d] = yj[d];
    }
}

/* addTriangularBlock: upper triangle of the diagonal block into M (sx*sx,
   column-major) (matsparsesym1.cc:262-285) */
static void sp_add_triangular_block(const spsym* m, double* M, uint32_t x, uint32_t sx)
{
    for (uint32_t j = 0; j < sx; ++j) {
        for (uint32_t k = 0; k < m->colSize[j + x]; ++k) {
            uint32_t i = m->col[j + x][k].line;
            if (x <= i) {
                i -= x;
                if (i < sx) {
                    if (i < j) M[i + sx * j] += m->col[j + x][k].val;
                    else       M[j + sx * i] += m->col[j + x][k].val;
                }
            }
        }
    }
}

/* addDiagonalBlock: full diagonal block into M (matsparsesym1.cc:288-310) */
static void sp_add_diagonal_block(const spsym* m, double* M, uint32_t x, uint32_t sx)
{
    for (uint32_t j = 0; j < sx; ++j) {
        for (uint32_t k = 0; k < m->colSize[j + x]; ++k) {
            uint32_t i = m->col[j + x][k].line;
            if (x <= i) {
                i -= x;
                if (i < sx) {
                    M[i + sx * j] += m->col[j + x][k].val;
                    if (i != j) M[j + sx * i] += m->col[j + x][k].val;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* LAPACK restatements                                                       */

/* dgetrf with the unblocked (dgetf2) right-looking algorithm: partial
   pivoting on the first entry of largest magnitude (idamax), ipiv 1-based.
   Returns info (0, or index of the first exactly-zero pivot). */
int orc_dgetrf(int n, double* a, int lda, int* ipiv)
{
    int info = 0;
    for (int j = 0; j < n; ++j) {
        int p = j;
        double amax = fabs(a[j + lda * j]);
        for (int i = j + 1; i < n; ++i) {
            double v = fabs(a[i + lda * j]);
            if (v > amax) { amax = v; p = i; }
        }
        ipiv[j] = p + 1;
        if (a[p + lda * j] != 0.0) {
            if (p != j)
                for (int k = 0; k < n; ++k) {
                    double t = a[j + lda * k];
                    a[j + lda * k] = a[p + lda * k];
                    a[p + lda * k] = t;
                }
            double piv = a[j + lda * j];
            if (fabs(piv) >= DBL_MIN) {
                double r = 1.0 / piv;
                for (int i = j + 1; i < n; ++i) a[i + lda * j] *= r;
            } else {
                for (int i = j + 1; i < n; ++i) a[i + lda * j] /= piv;
            }
        } else if (info == 0) {
            info = j + 1;
        }
        for (int k = j + 1; k < n; ++k) {
            double u = a[j + lda * k];
            if (u != 0.0)
                for (int i = j + 1; i < n; ++i) a[i + lda * k] -= a[i + lda * j] * u;
        }
    }
    return info;
}

/* dgetri: inv(U) in place (dtrti2, upper, non-unit), then solve
   inv(A) * L = inv(U), then undo the column interchanges. */
int orc_dgetri(int n, double* a, int lda, const int* ipiv)
{
    for (int j = 0; j < n; ++j)
        if (a[j + lda * j] == 0.0) return j + 1;
    /* dtrti2 upper non-unit */
    for (int j = 0; j < n; ++j) {
        a[j + lda * j] = 1.0 / a[j + lda * j];
        double ajj = -a[j + lda * j];
        /* x(0:j-1) = triu(A(0:j-1,0:j-1)) * x  (dtrmv upper, no-trans, non-unit) */
        for (int k = 0; k < j; ++k) {
            double t = a[k + lda * j];
            if (t != 0.0) {
                for (int i = 0; i < k; ++i) a[i + lda * j] += t * a[i + lda * k];
                a[k + lda * j] = t * a[k + lda * k];
            }
        }
        for (int i = 0; i < j; ++i) a[i + lda * j] *= ajj;
    }
    double* work = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int j = n - 1; j >= 0; --j) {
        for (int i = j + 1; i < n; ++i) {
            work[i] = a[i + lda * j];
            a[i + lda * j] = 0.0;
        }
        /* A(:,j) -= A(:,j+1:n) * work(j+1:n)  (dgemv) */
        for (int k = j + 1; k < n; ++k) {
            double w = work[k];
            if (w != 0.0)
                for (int i = 0; i < n; ++i) a[i + lda * j] -= a[i + lda * k] * w;
        }
    }
    free(work);
    for (int j = n - 2; j >= 0; --j) {
        int jp = ipiv[j] - 1;
        if (jp != j)
            for (int i = 0; i < n; ++i) {
                double t = a[i + lda * j];
                a[i + lda * j] = a[i + lda * jp];
                a[i + lda * jp] = t;
            }
    }
    return 0;
}

/* dpttrf: L*D*L' of a symmetric positive definite tridiagonal matrix;
   d (n) becomes D, e (n-1) becomes the subdiagonal of L. */
int orc_dpttrf(int n, double* d, double* e)
{
    for (int i = 0; i < n - 1; ++i) {
        if (d[i] <= 0.0) return i + 1;
        double ei = e[i];
        e[i] = ei / d[i];
        d[i + 1] -= e[i] * ei;
    }
    if (n > 0 && d[n - 1] <= 0.0) return n;
    return 0;
}

/* dptts2 with one right-hand side: solve L*D*L' x = b in place. */
void orc_dptts2(int n, const double* d, const double* e, double* b)
{
    if (n <= 1) {
        if (n == 1) b[0] *= 1.0 / d[0];
        return;
    }
    for (int i = 1; i < n; ++i) b[i] -= b[i - 1] * e[i - 1];
    b[n - 1] /= d[n - 1];
    for (int i = n - 2; i >= 0; --i) b[i] = b[i] / d[i] - b[i + 1] * e[i];
}

/* dgtsv with one right-hand side: Gaussian elimination with partial pivoting
   on a general tridiagonal matrix (dl: n-1 sub, d: n diag, du: n-1 super). */
int orc_dgtsv(int n, double* dl, double* d, double* du, double* b)
{
    if (n <= 0) return 0;
    for (int i = 0; i + 2 < n; ++i) {
        if (fabs(d[i]) >= fabs(dl[i])) {
            if (d[i] == 0.0) return i + 1;
            double f = dl[i] / d[i];
            d[i + 1] -= f * du[i];
            b[i + 1] -= f * b[i];
            dl[i] = 0.0;
        } else {
            double f = d[i] / dl[i];
            d[i] = dl[i];
            double t = d[i + 1];
            d[i + 1] = du[i] - f * t;
            dl[i] = du[i + 1];
            du[i + 1] = -f * dl[i];
            du[i] = t;
            t = b[i];
            b[i] = b[i + 1];
            b[i + 1] = t - f * b[i + 1];
        }
    }
    if (n > 1) {
        int i = n - 2;
        if (fabs(d[i]) >= fabs(dl[i])) {
            if (d[i] == 0.0) return i + 1;
            double f = dl[i] / d[i];
            d[i + 1] -= f * du[i];
            b[i + 1] -= f * b[i];
        } else {
            double f = d[i] / dl[i];
            d[i] = dl[i];
            double t = d[i + 1];
            d[i + 1] = du[i] - f * t;
            du[i] = t;
            t = b[i];
            b[i] = b[i + 1];
            b[i + 1] = t - f * b[i + 1];
        }
    }
    if (d[n - 1] == 0.0) return n;
    b[n - 1] /= d[n - 1];
    if (n > 1) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / d[n - 2];
    for (int i = n - 3; i >= 0; --i)
        b[i] = (b[i] - du[i] * b[i + 1] - dl[i] * b[i + 2]) / d[i];
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Meca state                                                                */

typedef struct {
    uint32_t np, idx;       /* nbPoints, matIndex */
    uint32_t seg;           /* offset of the first segment in the segment arrays */
    double   cut;           /* fnCut = segmentation() */
    double   rigidity;      /* FiberProp::rigidity */
    double   rf_rigidity;   /* rigidity / cut^3 (fiber.cc:459) */
    double   mobility;      /* rfMobility (fiber.cc:446) */
    double   hr0, hr1;      /* FiberProp::hydrodynamic_radius (0: the Meca's parameters) */
    double   height;        /* FiberProp::cylinder_height with surface_effect (0: volume drag) */
    double*  blk;           /* preconditioner block (mecable.cc:15-34) */
    uint32_t blk_alloc;
    int      use_blk;
    int      kind;          /* 0 RigidFiber, 1 Solid */
    /* Solid (solid.h): soRadius, soShape (centred reference), soShapeSqr,
       soReshapeTimer, soDrag, soDragRot, soMom2D, soCenter, soMom */
    double*  radius;
    double*  shape;
    double   shape_sqr;
    int      timer;
    double   drag, drag_rot, mom2d;
    double   center[3];
    double   mom[9];
} orc_fiber;

struct orc_meca {
    int        dim;
    orc_params prm;
    double     time_step;
    orc_fiber* fib;
    int        nfib, fib_alloc;
    uint32_t   npts, nseg, largest;
    uint32_t   vec_alloc;
    double    *vPTS, *vSOL, *vBAS, *vRHS, *vFOR, *vTMP;
    double    *pos;         /* Mecable-held points (psPos) */
    double    *diff;        /* rfDiff, DIM per segment */
    double    *jd, *je;     /* mtJJt, mtJJt2 after pttrf */
    double    *lag, *ltmp;  /* rfLag, rfTMP */
    double    *pdiff;       /* mtJJtiJforce */
    uint32_t   seg_alloc;
    spsym      mB, mC;
    int        use_mB, use_mC;
    double     noise_level;
    int        prepared, setup_done;
};

static void* xrealloc(void* p, size_t n) { return realloc(p, n ? n : 1); }

orc_meca* orc_create(int dim)
{
    if (dim != 2 && dim != 3) return NULL;
    orc_meca* m = (orc_meca*)calloc(1, sizeof(orc_meca));
    m->dim = dim;
    m->prm.time_step = 0.005;
    m->prm.kT = 0.0042;
    m->prm.viscosity = 1.0;
    m->prm.hydro_radius0 = 0.0125;
    m->prm.hydro_radius1 = 10.0;
    return m;
}

void orc_destroy(orc_meca* m)
{
    if (!m) return;
    for (int f = 0; f < m->fib_alloc; ++f) {
        free(m->fib[f].blk);
        free(m->fib[f].radius);
        free(m->fib[f].shape);
    }
    free(m->fib);
    free(m->vPTS); free(m->vSOL); free(m->vBAS); free(m->vRHS); free(m->vFOR); free(m->vTMP);
    free(m->pos); free(m->diff); free(m->jd); free(m->je); free(m->lag); free(m->ltmp); free(m->pdiff);
    sp_free(&m->mB); sp_free(&m->mC);
    free(m);
}

void orc_set_params(orc_meca* m, const orc_params* p) { m->prm = *p; }

/* Meca::clear (meca.cc:55-60) */
void orc_clear(orc_meca* m)
{
    for (int f = 0; f < m->fib_alloc; ++f) m->fib[f].np = 0;
    m->nfib = 0;
    m->npts = 0;
    m->nseg = 0;
    m->largest = 0;
    m->prepared = 0;
    m->setup_done = 0;
}

/* Meca::add (meca.cc:67-79): matIndex is the running prefix sum */
static orc_fiber* new_mecable(orc_meca* m)
{
    if (m->nfib == m->fib_alloc) {
        int na = m->fib_alloc ? 2 * m->fib_alloc : 64;
        m->fib = (orc_fiber*)realloc(m->fib, sizeof(orc_fiber) * na);
        memset(m->fib + m->fib_alloc, 0, sizeof(orc_fiber) * (na - m->fib_alloc));
        m->fib_alloc = na;
    }
    return &m->fib[m->nfib];
}

int orc_add_fiber(orc_meca* m, uint32_t np, double segmentation, double rigidity)
{
    return orc_add_fiber_ex(m, np, segmentation, rigidity, 0.0, 0.0, 0.0);
}

int orc_add_fiber_ex(orc_meca* m, uint32_t np, double segmentation, double rigidity, double hr0, double hr1,
                     double height)
{
    if (np < 2) return -1;
    orc_fiber* f = new_mecable(m);
    f->kind = 0;
    f->hr0 = hr0;
    f->hr1 = hr1;
    f->height = height;
    f->np = np;
    f->idx = m->npts;
    f->seg = m->nseg;
    f->cut = segmentation;
    f->rigidity = rigidity;
    m->npts += np;
    m->nseg += np - 1;
    if (m->largest < np) m->largest = np;
    return m->nfib++;
}

/* A Solid with the given per-point radii (Solid::addSphere, solid.cc:352-362;
   radius 0: a point without drag), its reference shape fixed from `points`
   (Solid::fixShape, solid.cc:370-400, with PointSet::calculateMomentum,
   point_set.cc:288-317) and the initial soReshapeTimer (solid.cc:28 draws it
   in [0, 7); the caller gives it). */
int orc_add_solid(orc_meca* m, uint32_t np, const double* radius, const double* points, int timer)
{
    const int D = m->dim;
    if (np < 1) return -1;
    orc_fiber* f = new_mecable(m);
    f->kind = 1;
    f->np = np;
    f->idx = m->npts;
    f->seg = m->nseg;
    f->cut = 0.0;
    f->rigidity = 0.0;
    f->radius = (double*)realloc(f->radius, sizeof(double) * np);
    f->shape = (double*)realloc(f->shape, sizeof(double) * D * np);
    memcpy(f->radius, radius, sizeof(double) * np);
    double cc[3] = { 0, 0, 0 }, pp[3] = { 0, 0, 0 };
    for (int d = 0; d < D; ++d) {
        cc[d] = points[d];
        pp[d] = points[d] * points[d];
    }
    for (uint32_t p = 1; p < np; ++p)
        for (int d = 0; d < D; ++d) {
            cc[d] += points[D * p + d];
            pp[d] += points[D * p + d] * points[D * p + d];
        }
    const double N = (double)np;
    for (int d = 0; d < D; ++d) cc[d] *= 1.0 / N;
    f->shape_sqr = 0.0;
    for (int d = 0; d < D; ++d) f->shape_sqr += pp[d] - N * cc[d] * cc[d];
    for (uint32_t p = 0; p < np; ++p)
        for (int d = 0; d < D; ++d) f->shape[D * p + d] = points[D * p + d] - cc[d];
    f->timer = timer;
    m->npts += np;
    if (m->largest < np) m->largest = np;
    return m->nfib++;
}

uint32_t orc_nb_points(const orc_meca* m) { return m->npts; }
uint32_t orc_mat_index(const orc_meca* m, int f) { return m->fib[f].idx; }

/* FiberNaked::bestNbPoints (fiber_naked.cc:22-30) */
uint32_t orc_best_nb_points(double ratio)
{
    unsigned n = (unsigned)(int)ratio;
    if ((2.0 * n + 1.0) * ratio > 2.0 * n * (n + 1.0)) return n + 2;
    return n + 1;
}

/* Fiber::dragCoefficientVolume (fiber.cc:315-363), dragCoefficientSurface
   (fiber.cc:401-414) and setDragCoefficient (428-447) */
static double fiber_mobility(const orc_meca* m, const orc_fiber* f)
{
    const double r0 = f->hr0 > 0 ? f->hr0 : m->prm.hydro_radius0;
    const double r1 = f->hr1 > 0 ? f->hr1 : m->prm.hydro_radius1;
    double len = f->cut * (double)(f->np - 1);      /* length() = nbSegments()*fnCut */
    double drag;
    if (f->height > 0) {
        drag = 2 * M_PI * m->prm.viscosity * len / acosh(1 + f->height / r0);
    } else {
        double lenc = len;
        if (lenc > r1) lenc = r1;
        double drag_sphere = 6.0 * r0;
        double drag_cyl = 3.0 * len / (log(0.5 * lenc / r0) + 0.312);
        drag = M_PI * m->prm.viscosity * (drag_cyl > drag_sphere ? drag_cyl : drag_sphere);
    }
    return (double)f->np / drag;
}

static int ensure_vectors(orc_meca* m)
{
    if (m->npts > m->vec_alloc) {
        uint32_t a = (m->npts + 3u) & ~3u;    /* meca.cc:511-522 */
        size_t nb = sizeof(double) * (size_t)m->dim * a;
        m->vPTS = (double*)xrealloc(m->vPTS, nb);
        m->vSOL = (double*)xrealloc(m->vSOL, nb);
        m->vBAS = (double*)xrealloc(m->vBAS, nb);
        m->vRHS = (double*)xrealloc(m->vRHS, nb);
        m->vFOR = (double*)xrealloc(m->vFOR, nb);
        m->vTMP = (double*)xrealloc(m->vTMP, nb);
        m->pos  = (double*)xrealloc(m->pos, nb);
        memset(m->vSOL, 0, nb); memset(m->vRHS, 0, nb); memset(m->vFOR, 0, nb);
        m->vec_alloc = a;
    }
    if (m->nseg > m->seg_alloc) {
        uint32_t a = m->nseg;
        m->diff  = (double*)xrealloc(m->diff, sizeof(double) * m->dim * a);
        m->jd    = (double*)xrealloc(m->jd, sizeof(double) * a);
        m->je    = (double*)xrealloc(m->je, sizeof(double) * a);
        m->lag   = (double*)xrealloc(m->lag, sizeof(double) * a);
        m->ltmp  = (double*)xrealloc(m->ltmp, sizeof(double) * a);
        m->pdiff = (double*)xrealloc(m->pdiff, sizeof(double) * a);
        memset(m->lag, 0, sizeof(double) * a);
        m->seg_alloc = a;
    }
    return 0;
}

/* RigidFiber::storeDifferences (rigid_fiber.cc:111-117) and
   RigidFiber::makeProjection (rigid_fiber_project.cc:50-87) */
static int fiber_make_projection(orc_meca* m, orc_fiber* f)
{
    const int D = m->dim;
    const double* P = m->pos + D * f->idx;
    double* dif = m->diff + D * f->seg;
    const double sc = 1.0 / f->cut;
    const uint32_t S = f->np - 1;
    for (uint32_t p = 0; p < D * S; ++p) dif[p] = sc * (P[p + D] - P[p]);
    double* d = m->jd + f->seg;
    double* e = m->je + f->seg;
    for (uint32_t s = 0; s < S; ++s) {
        const double* X = dif + D * s;
        double dd = 0.0;
        for (int k = 0; k < D; ++k) dd += X[k] * X[k];
        d[s] = 2.0 * dd;
        if (s + 1 < S) {
            double od = 0.0;
            for (int k = 0; k < D; ++k) od += X[k] * X[k + D];
            e[s] = -od;
        }
    }
    return orc_dpttrf((int)S, d, e);
}

/* LAPACK dpotf2('U') on a 3x3 column-major matrix: A = U'U, U in the upper
   triangle; returns j+1 if the leading minor j+1 is not positive */
static int potf2_upper3(double* A)
{
    for (int j = 0; j < 3; ++j) {
        double ajj = A[j + 3 * j];
        for (int k = 0; k < j; ++k) ajj -= A[k + 3 * j] * A[k + 3 * j];
        if (!(ajj > 0.0)) return j + 1;
        ajj = sqrt(ajj);
        A[j + 3 * j] = ajj;
        for (int i = j + 1; i < 3; ++i) {
            double v = A[j + 3 * i];
            for (int k = 0; k < j; ++k) v -= A[k + 3 * j] * A[k + 3 * i];
            A[j + 3 * i] = v / ajj;
        }
    }
    return 0;
}

/* LAPACK dpotrs('U') with one right-hand side: U'U x = b */
static void potrs_upper3(const double* U, double* b)
{
    for (int i = 0; i < 3; ++i) {
        double v = b[i];
        for (int k = 0; k < i; ++k) v -= U[k + 3 * i] * b[k];
        b[i] = v / U[i + 3 * i];
    }
    for (int i = 2; i >= 0; --i) {
        double v = b[i];
        for (int k = i + 1; k < 3; ++k) v -= U[i + 3 * k] * b[k];
        b[i] = v / U[i + 3 * i];
    }
}

/* Solid::prepareMecable (solid.cc:707-711): setDragCoefficient (665-704) then
   makeProjection (2D: 757-760, 3D: 789-847); soCenter = centroid() (181-200),
   the radius-weighted centre of the points */
static int solid_prepare(orc_meca* m, orc_fiber* f)
{
    const int D = m->dim;
    const double* P = m->pos + D * f->idx;
    const double eta6 = 6 * M_PI * m->prm.viscosity;
    double drag = 0.0, rot = 0.0, roti = 0.0, cen[3] = { 0, 0, 0 };
    for (uint32_t p = 0; p < f->np; ++p) {
        const double r = f->radius[p];
        if (r > 0) {
            drag += r;
            rot += r * r * r;
            for (int d = 0; d < D; ++d) cen[d] += r * P[D * p + d];
            if (D == 2) roti += r * (P[D * p] * P[D * p] + P[D * p + 1] * P[D * p + 1]);
        }
    }
    if (drag < REAL_EPSILON) return -1;                /* "zero drag in Solid" */
    for (int d = 0; d < D; ++d) cen[d] /= drag;
    f->drag = drag * eta6;
    f->drag_rot = rot * (8 * M_PI * m->prm.viscosity);
    for (int d = 0; d < 3; ++d) f->center[d] = d < D ? cen[d] : 0.0;
    if (D == 2) {
        f->mom2d = f->drag_rot + eta6 * roti - f->drag * (cen[0] * cen[0] + cen[1] * cen[1]);
        if (f->mom2d < REAL_EPSILON) return -1;         /* "zero rotational drag in Solid" */
        return 0;
    }
    double m0 = 0, m3 = 0, m6 = 0, m4 = 0, m7 = 0, m8 = 0;
    for (uint32_t p = 0; p < f->np; ++p) {
        const double r = f->radius[p];
        if (r > 0) {
            const double* x = P + 3 * p;
            const double px = r * x[0], py = r * x[1], pz = r * x[2];
            m0 += px * x[0];
            m3 += px * x[1];
            m6 += px * x[2];
            m4 += py * x[1];
            m7 += py * x[2];
            m8 += pz * x[2];
        }
    }
    m0 *= eta6; m3 *= eta6; m6 *= eta6; m4 *= eta6; m7 *= eta6; m8 *= eta6;
    const double* c = f->center;
    const double diag = m0 + m4 + m8 + f->drag_rot - f->drag * (c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    double* M = f->mom;
    memset(M, 0, sizeof(double) * 9);
    M[0 + 3 * 0] = diag - m0 + f->drag * c[0] * c[0];
    M[0 + 3 * 1] = -m3 + f->drag * c[0] * c[1];
    M[0 + 3 * 2] = -m6 + f->drag * c[0] * c[2];
    M[1 + 3 * 1] = diag - m4 + f->drag * c[1] * c[1];
    M[1 + 3 * 2] = -m7 + f->drag * c[1] * c[2];
    M[2 + 3 * 2] = diag - m8 + f->drag * c[2] * c[2];
    return potf2_upper3(M) ? -1 : 0;                   /* "failed to factorize Solid momentum matrix" */
}

/* Meca::prepare (meca.cc:488-537) with Fiber::prepareMecable (fiber.cc:450-460) */
int orc_prepare(orc_meca* m, const double* points)
{
    ensure_vectors(m);
    const int D = m->dim;
    sp_allocate(&m->mB, m->npts);
    sp_allocate(&m->mC, D * m->npts);
    sp_zero(&m->mB);
    sp_zero(&m->mC);
    memset(m->vBAS, 0, sizeof(double) * D * m->npts);
    m->time_step = m->prm.time_step;
    if (points) memcpy(m->pos, points, sizeof(double) * D * m->npts);
    /* putPoints: vPTS <- Mecable points */
    memcpy(m->vPTS, m->pos, sizeof(double) * D * m->npts);
    for (int i = 0; i < m->nfib; ++i) {
        orc_fiber* f = &m->fib[i];
        if (f->kind == 1) {
            if (solid_prepare(m, f)) return -3;        /* zero drag, or momentum not factorisable */
            continue;
        }
        f->mobility = fiber_mobility(m, f);
        if (fiber_make_projection(m, f)) return -2;   /* "could not build Fiber projection" */
        f->rf_rigidity = f->rigidity / (f->cut * f->cut * f->cut);
    }
    m->prepared = 1;
    m->setup_done = 0;
    return 0;
}

static int mecable_of(const orc_meca* m, uint32_t idx)
{
    int lo = 0, hi = m->nfib - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (m->fib[mid].idx <= idx) lo = mid; else hi = mid - 1;
    }
    return lo;
}

/* FiberNaked::interpolate (fiber_naked.cc:1050-1065), fnAbscissa = 0 */
void orc_interpolate(const orc_meca* m, int fi, double ab, uint32_t* mat1, uint32_t* mat2, double* coef)
{
    const orc_fiber* f = &m->fib[fi];
    if (ab <= 0) { *mat1 = f->idx; *mat2 = f->idx + 1; *coef = 0.0; return; }
    double n;
    double co = modf(ab / f->cut, &n);
    unsigned rd = (unsigned)n;
    if (rd < f->np - 1) { *mat1 = f->idx + rd; *mat2 = f->idx + rd + 1; *coef = co; }
    else { *mat1 = f->idx + f->np - 2; *mat2 = f->idx + f->np - 1; *coef = 1.0; }
}

/* ------------------------------------------------------------------------ */
/* assembly (meca_inter.cc)                                                  */

/* addPureForce(PointExact) (meca_inter.cc:123-134) */
void orc_add_pure_force(orc_meca* m, uint32_t a, const double* force)
{
    for (int d = 0; d < m->dim; ++d) m->vBAS[m->dim * a + d] += force[d];
}

/* interLink(PointExact, PointExact) (meca_inter.cc:196-230), no periodic space.
   neighbours: same Mecable and adjacent or equal points (point_exact.cc:43-47). */
void orc_inter_link_ee(orc_meca* m, uint32_t a, uint32_t b, double w)
{
    int fa = mecable_of(m, a), fb = mecable_of(m, b);
    if (fa == fb && (a == b || a == b + 1 || a + 1 == b)) return;
    *sp_ref(&m->mB, a, a) -= w;
    *sp_ref(&m->mB, a, b) += w;
    *sp_ref(&m->mB, b, b) -= w;
}

/* interLink(PointInterpolated, PointExact) (meca_inter.cc:240-275) */
void orc_inter_link_ie(orc_meca* m, uint32_t a1, uint32_t a2, double ca, uint32_t b, double w)
{
    if (mecable_of(m, a1) == mecable_of(m, b) && (a1 == b || a2 == b)) return;
    const double c[3] = { 1.0 - ca, ca, -1.0 };
    const double cw[3] = { w * c[0], w * c[1], -w };
    const uint32_t inx[3] = { a1, a2, b };
    for (int k = 0; k < 3; ++k)
        for (int l = k; l < 3; ++l)
            *sp_ref(&m->mB, inx[k], inx[l]) -= c[k] * cw[l];
}

/* interLink(PointInterpolated, PointInterpolated) (meca_inter.cc:285-321) */
void orc_inter_link_ii(orc_meca* m, uint32_t a1, uint32_t a2, double ca,
                       uint32_t b1, uint32_t b2, double cb, double w)
{
    if (mecable_of(m, a1) == mecable_of(m, b1) &&
        (a1 == b1 || a1 == b2 || a2 == b1 || a2 == b2)) return;
    const double c[4] = { 1.0 - ca, ca, -(1.0 - cb), -cb };
    const double cw[4] = { w * c[0], w * c[1], w * c[2], w * c[3] };
    const uint32_t inx[4] = { a1, a2, b1, b2 };
    for (int j = 0; j < 4; ++j)
        for (int i = j; i < 4; ++i)
            *sp_ref(&m->mB, inx[i], inx[j]) -= c[j] * cw[i];
}

/* Couple::setInteractions for a list of couples (couple.cc:87-92): each
   doubly-attached couple (state & 3 == 3; state NULL: all) links its two
   hands, interpolated on their fibers (fiber_naked.cc:1050-1065), with
   interLink(PointInterpolated, PointInterpolated); one call per step so that
   the CPU baseline is not timed on its caller's loop */
void orc_inter_couples(orc_meca* m, uint32_t n, const uint32_t* fa, const double* aa, const uint32_t* fb,
                       const double* ab, const double* w, const uint8_t* state)
{
    for (uint32_t k = 0; k < n; ++k) {
        if (state && (state[k] & 3) != 3) continue;
        uint32_t a1, a2, b1, b2;
        double ca, cb;
        orc_interpolate(m, (int)fa[k], aa[k], &a1, &a2, &ca);
        orc_interpolate(m, (int)fb[k], ab[k], &b1, &b2, &cb);
        orc_inter_link_ii(m, a1, a2, ca, b1, b2, cb, w[k]);
    }
}

/* interLongLink(PointExact, PointExact) (meca_inter.cc:334-415), no periodic space */
void orc_inter_long_link_ee(orc_meca* m, uint32_t a, uint32_t b, double len, double w)
{
    const int D = m->dim;
    int fa = mecable_of(m, a), fb = mecable_of(m, b);
    if (fa == fb && (a == b || a == b + 1 || a + 1 == b)) return;
    const uint32_t iA = D * a, iB = D * b;
    double ab[3] = { 0, 0, 0 };
    for (int d = 0; d < D; ++d) ab[d] = m->vPTS[iB + d] - m->vPTS[iA + d];
    double abn = 0.0;
    for (int d = 0; d < D; ++d) abn += ab[d] * ab[d];
    abn = sqrt(abn);
    if (abn < REAL_EPSILON) return;
    for (int d = 0; d < D; ++d) ab[d] /= abn;
    for (int d = 0; d < D; ++d) {
        m->vBAS[iA + d] -= w * ab[d] * len;
        m->vBAS[iB + d] += w * ab[d] * len;
    }
    len /= abn;
    const int cooked = (len > 1.0);
    for (int i = 0; i < D; ++i) {
        double v = cooked ? w * ab[i] * ab[i] : w * (1.0 + len * (ab[i] * ab[i] - 1.0));
        *sp_ref(&m->mC, iA + i, iA + i) -= v;
        *sp_ref(&m->mC, iA + i, iB + i) += v;
        *sp_ref(&m->mC, iB + i, iB + i) -= v;
        for (int j = i + 1; j < D; ++j) {
            v = cooked ? w * ab[i] * ab[j] : w * len * ab[i] * ab[j];
            *sp_ref(&m->mC, iA + i, iA + j) -= v;
            *sp_ref(&m->mC, iA + i, iB + j) += v;
            *sp_ref(&m->mC, iA + j, iB + i) += v;
            *sp_ref(&m->mC, iB + i, iB + j) -= v;
        }
    }
}

/* interClamp(PointExact) (meca_inter.cc:1883-1907), no periodic space */
void orc_inter_clamp_e(orc_meca* m, uint32_t a, const double* g, double w)
{
    *sp_ref(&m->mB, a, a) -= w;
    for (int d = 0; d < m->dim; ++d) m->vBAS[m->dim * a + d] += w * g[d];
}

/* interClamp(PointInterpolated) (meca_inter.cc:1918-1963) */
void orc_inter_clamp_i(orc_meca* m, uint32_t a1, uint32_t a2, double ca, const double* g, double w)
{
    const double c2 = ca, c2w = w * c2;
    const double c1 = 1.0 - ca, c1w = w * c1;
    *sp_ref(&m->mB, a1, a1) -= c1w * c1;
    *sp_ref(&m->mB, a1, a2) -= c2w * c1;
    *sp_ref(&m->mB, a2, a2) -= c2w * c2;
    for (int d = 0; d < m->dim; ++d) {
        m->vBAS[m->dim * a1 + d] += c1w * g[d];
        m->vBAS[m->dim * a2 + d] += c2w * g[d];
    }
}

/* interLongClamp (meca_inter.cc:1977-2022) */
void orc_inter_long_clamp(orc_meca* m, uint32_t a, const double* center, double len, double w)
{
    const int D = m->dim;
    const uint32_t inx = D * a;
    double axis[3] = { 0, 0, 0 };
    for (int d = 0; d < D; ++d) axis[d] = m->vPTS[inx + d] - center[d];
    double an = 0.0;
    for (int d = 0; d < D; ++d) an += axis[d] * axis[d];
    an = sqrt(an);
    if (an < REAL_EPSILON) return;
    for (int d = 0; d < D; ++d) axis[d] /= an;
    double ac = 0.0;
    for (int d = 0; d < D; ++d) ac += axis[d] * center[d];
    if (len < an) {
        len /= an;
        const double facX = w * len * (an + ac);
        const double facC = w * (1.0 - len);
        for (int i = 0; i < D; ++i) {
            *sp_ref(&m->mC, inx + i, inx + i) += w * (len * (1.0 - axis[i] * axis[i]) - 1.0);
            for (int j = i + 1; j < D; ++j)
                *sp_ref(&m->mC, inx + i, inx + j) -= w * len * axis[i] * axis[j];
            m->vBAS[inx + i] += facX * axis[i] + facC * center[i];
        }
    } else {
        const double facX = w * (len + ac);
        for (int i = 0; i < D; ++i) {
            for (int j = i; j < D; ++j)
                *sp_ref(&m->mC, inx + i, inx + j) -= w * axis[i] * axis[j];
            m->vBAS[inx + i] += facX * axis[i];
        }
    }
}

/* interPlane(PointExact) (meca_inter.cc:2301-2326) */
void orc_inter_plane_e(orc_meca* m, uint32_t a, const double* dir, const double* g, double w)
{
    const int D = m->dim;
    const uint32_t inx = D * a;
    double gd = 0.0;
    for (int d = 0; d < D; ++d) gd += g[d] * dir[d];
    const double pr = gd * w;
    for (int i = 0; i < D; ++i) {
        for (int j = i; j < D; ++j)
            *sp_ref(&m->mC, inx + i, inx + j) -= dir[i] * dir[j] * w;
        m->vBAS[inx + i] += pr * dir[i];
    }
}

/* ------------------------------------------------------------------------ */
/* position-dependent emitters on interpolated points (meca_inter.cc)       */
/*                                                                          */
/* A PointInterpolated (point_interpolated.h:97-124) is (p1, p2, c) on one  */
/* Mecable: pos = P1 + c (P2 - P1), diff = P2 - P1, coef1 = c, coef2 = 1-c. */
/* Vector arithmetic as vector2.h / vector3.h: norm = sqrt of the squares  */
/* summed in order, normalized = v * (1 / norm), dot products in order.     */
/* No periodic space (modulo == 0).  Random::sign_exc (random.cc:73-82)    */
/* draws a random sign at exactly 0; this restatement returns +1 there.    */
/* BLAS calls follow the reference BLAS (netlib dgemm, dspr, dsyrk) order. */

static const double* pt_of(const orc_meca* m, uint32_t p) { return m->vPTS + (size_t)m->dim * p; }

static void pi_pos(const orc_meca* m, uint32_t p1, uint32_t p2, double c, double* out)
{
    const double *x = pt_of(m, p1), *y = pt_of(m, p2);
    for (int d = 0; d < m->dim; ++d) out[d] = x[d] + c * (y[d] - x[d]);
}

static void pi_diff(const orc_meca* m, uint32_t p1, uint32_t p2, double* out)
{
    const double *x = pt_of(m, p1), *y = pt_of(m, p2);
    for (int d = 0; d < m->dim; ++d) out[d] = y[d] - x[d];
}

static double v_dot(int D, const double* a, const double* b)
{
    double s = a[0] * b[0];
    for (int d = 1; d < D; ++d) s = s + a[d] * b[d];
    return s;
}

static double v_norm(int D, const double* a) { return sqrt(v_dot(D, a, a)); }

static void pi_dir(const orc_meca* m, uint32_t p1, uint32_t p2, double* out)
{
    pi_diff(m, p1, p2, out);
    const double sn = 1.0 / v_norm(m->dim, out);
    for (int d = 0; d < m->dim; ++d) out[d] = out[d] * sn;
}

static int sign_exc(double a) { return a < 0 ? -1 : 1; }

/* vecProd in 2D (vector2.h:372-375) */
static double cross2(const double* a, const double* b) { return a[0] * b[1] - a[1] * b[0]; }

/* PointInterpolated::overlapping (point_interpolated.cc:21-33) */
static int overlap_ie(const orc_meca* m, uint32_t a1, uint32_t a2, uint32_t b)
{
    return mecable_of(m, a1) == mecable_of(m, b) && (a1 == b || a2 == b);
}

static int overlap_ii(const orc_meca* m, uint32_t a1, uint32_t a2, uint32_t b1, uint32_t b2)
{
    return mecable_of(m, a1) == mecable_of(m, b1) && (a1 == b1 || a1 == b2 || a2 == b1 || a2 == b2);
}

#define MC(i, j) (*sp_ref(&m->mC, (i), (j)))
#define MB(i, j) (*sp_ref(&m->mB, (i), (j)))

/* addPureForce(PointInterpolated) (meca_inter.cc:140-156) */
void orc_add_pure_force_i(orc_meca* m, uint32_t a1, uint32_t a2, double ca, const double* force)
{
    const int D = m->dim;
    const double a = ca, b = 1.0 - ca;
    for (int d = 0; d < D; ++d) {
        m->vBAS[D * a1 + d] += b * force[d];
        m->vBAS[D * a2 + d] += a * force[d];
    }
}

/* interLongLink(PointInterpolated, PointExact) (meca_inter.cc:425-505) and
   (PointInterpolated, PointInterpolated) (516-596): one body over the n = 3
   or 4 points with coefficients c and the target at the end */
static void long_link_n(orc_meca* m, int n, const uint32_t* pts, const double* c, const double* ab_in,
                        double len, double w)
{
    const int D = m->dim;
    double cw[4], ab[3] = { 0, 0, 0 };
    uint32_t inx[4];
    for (int k = 0; k < n; ++k) { cw[k] = w * c[k]; inx[k] = D * pts[k]; }
    for (int d = 0; d < D; ++d) ab[d] = ab_in[d];
    const double abn = v_norm(D, ab);
    if (abn < REAL_EPSILON) return;
    for (int d = 0; d < D; ++d) ab[d] /= abn;
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < D; ++d) m->vBAS[inx[i] + d] -= cw[i] * ab[d] * len;
    len /= abn;
    const int cooked = (len > 1.0);
    for (int i = 0; i < D; ++i) {
        double v = cooked ? ab[i] * ab[i] : 1.0 + len * (ab[i] * ab[i] - 1.0);
        for (int k = 0; k < n; ++k)
            for (int l = k; l < n; ++l) MC(inx[k] + i, inx[l] + i) -= c[k] * cw[l] * v;
        for (int j = i + 1; j < D; ++j) {
            v = cooked ? ab[i] * ab[j] : len * ab[i] * ab[j];
            for (int k = 0; k < n; ++k)
                for (int l = 0; l < n; ++l) MC(inx[k] + i, inx[l] + j) -= c[k] * cw[l] * v;
        }
    }
}

void orc_inter_long_link_ie(orc_meca* m, uint32_t a1, uint32_t a2, double ca, uint32_t b, double len, double w)
{
    if (overlap_ie(m, a1, a2, b)) return;
    const uint32_t pts[3] = { a1, a2, b };
    const double c[3] = { 1.0 - ca, ca, -1.0 };
    double pa[3], ab[3];
    pi_pos(m, a1, a2, ca, pa);
    for (int d = 0; d < m->dim; ++d) ab[d] = pt_of(m, b)[d] - pa[d];
    long_link_n(m, 3, pts, c, ab, len, w);
}

void orc_inter_long_link_ii(orc_meca* m, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2, double cb,
                            double len, double w)
{
    if (overlap_ii(m, a1, a2, b1, b2)) return;
    const uint32_t pts[4] = { a1, a2, b1, b2 };
    const double c[4] = { 1.0 - ca, ca, -(1.0 - cb), -cb };
    double pa[3], pb[3], ab[3];
    pi_pos(m, a1, a2, ca, pa);
    pi_pos(m, b1, b2, cb, pb);
    for (int d = 0; d < m->dim; ++d) ab[d] = pb[d] - pa[d];
    long_link_n(m, 4, pts, c, ab, len, w);
}

/* the 3D arm of the side links (meca_inter.cc:748-756, 921-929, 1489-1497):
   the part of (target - pta.pos()) orthogonal to the segment, scaled to len;
   returns 0 when that part is shorter than REAL_EPSILON */
static int side_arm3(const double* a, const double* as, double len, double* arm)
{
    const double s = v_dot(3, as, a) / v_dot(3, a, a);
    for (int d = 0; d < 3; ++d) arm[d] = as[d] - s * a[d];
    const double n = v_norm(3, arm);
    if (!(n > REAL_EPSILON)) return 0;
    const double f = len / n;
    for (int d = 0; d < 3; ++d) arm[d] = arm[d] * f;
    return 1;
}

/* interSideLinkS (PE: meca_inter.cc:671-726, PI: 834-886) and
   interSideSlidingLinkS (PE: 1374-1453, PI: 1626-1706) share one body:
   T = a a' + b b' (side link) or b b' (side sliding link), RB = arm or len b */
static void side_s(orc_meca* m, int n, const uint32_t* pts, const double* c, const double* a, const double* arm,
                   double len, double w, int sliding)
{
    double b[3];
    for (int d = 0; d < 3; ++d) b[d] = arm[d] / len;
    double xx, xy, xz, yy, yz, zz;
    if (sliding) {
        xx = b[0] * b[0]; xy = b[0] * b[1]; xz = b[0] * b[2];
        yy = b[1] * b[1]; yz = b[1] * b[2]; zz = b[2] * b[2];
    } else {
        xx = a[0] * a[0] + b[0] * b[0]; xy = a[0] * a[1] + b[0] * b[1]; xz = a[0] * a[2] + b[0] * b[2];
        yy = a[1] * a[1] + b[1] * b[1]; yz = a[1] * a[2] + b[1] * b[2]; zz = a[2] * a[2] + b[2] * b[2];
    }
    const double T[9] = { xx, xy, xz, xy, yy, yz, xz, yz, zz };
    double RB[3];
    for (int d = 0; d < 3; ++d) RB[d] = sliding ? len * b[d] : arm[d];
    double cw[4];
    uint32_t inx[4];
    for (int k = 0; k < n; ++k) { cw[k] = -w * c[k]; inx[k] = 3 * pts[k]; }
    for (int i = 0; i < n; ++i) {
        for (int x = 0; x < 3; ++x) m->vBAS[inx[i] + x] += cw[i] * RB[x];
        const double g = c[i] * cw[i];
        for (int x = 0; x < 3; ++x)
            for (int y = x; y < 3; ++y) MC(inx[i] + x, inx[i] + y) += g * T[x + 3 * y];
        for (int j = i + 1; j < n; ++j) {
            const double h = c[i] * cw[j];
            for (int x = 0; x < 3; ++x)
                for (int y = 0; y < 3; ++y) MC(inx[i] + x, inx[j] + y) += h * T[x + 3 * y];
        }
    }
}

/* interSideLink2D(PointInterpolated, PointExact) (meca_inter.cc:617-667) */
static void side_link2d_ie(orc_meca* m, uint32_t a1, uint32_t a2, double ca, uint32_t b, double arm, double w)
{
    double df[2];
    pi_diff(m, a1, a2, df);
    const double aa = ca, bb = 1.0 - ca;
    const double ee = arm / v_norm(2, df);
    uint32_t i1 = a1, i2 = a2, iE = b;
    const double aaw = aa * w, bbw = bb * w;
    const double eew = w * ee, eeeew = ee * eew;
    MB(i1, i1) -= bbw * bb + eeeew;
    MB(i1, i2) -= aaw * bb - eeeew;
    MB(i2, i2) -= aaw * aa + eeeew;
    MB(iE, iE) -= w;
    MB(i1, iE) += bbw;
    MB(i2, iE) += aaw;
    i1 *= 2; i2 *= 2; iE *= 2;
    MC(i1, i2 + 1) += eew;
    MC(i1 + 1, i2) -= eew;
    MC(i1, iE + 1) -= eew;
    MC(i1 + 1, iE) += eew;
    MC(i2, iE + 1) += eew;
    MC(i2 + 1, iE) -= eew;
}

/* interSideLink(PointInterpolated, PointExact) (meca_inter.cc:730-759) */
void orc_inter_side_link_ie(orc_meca* m, uint32_t a1, uint32_t a2, double ca, uint32_t b, double len, double w)
{
    if (overlap_ie(m, a1, a2, b)) return;
    double pa[3], df[3], as[3];
    pi_pos(m, a1, a2, ca, pa);
    pi_diff(m, a1, a2, df);
    for (int d = 0; d < m->dim; ++d) as[d] = pt_of(m, b)[d] - pa[d];
    if (m->dim == 2) {
        side_link2d_ie(m, a1, a2, ca, b, len * sign_exc(cross2(df, as)), w);
    } else if (m->dim == 3) {
        double arm[3], dir[3];
        if (!side_arm3(df, as, len, arm)) return;
        pi_dir(m, a1, a2, dir);
        const uint32_t pts[3] = { a1, a2, b };
        const double c[3] = { 1.0 - ca, ca, -1.0 };
        side_s(m, 3, pts, c, dir, arm, len, w, 0);
    }
}

/* interSideLink2D(PointInterpolated, PointInterpolated) (meca_inter.cc:764-830) */
static void side_link2d_ii(orc_meca* m, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2, double cb,
                           double arm, double w)
{
    double df[2];
    pi_diff(m, a1, a2, df);
    const double aa1 = ca, bb1 = 1.0 - ca;
    const double aa2 = cb, bb2 = 1.0 - cb;
    const double ee1 = arm / v_norm(2, df);
    uint32_t i11 = a1, i12 = a2, i21 = b1, i22 = b2;
    const double aa1w = aa1 * w, aa2w = aa2 * w;
    const double bb1w = bb1 * w, bb2w = bb2 * w;
    const double ee1w = ee1 * w, ee1ee1w = ee1 * ee1w;
    MB(i11, i11) -= bb1w * bb1 + ee1ee1w;
    MB(i11, i12) += -aa1w * bb1 + ee1ee1w;
    MB(i12, i12) -= aa1w * aa1 + ee1ee1w;
    MB(i21, i21) -= bb2w * bb2;
    MB(i21, i22) += -aa2w * bb2;
    MB(i22, i22) -= aa2w * aa2;
    MB(i11, i21) += bb1w * bb2;
    MB(i11, i22) += bb1w * aa2;
    MB(i12, i22) += aa1w * aa2;
    MB(i12, i21) += aa1w * bb2;
    i11 *= 2; i12 *= 2; i21 *= 2; i22 *= 2;
    MC(i11, i12 + 1) += ee1w;
    MC(i11 + 1, i12) -= ee1w;
    const double ee1bb2w = ee1w * bb2;
    const double ee1aa2w = ee1w * aa2;
    MC(i11, i21 + 1) -= ee1bb2w;
    MC(i11, i22 + 1) -= ee1aa2w;
    MC(i11 + 1, i21) += ee1bb2w;
    MC(i11 + 1, i22) += ee1aa2w;
    MC(i12, i21 + 1) += ee1bb2w;
    MC(i12, i22 + 1) += ee1aa2w;
    MC(i12 + 1, i21) -= ee1bb2w;
    MC(i12 + 1, i22) -= ee1aa2w;
}

/* interSideLink(PointInterpolated, PointInterpolated) (meca_inter.cc:903-935) */
void orc_inter_side_link_ii(orc_meca* m, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2, double cb,
                            double len, double w)
{
    if (overlap_ii(m, a1, a2, b1, b2)) return;
    double pa[3], pb[3], df[3], as[3];
    pi_pos(m, a1, a2, ca, pa);
    pi_pos(m, b1, b2, cb, pb);
    pi_diff(m, a1, a2, df);
    for (int d = 0; d < m->dim; ++d) as[d] = pb[d] - pa[d];
    if (m->dim == 2) {
        side_link2d_ii(m, a1, a2, ca, b1, b2, cb, len * sign_exc(cross2(df, as)), w);
    } else if (m->dim == 3) {
        double arm[3], dir[3];
        if (!side_arm3(df, as, len, arm)) return;
        pi_dir(m, a1, a2, dir);
        const uint32_t pts[4] = { a1, a2, b1, b2 };
        const double c[4] = { 1.0 - ca, ca, -(1.0 - cb), -cb };
        side_s(m, 4, pts, c, dir, arm, len, w, 0);
    }
}

/* interSideSideLink (meca_inter.cc:1038-1059) and interSideSideLink2D
   (943-1021); the reference has no 3D version (it throws): -1 */
int orc_inter_side_side_link(orc_meca* m, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2, double cb,
                             double len, double w)
{
    if (m->dim != 2) return -1;
    double pa[2], pb[2], da[2], db[2], dir[2];
    pi_pos(m, a1, a2, ca, pa);
    pi_pos(m, b1, b2, cb, pb);
    pi_diff(m, a1, a2, da);
    pi_diff(m, b1, b2, db);
    for (int d = 0; d < 2; ++d) dir[d] = pb[d] - pa[d];
    const int side1 = sign_exc(cross2(da, dir));
    const int side2 = sign_exc(cross2(dir, db));
    if (overlap_ii(m, a1, a2, b1, b2)) return 0;
    const double aa1 = ca, bb1 = 1.0 - ca;
    const double aa2 = cb, bb2 = 1.0 - cb;
    const double ee1 = side1 * len / (2 * v_norm(2, da));
    const double ee2 = side2 * len / (2 * v_norm(2, db));
    uint32_t i11 = a1, i12 = a2, i21 = b1, i22 = b2;
    const double aa1w = aa1 * w, aa2w = aa2 * w;
    const double bb1w = bb1 * w, bb2w = bb2 * w;
    const double ee1w = ee1 * w, ee1ee1w = ee1 * ee1w;
    const double ee2w = ee2 * w, ee2ee2w = ee2 * ee2w;
    const double ee1ee2w = ee1 * ee2w;
    MB(i11, i11) -= bb1w * bb1 + ee1ee1w;
    MB(i11, i12) -= aa1w * bb1 - ee1ee1w;
    MB(i12, i12) -= aa1w * aa1 + ee1ee1w;
    MB(i21, i21) -= bb2w * bb2 + ee2ee2w;
    MB(i21, i22) -= aa2w * bb2 - ee2ee2w;
    MB(i22, i22) -= aa2w * aa2 + ee2ee2w;
    MB(i11, i21) += bb1w * bb2 + ee1ee2w;
    MB(i11, i22) += bb1w * aa2 - ee1ee2w;
    MB(i12, i22) += aa1w * aa2 + ee1ee2w;
    MB(i12, i21) += aa1w * bb2 - ee1ee2w;
    i11 *= 2; i12 *= 2; i21 *= 2; i22 *= 2;
    MC(i11, i12 + 1) += ee1w;
    MC(i11 + 1, i12) -= ee1w;
    MC(i21, i22 + 1) += ee2w;
    MC(i21 + 1, i22) -= ee2w;
    const double ee1aa2w = ee1w * aa2;
    const double ee1bb2w = ee1w * bb2;
    const double ee2aa1w = ee2w * aa1;
    const double ee2bb1w = ee2w * bb1;
    MC(i11, i21 + 1) += ee2bb1w - ee1bb2w;
    MC(i11, i22 + 1) -= ee2bb1w + ee1aa2w;
    MC(i11 + 1, i21) += ee1bb2w - ee2bb1w;
    MC(i11 + 1, i22) += ee1aa2w + ee2bb1w;
    MC(i12, i21 + 1) += ee2aa1w + ee1bb2w;
    MC(i12, i22 + 1) += ee1aa2w - ee2aa1w;
    MC(i12 + 1, i21) -= ee1bb2w + ee2aa1w;
    MC(i12 + 1, i22) -= ee1aa2w - ee2aa1w;
    return 0;
}

/* interSlidingLink(PointInterpolated, PointExact) (meca_inter.cc:1075-1140) */
void orc_inter_sliding_link_ie(orc_meca* m, uint32_t a1, uint32_t a2, double ca, uint32_t b, double w)
{
    if (overlap_ie(m, a1, a2, b)) return;
    const int D = m->dim;
    const double aa = ca, bb = 1.0 - ca;
    const double bbbb = bb * bb, aaaa = aa * aa, aabb = aa * bb;
    double dir[3];
    pi_dir(m, a1, a2, dir);
    const uint32_t i1 = D * a1, i2 = D * a2, iE = D * b;
    for (int x = 0; x < D; ++x) {
        for (int y = x; y < D; ++y) {
            const double P = w * ((x == y) - dir[x] * dir[y]);
            MC(i1 + x, i1 + y) -= bbbb * P;
            MC(i1 + x, i2 + y) -= aabb * P;
            MC(i2 + x, i2 + y) -= aaaa * P;
            MC(i1 + x, iE + y) += bb * P;
            MC(i2 + x, iE + y) += aa * P;
            MC(iE + x, iE + y) -= P;
            if (x != y) {
                MC(i1 + y, i2 + x) -= aabb * P;
                MC(i1 + y, iE + x) += bb * P;
                MC(i2 + y, iE + x) += aa * P;
            }
        }
    }
}

/* interSlidingLink(PointInterpolated, PointInterpolated) (meca_inter.cc:1153-1258) */
void orc_inter_sliding_link_ii(orc_meca* m, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2, double cb,
                               double w)
{
    if (overlap_ii(m, a1, a2, b1, b2)) return;
    const int D = m->dim;
    const double c[4] = { 1.0 - ca, ca, -(1.0 - cb), -cb };
    double cw[4];
    for (int k = 0; k < 4; ++k) cw[k] = -w * c[k];
    const uint32_t inx[4] = { D * a1, D * a2, D * b1, D * b2 };
    double dir[3];
    pi_dir(m, a1, a2, dir);
    if (D == 2) {
        const double dd[3] = { dir[0] * dir[0], dir[0] * dir[1], dir[1] * dir[1] };
        for (int j = 0; j < 4; ++j) {
            const double g = c[j] * cw[j];
            MC(inx[j], inx[j]) += (1.0 - dd[0]) * g;
            MC(inx[j] + 1, inx[j] + 1) += (1.0 - dd[2]) * g;
            MC(inx[j], inx[j] + 1) += (-dd[1]) * g;
            for (int i = j + 1; i < 4; ++i) {
                const double h = c[j] * cw[i];
                MC(inx[i], inx[j]) += (1.0 - dd[0]) * h;
                MC(inx[i] + 1, inx[j] + 1) += (1.0 - dd[2]) * h;
                MC(inx[i], inx[j] + 1) += (-dd[1]) * h;
                MC(inx[i] + 1, inx[j]) += (-dd[1]) * h;
            }
        }
    } else if (D == 3) {
        const double dd[6] = { dir[0] * dir[0], dir[0] * dir[1], dir[0] * dir[2],
                               dir[1] * dir[1], dir[1] * dir[2], dir[2] * dir[2] };
        for (int j = 0; j < 4; ++j) {
            const double g = c[j] * cw[j];
            MC(inx[j], inx[j]) += (1.0 - dd[0]) * g;
            MC(inx[j] + 1, inx[j] + 1) += (1.0 - dd[3]) * g;
            MC(inx[j] + 2, inx[j] + 2) += (1.0 - dd[5]) * g;
            MC(inx[j], inx[j] + 1) += (-dd[1]) * g;
            MC(inx[j], inx[j] + 2) += (-dd[2]) * g;
            MC(inx[j] + 1, inx[j] + 2) += (-dd[4]) * g;
            for (int i = j + 1; i < 4; ++i) {
                const double h = c[j] * cw[i];
                MC(inx[i], inx[j]) += (1.0 - dd[0]) * h;
                MC(inx[i] + 1, inx[j] + 1) += (1.0 - dd[3]) * h;
                MC(inx[i] + 2, inx[j] + 2) += (1.0 - dd[5]) * h;
                MC(inx[i], inx[j] + 1) += (-dd[1]) * h;
                MC(inx[i], inx[j] + 2) += (-dd[2]) * h;
                MC(inx[i] + 1, inx[j] + 2) += (-dd[4]) * h;
                MC(inx[i] + 1, inx[j]) += (-dd[1]) * h;
                MC(inx[i] + 2, inx[j]) += (-dd[2]) * h;
                MC(inx[i] + 2, inx[j] + 1) += (-dd[4]) * h;
            }
        }
    }
}

/* netlib dgemm('N','N', 2, n, 2, alpha, P, 2, T, 2, 0, PT, 2) then
   dgemm('T','N', n, n, 2, 1, T, 2, PT, 2, 0, TPT, n): the reference BLAS
   order (C(i,j) accumulated over l from 0, the transposed product as one
   dot of length 2) */
static void tpt2(int n, double alpha, const double* P, const double* T, double* TPT)
{
    double PT[16];
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < 2; ++i) PT[i + 2 * j] = 0.0;
        for (int l = 0; l < 2; ++l) {
            const double t = alpha * T[l + 2 * j];
            for (int i = 0; i < 2; ++i) PT[i + 2 * j] = PT[i + 2 * j] + t * P[i + 2 * l];
        }
    }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double t = 0.0;
            for (int l = 0; l < 2; ++l) t = t + T[l + 2 * i] * PT[l + 2 * j];
            TPT[i + n * j] = 1.0 * t;
        }
}

/* interSideSlidingLink2D (PE: meca_inter.cc:1266-1366, PI: 1508-1618), the
   active '#if (1)' branch: TPT = T' (-w P) T over the n = 6 or 8 coordinates */
static void side_sliding2d(orc_meca* m, int n, const uint32_t* pts, double ca, double cb, double arm, double w)
{
    const uint32_t a1 = pts[0], a2 = pts[1];
    double dir[2], df[2];
    pi_dir(m, a1, a2, dir);
    pi_diff(m, a1, a2, df);
    const double aa = ca, bb = 1.0 - ca;
    const double ee = arm / v_norm(2, df);
    double T[16] = { bb, -ee, ee, bb, aa, ee, -ee, aa, -1, 0, 0, -1, 0, 0, 0, 0 };
    if (n == 8) {
        const double cc = cb, dd = 1.0 - cb;
        const double t2[8] = { -dd, 0, 0, -dd, -cc, 0, 0, -cc };
        for (int k = 0; k < 8; ++k) T[8 + k] = t2[k];
    }
    const double P[4] = { 1 - dir[0] * dir[0], -dir[0] * dir[1], -dir[0] * dir[1], 1 - dir[1] * dir[1] };
    double TPT[64];
    tpt2(n, -w, P, T, TPT);
    uint32_t inx[8];
    for (int k = 0; k < n / 2; ++k) { inx[2 * k] = 2 * pts[k]; inx[2 * k + 1] = 2 * pts[k] + 1; }
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) MC(inx[i], inx[j]) += TPT[i + n * j];
}

/* interSideSlidingLink(PointInterpolated, PointExact) (meca_inter.cc:1469-1499) */
void orc_inter_side_sliding_link_ie(orc_meca* m, uint32_t a1, uint32_t a2, double ca, uint32_t b, double len,
                                    double w)
{
    double pa[3], df[3], as[3];
    pi_pos(m, a1, a2, ca, pa);
    pi_diff(m, a1, a2, df);
    for (int d = 0; d < m->dim; ++d) as[d] = pt_of(m, b)[d] - pa[d];
    if (m->dim == 2) {
        const double arm = len * sign_exc(cross2(df, as));
        if (overlap_ie(m, a1, a2, b)) return;
        const uint32_t pts[3] = { a1, a2, b };
        side_sliding2d(m, 6, pts, ca, 0.0, arm, w);
    } else if (m->dim == 3) {
        double arm[3], dir[3];
        if (!side_arm3(df, as, len, arm)) return;
        if (overlap_ie(m, a1, a2, b)) return;
        pi_dir(m, a1, a2, dir);
        const uint32_t pts[3] = { a1, a2, b };
        const double c[3] = { 1.0 - ca, ca, -1.0 };
        side_s(m, 3, pts, c, dir, arm, len, w, 1);
    }
}

/* interSideSlidingLink(PointInterpolated, PointInterpolated) (meca_inter.cc:1724-1751) */
void orc_inter_side_sliding_link_ii(orc_meca* m, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2,
                                    double cb, double len, double w)
{
    double pa[3], pb[3], df[3], as[3];
    pi_pos(m, a1, a2, ca, pa);
    pi_pos(m, b1, b2, cb, pb);
    pi_diff(m, a1, a2, df);
    for (int d = 0; d < m->dim; ++d) as[d] = pb[d] - pa[d];
    if (m->dim == 2) {
        const double arm = len * sign_exc(cross2(df, as));
        if (overlap_ii(m, a1, a2, b1, b2)) return;
        const uint32_t pts[4] = { a1, a2, b1, b2 };
        side_sliding2d(m, 8, pts, ca, cb, arm, w);
    } else if (m->dim == 3) {
        double arm[3], dir[3];
        if (!side_arm3(df, as, len, arm)) return;
        if (overlap_ii(m, a1, a2, b1, b2)) return;
        pi_dir(m, a1, a2, dir);
        const uint32_t pts[4] = { a1, a2, b1, b2 };
        const double c[4] = { 1.0 - ca, ca, -(1.0 - cb), -cb };
        side_s(m, 4, pts, c, dir, arm, len, w, 1);
    }
}

/* interTorque2D (meca_inter.cc:1772-1869); the reference has it in 2D only: -1 */
int orc_inter_torque(orc_meca* m, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2, double cb,
                     double cosinus, double sinus, double w)
{
    (void)ca; (void)cb;
    if (m->dim != 2) return -1;
    if (overlap_ii(m, a1, a2, b1, b2)) return 0;
    const uint32_t index[8] = { 2 * a1, 2 * a1 + 1, 2 * a2, 2 * a2 + 1, 2 * b1, 2 * b1 + 1, 2 * b2, 2 * b2 + 1 };
    double ab[2], cd[2];
    pi_diff(m, a1, a2, ab);
    pi_diff(m, b1, b2, cd);
    const double *a = pt_of(m, a1), *b = pt_of(m, a2), *c = pt_of(m, b1), *d = pt_of(m, b2);
    const double coord[8] = { a[0], a[1], b[0], b[1], c[0], c[1], d[0], d[1] };
    double ce[2];
    ce[0] = cd[0] * cosinus + cd[1] * sinus;
    ce[1] = -cd[0] * sinus + cd[1] * cosinus;
    const double abn = v_norm(2, ab), abnS = v_dot(2, ab, ab);
    const double cdn = v_norm(2, cd), cdnS = v_dot(2, cd, cd);
    if (abn < REAL_EPSILON || cdn < REAL_EPSILON) return 0;
    for (int k = 0; k < 2; ++k) { ab[k] /= abn; cd[k] /= cdn; ce[k] /= cdn; }
    const double h[8] = { ab[1] / abn, -ab[0] / abn, -ab[1] / abn, ab[0] / abn,
                          -cd[1] / cdn, cd[0] / cdn, cd[1] / cdn, -cd[0] / cdn };
    const double dangle = atan2(ab[0] * ce[1] - ab[1] * ce[0], v_dot(2, ab, ce));
    const double w1 = w;
    const double w2 = w * dangle;
    double mm[36];
    for (int k = 0; k < 36; ++k) mm[k] = 0.0;
    /* dspr('L', 8, -w1, h, 1, mm): lower packed, column by column */
    for (int j = 0, kk = 0; j < 8; kk += 8 - j, ++j) {
        if (h[j] != 0.0) {
            const double t = -w1 * h[j];
            for (int i = j, k = kk; i < 8; ++i, ++k) mm[k] = mm[k] + h[i] * t;
        }
    }
    const double Da = w2 * (-2 * ab[0] * ab[1]) / abnS;
    const double da = w2 * (ab[0] * ab[0] - ab[1] * ab[1]) / abnS;
    const double Dc = w2 * (-2 * cd[0] * cd[1]) / cdnS;
    const double dc = w2 * (cd[0] * cd[0] - cd[1] * cd[1]) / cdnS;
    const double ea[4] = { -Da, -da, Da, da };
    const double ec[4] = { Dc, dc, -Dc, -dc };
    int sa = 0, sc = 26;
    for (int j = 0; j < 4; ++j) {
        const double sg = (j & 1) ? -1.0 : 1.0;
        for (int i = j; i < 4; ++i) {
            mm[i + sa] += sg * ea[i - j];
            mm[i + sc] += sg * ec[i - j];
        }
        sa += 7 - j;
        sc += 3 - j;
    }
    for (int i = 0; i < 8; ++i) {
        m->vBAS[index[i]] += w2 * h[i];
        for (int j = 0; j < 8; ++j) {
            int k;
            if (j < i) {
                k = (int)(j * (7.5 - 0.5 * j) + i);
            } else {
                k = (int)(i * (7.5 - 0.5 * i) + j);
                MC(index[i], index[j]) += mm[k];
            }
            m->vBAS[index[i]] -= mm[k] * coord[j];
        }
    }
    return 0;
}

/* interSideClamp (meca_inter.cc:2200-2226) with interSideClamp2D (2032-2074)
   and interSideClamp3D (2089-2186) */
void orc_inter_side_clamp(orc_meca* m, uint32_t a1, uint32_t a2, double ca, const double* g, double len, double w)
{
    const int D = m->dim;
    double pa[3], df[3];
    pi_pos(m, a1, a2, ca, pa);
    pi_diff(m, a1, a2, df);
    if (D == 2) {
        double gs[2] = { g[0] - pa[0], g[1] - pa[1] };
        const double arm = len * sign_exc(cross2(df, gs));
        const double aa = ca, aaw = w * aa;
        const double bb = 1.0 - ca, bbw = w * bb;
        const double ee = arm / v_norm(2, df);
        const double eew = w * ee;
        const double ee2w = w * ee * ee;
        uint32_t i1 = a1, i2 = a2;
        MB(i1, i1) -= bb * bbw + ee2w;
        MB(i1, i2) -= aa * bbw - ee2w;
        MB(i2, i2) -= aa * aaw + ee2w;
        i1 *= 2; i2 *= 2;
        MC(i1, i2 + 1) += eew;
        MC(i1 + 1, i2) -= eew;
        m->vBAS[i1] += bbw * g[0] - eew * g[1];
        m->vBAS[i1 + 1] += bbw * g[1] + eew * g[0];
        m->vBAS[i2] += aaw * g[0] + eew * g[1];
        m->vBAS[i2 + 1] += aaw * g[1] - eew * g[0];
    } else if (D == 3) {
        const double ag[3] = { pa[0] - g[0], pa[1] - g[1], pa[2] - g[2] };
        double arm[3] = { ag[1] * df[2] - ag[2] * df[1], ag[2] * df[0] - ag[0] * df[2], ag[0] * df[1] - ag[1] * df[0] };
        const double n = v_norm(3, arm);
        if (!(n > REAL_EPSILON)) return;
        const double f = len / n;
        for (int d = 0; d < 3; ++d) arm[d] = arm[d] * f;
        const double aa = 1.0 - ca, bb = ca;
        const double s = 1.0 / v_norm(3, df);
        const double ex = s * arm[0], ey = s * arm[1], ez = s * arm[2];
        const uint32_t i1 = 3 * a1, i2 = 3 * a2;
        const uint32_t inx[6] = { i1, i1 + 1, i1 + 2, i2, i2 + 1, i2 + 2 };
        const double T[18] = { aa, ez, -ey, bb, -ez, ey,
                               -ez, aa, ex, ez, bb, -ex,
                               ey, -ex, aa, -ey, ex, bb };
        const double a2 = aa * aa, b2 = bb * bb, ab = aa * bb;
        const double exx = ex * ex, exy = ex * ey, exz = ex * ez;
        const double eyy = ey * ey, eyz = ey * ez, ezz = ez * ez;
        double TT[36] = { eyy + ezz + a2, 0, 0, 0, 0, 0,
                          -exy, exx + ezz + a2, 0, 0, 0, 0,
                          -exz, -eyz, exx + eyy + a2, 0, 0, 0,
                          -ezz - eyy + ab, ez + exy, exz - ey, eyy + ezz + b2, 0, 0,
                          -ez + exy, -ezz - exx + ab, eyz + ex, -exy, exx + ezz + b2, 0,
                          exz + ey, eyz - ex, -eyy - exx + ab, -exz, -eyz, exx + eyy + b2 };
        const double sca = 1.0 / v_norm(3, arm);
        const double aan = aa * sca, bbn = bb * sca;
        const double TP[6] = { aan * ex, aan * ey, aan * ez, bbn * ex, bbn * ey, bbn * ez };
        /* dsyrk('U','N', 6, 1, w, TP, 6, -w, TT, 6) */
        for (int j = 0; j < 6; ++j) {
            for (int i = 0; i <= j; ++i) TT[i + 6 * j] = -w * TT[i + 6 * j];
            if (TP[j] != 0.0) {
                const double t = w * TP[j];
                for (int i = 0; i <= j; ++i) TT[i + 6 * j] = TT[i + 6 * j] + t * TP[i];
            }
        }
        for (int i = 0; i < 6; ++i)
            for (int j = i; j < 6; ++j) MC(inx[i], inx[j]) += TT[i + 6 * j];
        const double ws = v_dot(3, arm, g) * sca * sca;
        const double gx = w * (g[0] - ws * arm[0]);
        const double gy = w * (g[1] - ws * arm[1]);
        const double gz = w * (g[2] - ws * arm[2]);
        for (int i = 0; i < 6; ++i) m->vBAS[inx[i]] += T[i] * gx + T[i + 6] * gy + T[i + 12] * gz;
    }
}

/* interSlidingClamp (meca_inter.cc:2241-2283) */
void orc_inter_sliding_clamp(orc_meca* m, uint32_t a1, uint32_t a2, double ca, const double* g, double w)
{
    const int D = m->dim;
    const double aa = ca, bb = 1.0 - ca;
    const double bbbb = bb * bb, aaaa = aa * aa, aabb = aa * bb;
    double dir[3];
    pi_dir(m, a1, a2, dir);
    const uint32_t i1 = D * a1, i2 = D * a2;
    for (int x = 0; x < D; ++x) {
        for (int y = x; y < D; ++y) {
            const double P = w * ((x == y) - dir[x] * dir[y]);
            MC(i1 + x, i1 + y) -= bbbb * P;
            MC(i2 + x, i2 + y) -= aaaa * P;
            MC(i1 + x, i2 + y) -= aabb * P;
            if (x != y) MC(i1 + y, i2 + x) -= aabb * P;
        }
    }
    const double pr = v_dot(D, g, dir);
    for (int x = 0; x < D; ++x) {
        m->vBAS[i1 + x] += w * bb * (g[x] - pr * dir[x]);
        m->vBAS[i2 + x] += w * aa * (g[x] - pr * dir[x]);
    }
}

/* interPlane(PointInterpolated) (meca_inter.cc:2340-2378) */
void orc_inter_plane_i(orc_meca* m, uint32_t a1, uint32_t a2, double ca, const double* dir, const double* g,
                       double w)
{
    const int D = m->dim;
    const double aa = ca, bb = 1.0 - ca;
    const double bbbb = bb * bb, aaaa = aa * aa, aabb = aa * bb;
    const uint32_t i1 = D * a1, i2 = D * a2;
    for (int x = 0; x < D; ++x) {
        for (int y = x; y < D; ++y) {
            const double P = w * dir[x] * dir[y];
            MC(i1 + x, i1 + y) -= bbbb * P;
            MC(i2 + x, i2 + y) -= aaaa * P;
            MC(i1 + x, i2 + y) -= aabb * P;
            if (x != y) MC(i1 + y, i2 + x) -= aabb * P;
        }
    }
    const double pr = w * v_dot(D, g, dir);
    for (int x = 0; x < D; ++x) {
        m->vBAS[i1 + x] += pr * bb * dir[x];
        m->vBAS[i2 + x] += pr * aa * dir[x];
    }
}

#undef MC
#undef MB

/* ------------------------------------------------------------------------ */
/* steric interactions: Simul::setStericInteractions (simul_solve.cc:12-188) */
/* with PointGrid (point_grid.cc, point_grid.h) over Grid<DIM> with side     */
/* regions (grid.h:354-370, 600-861).                                        */
/*                                                                           */
/* Items: the segments of the steric fibers ("loci", in fiber order) and the */
/* points of the steric Solids with a radius above REAL_EPSILON ("points"),  */
/* each kept in the list of the cell that holds its centre, in insertion     */
/* order.  Where the reference compares two loci by the address of their     */
/* records (`&bb < &aa`, point_grid.cc:205,295), which depends on the memory */
/* the per-cell arrays were given, this restatement compares their ranks in  */
/* the cell-major order (cell index, then insertion order).                  */

typedef struct {
    uint32_t pt;      /* matIndex of the first point of the segment */
    uint32_t mec;     /* Mecable */
    int first, last;  /* FiberLocus::isFirst / isLast */
    double cut;       /* FiberLocus::len() = segmentation */
    double rad, rng;  /* FatLocus::radius, range */
} st_locus;

typedef struct {
    uint32_t pt;      /* matIndex of the point */
    double rad;
} st_point;

typedef struct {
    int D;
    int gdim[3];
    double inf[3], delta[3];
    uint32_t ncells;
    st_locus* loc; uint32_t nloc;
    st_point* pnt; uint32_t npnt;
    uint32_t* loc_start; uint32_t* loc_list;  /* per cell: [start[c], start[c+1]) in loc_list */
    uint32_t* pnt_start; uint32_t* pnt_list;
    uint32_t* loc_rank;                        /* cell-major rank of each locus */
    double push, pull;
    orc_meca* m;
    orc_steric_call* log; uint32_t nlog, caplog;
} st_grid;

static const double* st_pos(const st_grid* g, uint32_t p) { return g->m->vPTS + (size_t)g->D * p; }

static void st_record(st_grid* g, int kind, uint32_t a1, uint32_t a2, double ca, uint32_t b1, uint32_t b2, double cb,
                      double len, double w)
{
    if (g->log && g->nlog < g->caplog) {
        orc_steric_call* c = &g->log[g->nlog];
        c->kind = kind; c->p[0] = a1; c->p[1] = a2; c->p[2] = b1; c->p[3] = b2;
        c->ca = ca; c->cb = cb; c->len = len; c->w = w;
    }
    ++g->nlog;
}

/* the three Meca calls the steric checks make */
static void st_long_link(st_grid* g, uint32_t a, uint32_t b, double len, double w)
{
    st_record(g, 0, a, b, 0.0, 0, 0, 0.0, len, w);
    orc_inter_long_link_ee(g->m, a, b, len, w);
}

static void st_side_sliding_ie(st_grid* g, uint32_t a1, double ca, uint32_t b, double len, double w)
{
    st_record(g, 1, a1, a1 + 1, ca, b, 0, 0.0, len, w);
    orc_inter_side_sliding_link_ie(g->m, a1, a1 + 1, ca, b, len, w);
}

static void st_side_sliding_ii(st_grid* g, uint32_t a1, double ca, uint32_t b1, double cb, double len, double w)
{
    st_record(g, 2, a1, a1 + 1, ca, b1, b1 + 1, cb, len, w);
    orc_inter_side_sliding_link_ii(g->m, a1, a1 + 1, ca, b1, b1 + 1, cb, len, w);
}

/* Grid::index (grid.h:445-455) with the bounded image (grid.h:147-150) */
static uint32_t st_cell(const st_grid* g, const double* w)
{
    uint32_t inx = 0;
    for (int d = g->D - 1; d >= 0; --d) {
        const double c = 0.0 + (w[d] - g->inf[d]) * g->delta[d];
        const int s = g->gdim[d];
        const int i = c <= 0 ? 0 : (c >= s ? s - 1 : (int)c);
        inx = (uint32_t)g->gdim[d] * inx + (uint32_t)i;
    }
    return inx;
}

/* FiberLocus::projectPoint0 (fiber_locus.cc:25-49): dis is set only when the
   projection falls inside [0, len] */
static void st_project0(const st_grid* g, const st_locus* L, const double* w, double* abs_, double* dis)
{
    const int D = g->D;
    const double *p1 = st_pos(g, L->pt), *p2 = st_pos(g, L->pt + 1);
    double dx[3] = { 0, 0, 0 }, aw[3] = { 0, 0, 0 };
    for (int d = 0; d < D; ++d) { dx[d] = p2[d] - p1[d]; aw[d] = w[d] - p1[d]; }
    const double ls = L->cut;
    *abs_ = v_dot(D, aw, dx) / ls;
    if (0 <= *abs_ && *abs_ <= ls) *dis = v_dot(D, aw, aw) - *abs_ * *abs_;
}

/* checkPP (point_grid.cc:103-118): two spheres (Solid points or fiber ends) */
static void st_check_pp(st_grid* g, uint32_t pa, const double* wa, double ra, uint32_t pb, const double* wb, double rb)
{
    const int D = g->D;
    const double len = ra + rb;
    double vab[3] = { 0, 0, 0 };
    for (int d = 0; d < D; ++d) vab[d] = wb[d] - wa[d];
    if (v_dot(D, vab, vab) < len * len) st_long_link(g, pa, pb, len, g->push);
}

/* checkPL (point_grid.cc:125-172): a sphere against a segment */
static void st_check_pl(st_grid* g, const st_point* aa, const st_locus* bb)
{
    const int D = g->D;
    const double* apos = st_pos(g, aa->pt);
    const double len = aa->rad + bb->rad;
    double a, d = 0.0;
    st_project0(g, bb, apos, &a, &d);
    if (0 <= a) {
        if (a <= bb->cut) {
            if (d < len * len) st_side_sliding_ie(g, bb->pt, a / bb->cut, aa->pt, len, g->push);
        } else if (bb->last) {
            st_check_pp(g, aa->pt, apos, aa->rad, bb->pt + 1, st_pos(g, bb->pt + 1), bb->rad);
        }
    } else {
        if (bb->first) {
            st_check_pp(g, aa->pt, apos, aa->rad, bb->pt, st_pos(g, bb->pt), bb->rad);
        } else {
            /* projection on the previous segment: interact with the node if it falls on its right */
            const double* q1 = st_pos(g, bb->pt);
            const double* q0 = st_pos(g, bb->pt - 1);
            double vab[3] = { 0, 0, 0 }, dp[3] = { 0, 0, 0 };
            for (int k = 0; k < D; ++k) { vab[k] = apos[k] - q1[k]; dp[k] = q1[k] - q0[k]; }
            if (v_dot(D, vab, dp) >= 0) {
                if (v_dot(D, vab, vab) < len * len) st_long_link(g, aa->pt, bb->pt, len, g->push);
            }
        }
    }
}

/* checkLL1 (point_grid.cc:178-242): the first point of bb against segment aa */
static void st_check_ll1(st_grid* g, const st_locus* aa, uint32_t ra, const st_locus* bb, uint32_t rb)
{
    const int D = g->D;
    const double ran = aa->rng + bb->rad;
    double a, d = INFINITY;
    const double* b1 = st_pos(g, bb->pt);
    st_project0(g, aa, b1, &a, &d);
    if (d < ran * ran) {
        const double len = aa->rad + bb->rad;
        st_side_sliding_ie(g, aa->pt, a / aa->cut, bb->pt, len, d > len * len ? g->pull : g->push);
    } else if (a < 0) {
        const double* a1 = st_pos(g, aa->pt);
        if (aa->first) {
            if (rb < ra && bb->first) {
                double vab[3] = { 0, 0, 0 }, db[3] = { 0, 0, 0 };
                const double* b2 = st_pos(g, bb->pt + 1);
                for (int k = 0; k < D; ++k) { vab[k] = b1[k] - a1[k]; db[k] = b2[k] - b1[k]; }
                const double len = aa->rad + bb->rad;
                if (v_dot(D, vab, vab) < len * len && v_dot(D, vab, db) >= 0)
                    st_long_link(g, aa->pt, bb->pt, len, g->push);
            }
        } else {
            double vab[3] = { 0, 0, 0 }, dp[3] = { 0, 0, 0 };
            const double* a0 = st_pos(g, aa->pt - 1);
            for (int k = 0; k < D; ++k) { vab[k] = b1[k] - a1[k]; dp[k] = a1[k] - a0[k]; }
            if (v_dot(D, vab, dp) >= 0) {
                const double dd = v_dot(D, vab, vab);
                if (dd < ran * ran) {
                    const double len = aa->rad + bb->rad;
                    st_long_link(g, aa->pt, bb->pt, len, dd > len * len ? g->pull : g->push);
                }
            }
        }
    }
}

/* checkLL2 (point_grid.cc:248-315): the last point of bb against segment aa */
static void st_check_ll2(st_grid* g, const st_locus* aa, uint32_t ra, const st_locus* bb, uint32_t rb)
{
    const int D = g->D;
    const double ran = aa->rng + bb->rad;
    double a, d = INFINITY;
    const double* b2 = st_pos(g, bb->pt + 1);
    st_project0(g, aa, b2, &a, &d);
    if (d < ran * ran) {
        const double len = aa->rad + bb->rad;
        st_side_sliding_ie(g, aa->pt, a / aa->cut, bb->pt + 1, len, d > len * len ? g->pull : g->push);
    } else if (a < 0) {
        const double* a1 = st_pos(g, aa->pt);
        double vab[3] = { 0, 0, 0 };
        for (int k = 0; k < D; ++k) vab[k] = b2[k] - a1[k];
        if (aa->first) {
            double db[3] = { 0, 0, 0 };
            const double* b1 = st_pos(g, bb->pt);
            for (int k = 0; k < D; ++k) db[k] = b2[k] - b1[k];
            const double len = aa->rad + bb->rad;
            if (v_dot(D, vab, vab) < len * len && v_dot(D, vab, db) <= 0)
                st_long_link(g, aa->pt, bb->pt + 1, len, g->push);
        } else {
            double dp[3] = { 0, 0, 0 };
            const double* a0 = st_pos(g, aa->pt - 1);
            for (int k = 0; k < D; ++k) dp[k] = a1[k] - a0[k];
            if (v_dot(D, vab, dp) >= 0) {
                const double dd = v_dot(D, vab, vab);
                if (dd < ran * ran) {
                    const double len = aa->rad + bb->rad;
                    st_long_link(g, aa->pt, bb->pt + 1, len, dd > len * len ? g->pull : g->push);
                }
            }
        }
    } else if (rb < ra && aa->last && a > aa->cut) {
        const double* a2 = st_pos(g, aa->pt + 1);
        const double* b1 = st_pos(g, bb->pt);
        double vab[3] = { 0, 0, 0 }, db[3] = { 0, 0, 0 };
        for (int k = 0; k < D; ++k) { vab[k] = b2[k] - a2[k]; db[k] = b2[k] - b1[k]; }
        const double len = aa->rad + bb->rad;
        if (v_dot(D, vab, vab) < len * len && v_dot(D, vab, db) <= 0)
            st_long_link(g, aa->pt + 1, bb->pt + 1, len, g->push);
    }
}

/* FiberLocus::projectSegment (fiber_locus.cc:206-270) */
static int st_project_segment(const st_grid* g, const st_locus* A, const st_locus* B, double* abs1, double* abs2,
                              double* dist)
{
    const int D = g->D;
    const double *a1 = st_pos(g, A->pt), *a2 = st_pos(g, A->pt + 1);
    const double *b1 = st_pos(g, B->pt), *b2 = st_pos(g, B->pt + 1);
    double d1[3] = { 0, 0, 0 }, d2[3] = { 0, 0, 0 }, d12[3] = { 0, 0, 0 };
    for (int k = 0; k < D; ++k) { d1[k] = a2[k] - a1[k]; d2[k] = b2[k] - b1[k]; d12[k] = b1[k] - a1[k]; }
    const double len1 = A->cut, len2 = B->cut;
    const double beta = v_dot(D, d1, d2) / (len1 * len2);
    const double scal = 1.0 - beta * beta;
    if (scal > REAL_EPSILON) {
        const double d1d12 = v_dot(D, d1, d12) / (scal * len1);
        const double d2d12 = v_dot(D, d2, d12) / (scal * len2);
        *abs1 = d1d12 - beta * d2d12;
        if (*abs1 < 0 || len1 < *abs1) return 0;
        *abs2 = beta * d1d12 - d2d12;
        if (*abs2 < 0 || len2 < *abs2) return 0;
        double v[3] = { 0, 0, 0 };
        const double s = *abs2 / len2;
        for (int k = 0; k < D; ++k) v[k] = d12[k] + d2[k] * s;
        *dist = v_dot(D, v, v) - *abs1 * *abs1;
        return 1;
    }
    double p1 = v_dot(D, d12, d1) / len1;
    double p2 = p1 + beta * len2;
    if (p1 < 0 && p2 < 0) return 0;
    if (p1 > len1 && p2 > len1) return 0;
    const double p = p1;
    *dist = v_dot(D, d12, d12) - p1 * p1;
    if (p1 < 0) p1 = 0; else if (p1 > len1) p1 = len1;
    if (p2 < 0) p2 = 0; else if (p2 > len1) p2 = len1;
    *abs1 = 0.5 * (p1 + p2);
    *abs2 = (*abs1 - p) * beta;
    return 1;
}

/* checkLL (point_grid.cc:321-354) */
static void st_check_ll(st_grid* g, uint32_t ia, uint32_t ib)
{
    const st_locus *aa = &g->loc[ia], *bb = &g->loc[ib];
    const uint32_t ra = g->loc_rank[ia], rb = g->loc_rank[ib];
    st_check_ll1(g, aa, ra, bb, rb);
    if (aa->last) st_check_ll2(g, bb, rb, aa, ra);
    st_check_ll1(g, bb, rb, aa, ra);
    if (bb->last) st_check_ll2(g, aa, ra, bb, rb);
    if (g->D == 3) {
        const double ran = fmax(aa->rng + bb->rad, aa->rad + bb->rng);
        double a, b, d;
        if (st_project_segment(g, aa, bb, &a, &b, &d) && d < ran * ran) {
            const double len = aa->rad + bb->rad;
            st_side_sliding_ii(g, aa->pt, a / aa->cut, bb->pt, b / bb->cut, len, d > len * len ? g->pull : g->push);
        }
    }
}

static void st_check_pp_items(st_grid* g, uint32_t i, uint32_t j)
{
    const st_point *a = &g->pnt[i], *b = &g->pnt[j];
    st_check_pp(g, a->pt, st_pos(g, a->pt), a->rad, b->pt, st_pos(g, b->pt), b->rad);
}

/* Grid::createSideRegions(1) for one cell (grid.h:650-861): the shifts of
   makeRectangularGrid in its order, kept when inside and at an index offset >= 0 */
static int st_region(const st_grid* g, uint32_t indx, int* off)
{
    const int D = g->D;
    int ori[3] = { 0, 0, 0 };
    uint32_t r = indx;
    for (int d = 0; d < D; ++d) { ori[d] = (int)(r % (uint32_t)g->gdim[d]); r /= (uint32_t)g->gdim[d]; }
    int shifts[27][3];
    int nb = 1;
    for (int k = 0; k < 3; ++k) shifts[0][k] = 0;
    for (int d = 0; d < D; ++d) {
        int h = nb;
        for (int s = -1; s <= 1; s += 2)
            for (int n = 0; n < nb; ++n, ++h) {
                for (int e = 0; e < 3; ++e) shifts[h][e] = shifts[n][e];
                shifts[h][d] = s;
            }
        nb = h;
    }
    int cnt = 0;
    for (int q = 0; q < nb; ++q) {
        int cc[3], in = 1;
        for (int d = 0; d < D; ++d) {
            cc[d] = ori[d] + shifts[q][d];
            if (cc[d] < 0 || cc[d] >= g->gdim[d]) in = 0;
        }
        if (!in) continue;
        int idx = 0;
        for (int d = D - 1; d >= 0; --d) idx = g->gdim[d] * idx + cc[d];
        const int o = idx - (int)indx;
        if (o >= 0) off[cnt++] = o;
    }
    return cnt;
}

/* PointGrid::setInteractions (point_grid.cc:361-413) */
static void st_traverse(st_grid* g)
{
    for (uint32_t indx = 0; indx < g->ncells; ++indx) {
        int reg[27];
        const int nr = st_region(g, indx, reg);
        const uint32_t p0 = g->pnt_start[indx], p1 = g->pnt_start[indx + 1];
        const uint32_t l0 = g->loc_start[indx], l1 = g->loc_start[indx + 1];
        for (uint32_t ii = p0; ii < p1; ++ii) {
            for (uint32_t jj = ii + 1; jj < p1; ++jj) st_check_pp_items(g, g->pnt_list[ii], g->pnt_list[jj]);
            for (uint32_t jj = l0; jj < l1; ++jj) st_check_pl(g, &g->pnt[g->pnt_list[ii]], &g->loc[g->loc_list[jj]]);
        }
        for (uint32_t ii = l0; ii < l1; ++ii)
            for (uint32_t jj = ii + 1; jj < l1; ++jj) st_check_ll(g, g->loc_list[ii], g->loc_list[jj]);
        for (int r = 1; r < nr; ++r) {
            const uint32_t side = indx + (uint32_t)reg[r];
            const uint32_t q0 = g->pnt_start[side], q1 = g->pnt_start[side + 1];
            const uint32_t s0 = g->loc_start[side], s1 = g->loc_start[side + 1];
            for (uint32_t ii = p0; ii < p1; ++ii) {
                for (uint32_t jj = q0; jj < q1; ++jj) st_check_pp_items(g, g->pnt_list[ii], g->pnt_list[jj]);
                for (uint32_t jj = s0; jj < s1; ++jj) st_check_pl(g, &g->pnt[g->pnt_list[ii]], &g->loc[g->loc_list[jj]]);
            }
            for (uint32_t ii = l0; ii < l1; ++ii) {
                for (uint32_t jj = q0; jj < q1; ++jj) st_check_pl(g, &g->pnt[g->pnt_list[jj]], &g->loc[g->loc_list[ii]]);
                for (uint32_t jj = s0; jj < s1; ++jj) st_check_ll(g, g->loc_list[ii], g->loc_list[jj]);
            }
        }
    }
}

/* counting sort of items into cells, insertion order kept */
static void st_bucket(uint32_t ncells, uint32_t n, const uint32_t* cell, uint32_t** start, uint32_t** list)
{
    *start = (uint32_t*)calloc((size_t)ncells + 1, sizeof(uint32_t));
    *list = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) (*start)[cell[i] + 1]++;
    for (uint32_t c = 0; c < ncells; ++c) (*start)[c + 1] += (*start)[c];
    uint32_t* fill = (uint32_t*)malloc(((size_t)ncells + 1) * sizeof(uint32_t));
    memcpy(fill, *start, ((size_t)ncells + 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) (*list)[fill[cell[i]]++] = i;
    free(fill);
}

int orc_steric(orc_meca* m, const orc_steric_params* sp, const int* steric, const double* radius, const double* range,
               int brute_force, orc_steric_call* log, uint32_t caplog, uint32_t* ncalls)
{
    const int D = m->dim;
    st_grid g;
    memset(&g, 0, sizeof(g));
    g.D = D; g.m = m; g.push = sp->stiff_push; g.pull = sp->stiff_pull;
    g.log = log; g.caplog = caplog;
    /* Simul::estimateStericRange (simul_solve.cc:12-82) when no range is given,
       with the fibers' actual segmentation for FiberProp::segmentation */
    double step = sp->max_range;
    if (step <= 0) {
        double ran = 0, len = 0;
        for (int f = 0; f < m->nfib; ++f) {
            const orc_fiber* F = &m->fib[f];
            if (F->kind != 0 || !steric[f]) continue;
            if (len < 4.1 / 3 * F->cut) len = 4.1 / 3 * F->cut;
            if (ran < radius[f] + range[f]) ran = radius[f] + range[f];
        }
        ran = len + 2 * ran;
        for (int f = 0; f < m->nfib; ++f) {
            const orc_fiber* F = &m->fib[f];
            if (F->kind != 1 || !steric[f]) continue;
            for (uint32_t p = 0; p < F->np; ++p)
                if (ran < 2 * F->radius[p]) ran = 2 * F->radius[p];
        }
        step = ran;
    }
    if (!(step > REAL_EPSILON)) return -1;
    /* PointGrid::setGrid (point_grid.cc:25-64) with Grid::setDimensions (grid.h:281-307) */
    uint32_t ncells = 1;
    double width[3] = { 0, 0, 0 };
    for (int d = 0; d < D; ++d) {
        double ext = sp->extension[d];
        if (ext < 0) return -2;
        const double n = 2 * ext / step;
        g.gdim[d] = (int)ceil(n) + 2;
        ext = g.gdim[d] * 0.5 * step;
        g.inf[d] = -ext;
        const double sup = ext;
        width[d] = (sup - g.inf[d]) / (double)g.gdim[d];
        g.delta[d] = (double)g.gdim[d] / (sup - g.inf[d]);
        ncells *= (uint32_t)g.gdim[d];
    }
    g.ncells = ncells;
    /* Grid::minimumWidth(1) (grid.h:354-364) */
    double maxd = 0;
    for (int d = 0; d < D; ++d) if (g.gdim[d] > 1) maxd = width[d];
    for (int d = 0; d < D; ++d) if (g.gdim[d] > 1 && width[d] < maxd) maxd = width[d];
    /* items (simul_solve.cc:153-184), PointGrid::add (point_grid.cc:71-96) */
    uint32_t nl = 0, npn = 0;
    for (int f = 0; f < m->nfib; ++f) {
        const orc_fiber* F = &m->fib[f];
        if (!steric[f]) continue;
        if (F->kind == 0) nl += F->np - 1;
        else for (uint32_t p = 0; p < F->np; ++p) npn += F->radius[p] > REAL_EPSILON;
    }
    g.loc = (st_locus*)malloc((nl ? nl : 1) * sizeof(st_locus));
    g.pnt = (st_point*)malloc((npn ? npn : 1) * sizeof(st_point));
    uint32_t* lcell = (uint32_t*)malloc((nl ? nl : 1) * sizeof(uint32_t));
    uint32_t* pcell = (uint32_t*)malloc((npn ? npn : 1) * sizeof(uint32_t));
    int err = 0;
    for (int f = 0; f < m->nfib; ++f) {
        const orc_fiber* F = &m->fib[f];
        if (!steric[f]) continue;
        if (F->kind == 0) {
            for (uint32_t r = 0; r + 1 < F->np; ++r) {
                st_locus* L = &g.loc[g.nloc];
                L->pt = F->idx + r; L->mec = (uint32_t)f;
                L->first = r == 0; L->last = r + 2 == F->np;
                L->cut = F->cut; L->rad = radius[f]; L->rng = radius[f] + range[f];
                const double *p1 = st_pos(&g, L->pt), *p2 = st_pos(&g, L->pt + 1);
                double c[3] = { 0, 0, 0 };
                for (int d = 0; d < D; ++d) c[d] = 0.5 * (p1[d] + p2[d]);
                lcell[g.nloc] = st_cell(&g, c);
                if (maxd < sqrt(L->cut * L->cut + L->rng * L->rng)) err = -3;
                ++g.nloc;
            }
        } else {
            for (uint32_t p = 0; p < F->np; ++p) {
                if (!(F->radius[p] > REAL_EPSILON)) continue;
                st_point* P = &g.pnt[g.npnt];
                P->pt = F->idx + p; P->rad = F->radius[p];
                pcell[g.npnt] = st_cell(&g, st_pos(&g, P->pt));
                if (maxd < 2 * P->rad) err = -3;
                ++g.npnt;
            }
        }
    }
    if (!err) {
        st_bucket(ncells, g.nloc, lcell, &g.loc_start, &g.loc_list);
        st_bucket(ncells, g.npnt, pcell, &g.pnt_start, &g.pnt_list);
        g.loc_rank = (uint32_t*)malloc((g.nloc ? g.nloc : 1) * sizeof(uint32_t));
        for (uint32_t k = 0; k < g.nloc; ++k) g.loc_rank[g.loc_list[k]] = k;
        if (!brute_force) {
            st_traverse(&g);
        } else {
            /* every pair once, in the cell-major order of the items: the set of
               calls the grid must find when its cells are wide enough */
            for (uint32_t i = 0; i < g.npnt; ++i)
                for (uint32_t j = i + 1; j < g.npnt; ++j) st_check_pp_items(&g, g.pnt_list[i], g.pnt_list[j]);
            for (uint32_t i = 0; i < g.npnt; ++i)
                for (uint32_t j = 0; j < g.nloc; ++j) st_check_pl(&g, &g.pnt[g.pnt_list[i]], &g.loc[g.loc_list[j]]);
            for (uint32_t i = 0; i < g.nloc; ++i)
                for (uint32_t j = i + 1; j < g.nloc; ++j) st_check_ll(&g, g.loc_list[i], g.loc_list[j]);
        }
        free(g.loc_start); free(g.loc_list); free(g.pnt_start); free(g.pnt_list); free(g.loc_rank);
    }
    free(lcell); free(pcell); free(g.loc); free(g.pnt);
    if (ncalls) *ncalls = g.nlog;
    return err;
}

/* Fiber::setInteractions, CONFINE_INSIDE (fiber.cc:471-481), with
   SpaceSphere::inside (space_sphere.cc:113-116) and setInteraction (148-151) */
void orc_confine_sphere(orc_meca* m, double radius, double stiffness)
{
    const int D = m->dim;
    const double zero[3] = { 0, 0, 0 };
    for (int i = 0; i < m->nfib; ++i) {
        const orc_fiber* f = &m->fib[i];
        if (f->kind != 0) continue;
        for (uint32_t p = 0; p < f->np; ++p) {
            const double* x = m->pos + D * (f->idx + p);
            double r2 = 0.0;
            for (int d = 0; d < D; ++d) r2 += x[d] * x[d];
            if (!(r2 <= radius * radius))
                orc_inter_long_clamp(m, f->idx + p, zero, radius, stiffness);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* per-fiber operators                                                       */

/* projectForcesA / dptts2 / projectForcesB (rigid_fiber_project.cc:95-149):
   Y = sca * P X  with  P = I - J'(JJ')^-1 J */
static void fiber_project(const orc_meca* m, const orc_fiber* f, const double* X, double* Y,
                          double sca, double* tmp)
{
    const int D = m->dim;
    const uint32_t S = f->np - 1;
    const double* dif = m->diff + D * f->seg;
    for (uint32_t s = 0; s < S; ++s) {
        double t = 0.0;
        for (int d = 0; d < D; ++d) t += dif[D * s + d] * (X[D * s + D + d] - X[D * s + d]);
        tmp[s] = t;
    }
    orc_dptts2((int)S, m->jd + f->seg, m->je + f->seg, tmp);
    for (int d = 0; d < D; ++d) {
        const uint32_t kk = D * S;
        Y[d] = sca * (X[d] + dif[d] * tmp[0]);
        Y[kk + d] = sca * (X[kk + d] - dif[kk - D + d] * tmp[S - 1]);
    }
    for (uint32_t s = 1; s < S; ++s) {
        const uint32_t kk = D * s;
        for (int d = 0; d < D; ++d)
            Y[kk + d] = sca * (X[kk + d] + dif[kk + d] * tmp[s] - dif[kk - D + d] * tmp[s - 1]);
    }
}

/* Solid::setSpeedsFromForces (2D: solid.cc:763-782, 3D: 858-888): the
   total force T and torque about the origin R, reduced to rigid motion */
static void solid_speeds(const orc_meca* m, const orc_fiber* f, const double* X, double* Y, double sc)
{
    const double* P = m->vPTS + m->dim * f->idx;   /* psPos: the points put into vPTS by prepare */
    const double* c = f->center;
    if (m->dim == 2) {
        double TX = 0, TY = 0, R = 0;
        for (uint32_t p = 0; p < f->np; ++p) {
            TX += X[2 * p];
            TY += X[2 * p + 1];
            R += P[2 * p] * X[2 * p + 1] - P[2 * p + 1] * X[2 * p];
        }
        R = (sc / f->mom2d) * (R + (TX * c[1] - TY * c[0]));
        const double Tx = (sc / f->drag) * TX + c[1] * R;
        const double Ty = (sc / f->drag) * TY - c[0] * R;
        for (uint32_t p = 0; p < f->np; ++p) {
            Y[2 * p] = Tx - R * P[2 * p + 1];
            Y[2 * p + 1] = Ty + R * P[2 * p];
        }
        return;
    }
    double TX = 0, TY = 0, TZ = 0, RX = 0, RY = 0, RZ = 0;
    for (uint32_t p = 0; p < f->np; ++p) {
        const double* x = X + 3 * p;
        const double* q = P + 3 * p;
        TX += x[0];
        TY += x[1];
        TZ += x[2];
        RX += q[1] * x[2] - q[2] * x[1];
        RY += q[2] * x[0] - q[0] * x[2];
        RZ += q[0] * x[1] - q[1] * x[0];
    }
    double R[3] = { sc * (RX + (TY * c[2] - TZ * c[1])), sc * (RY + (TZ * c[0] - TX * c[2])),
                    sc * (RZ + (TX * c[1] - TY * c[0])) };
    potrs_upper3(f->mom, R);
    const double s = sc / f->drag;
    const double T[3] = { s * TX + (c[1] * R[2] - c[2] * R[1]), s * TY + (c[2] * R[0] - c[0] * R[2]),
                          s * TZ + (c[0] * R[1] - c[1] * R[0]) };
    for (uint32_t p = 0; p < f->np; ++p) {
        const double* q = P + 3 * p;
        Y[3 * p] = T[0] + R[1] * q[2] - R[2] * q[1];
        Y[3 * p + 1] = T[1] + R[2] * q[0] - R[0] * q[2];
        Y[3 * p + 2] = T[2] + R[0] * q[1] - R[1] * q[0];
    }
}

/* RigidFiber::setSpeedsFromForces (rigid_fiber.cc:124-132) */
static void fiber_speeds(const orc_meca* m, const orc_fiber* f, const double* X, double* Y,
                         double sc, int rhs)
{
    if (f->kind == 1) {
        solid_speeds(m, f, X, Y, sc);
        return;
    }
    double* tmp = (rhs ? m->lag : m->ltmp) + f->seg;
    fiber_project(m, f, X, Y, sc * f->mobility, tmp);
}

/* RigidFiber::addRigidity with add_rigidity1 (rigid_fiber.cc:220-229, 294-303) */
static void fiber_add_rigidity(const orc_meca* m, const orc_fiber* f, const double* X, double* Y)
{
    const int D = m->dim;
    if (f->kind != 0 || f->np <= 2) return;
    const uint32_t nbs = D * (f->np - 2);
    const double k = f->rf_rigidity;
    for (uint32_t j = 0; j < nbs; ++j) {
        double v = k * (X[j] - 2 * X[j + D] + X[j + 2 * D]);
        Y[j] -= v;
        Y[j + D] += v + v;
        Y[j + 2 * D] -= v;
    }
}

/* RigidFiber::addRigidityMatUp (rigid_fiber.cc:191-215) into an upper
   MatrixSymmetric of size ps (val[min + ps*max], matsym.cc:51-59) */
static void fiber_rigidity_mat_up(const orc_fiber* f, double* M, int ps)
{
    int sz = (int)f->np;
    if (f->kind != 0 || sz < 3) return;
    const double r = f->rf_rigidity;
#define UP(i, j) M[((i) < (j) ? (i) + ps * (j) : (j) + ps * (i))]
    int s = 0, e = sz;
    for (int i = s + 2; i < e - 2; ++i) UP(i, i) -= 6 * r;
    for (int i = s + 1; i < e - 2; ++i) UP(i, i + 1) += 4 * r;
    for (int i = s; i < e - 2; ++i) UP(i, i + 2) -= r;
    UP(s, s) -= r;
    UP(e - 1, e - 1) -= r;
    if (sz == 3) {
        UP(s + 1, s + 1) -= 4 * r;
    } else {
        UP(s + 1, s + 1) -= 5 * r;
        UP(e - 2, e - 2) -= 5 * r;
    }
    UP(s, s + 1) += 2 * r;
    UP(e - 2, e - 1) += 2 * r;
#undef UP
}

/* makeProjectionDiff (rigid_fiber_project.cc:168-219): compressive
   multipliers are dropped */
static void fiber_make_projection_diff(orc_meca* m, const orc_fiber* f)
{
    if (f->kind != 0) return;
    const double sc = 1.0 / f->cut;
    const uint32_t S = f->np - 1;
    for (uint32_t s = 0; s < S; ++s) {
        double l = m->lag[f->seg + s];
        m->pdiff[f->seg + s] = (l < 0) ? 0.0 : l * sc;
    }
}

/* add_projection / addProjectionDiff (rigid_fiber_project.cc:225-248) */
static void fiber_add_projection_diff(const orc_meca* m, const orc_fiber* f, const double* X, double* Y)
{
    const int D = m->dim;
    if (f->kind != 0) return;
    const uint32_t S = f->np - 1;
    const double* c = m->pdiff + f->seg;
    for (uint32_t s = 0; s < S; ++s) {
        if (c[s] != 0.0) {
            const uint32_t l = D * s, k = l + D;
            for (int d = 0; d < D; ++d) {
                double v = c[s] * (X[k + d] - X[l + d]);
                Y[l + d] += v;
                Y[k + d] -= v;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Meca operators                                                            */

/* addLinearForces (meca.cc:99-127) */
static void add_linear_forces(const orc_meca* m, const double* X, double* Y, int with_rigidity)
{
    const int D = m->dim;
    if (with_rigidity)
        for (int i = 0; i < m->nfib; ++i) {
            const orc_fiber* f = &m->fib[i];
            fiber_add_rigidity(m, f, X + D * f->idx, Y + D * f->idx);
        }
    if (m->use_mB) sp_mul_add_iso(&m->mB, D, X, Y);
    if (m->use_mC) sp_mul_add(&m->mC, X, Y);
}

/* computeForces(X, Y, rig) (meca.cc:143-147) */
void orc_forces_linear(const orc_meca* m, const double* X, double* Y, int with_rigidity)
{
    memcpy(Y, m->vBAS, sizeof(double) * m->dim * m->npts);
    add_linear_forces(m, X, Y, with_rigidity);
}

/* Meca::multiply (meca.cc:157-181): Y = X - dt * P (mB + mC + R + P') X */
void orc_multiply(const orc_meca* m, const double* X, double* Y)
{
    const int D = m->dim;
    const uint32_t n = D * m->npts;
    double* T = m->vTMP;
    memset(T, 0, sizeof(double) * n);
    add_linear_forces(m, X, T, 1);
    for (int i = 0; i < m->nfib; ++i) {
        const orc_fiber* f = &m->fib[i];
        const uint32_t o = D * f->idx;
        fiber_add_projection_diff(m, f, X + o, T + o);
        fiber_speeds(m, f, T + o, Y + o, -m->time_step, 0);
    }
    for (uint32_t k = 0; k < n; ++k) Y[k] += X[k];
}

/* Meca::duplicateMat (meca.cc:188-223) */
static void duplicate_mat(int D, int ps, const double* X, double* Y)
{
    const int bs = D * ps;
    memset(Y, 0, sizeof(double) * bs * bs);
    for (int i = 0; i < ps; ++i) {
        double xx = X[i + ps * i];
        int kk = (bs + 1) * D * i;
        for (int d = 0; d < D; ++d, kk += bs + 1) Y[kk] = xx;
        for (int j = i + 1; j < ps; ++j) {
            xx = X[i + ps * j];
            kk = D * (i + bs * j);
            int ll = D * (j + bs * i);
            for (int d = 0; d < D; ++d, kk += bs + 1, ll += bs + 1) {
                Y[kk] = xx;
                Y[ll] = xx;
            }
        }
    }
}

/* Meca::getBlock (meca.cc:231-283): I - dt P (B + C + R + P'), column-major */
static void get_block(const orc_meca* m, const orc_fiber* f, double* blk)
{
    const int D = m->dim;
    const int ps = (int)f->np, bs = D * ps;
    size_t n1 = (size_t)ps * ps > (size_t)bs ? (size_t)ps * ps : (size_t)bs;
    double* t1 = (double*)calloc(n1, sizeof(double));
    double* t2 = (double*)calloc((size_t)bs * bs, sizeof(double));
    fiber_rigidity_mat_up(f, t1, ps);
    if (m->use_mB) sp_add_triangular_block(&m->mB, t1, f->idx, ps);
    duplicate_mat(D, ps, t1, t2);
    if (m->use_mC) sp_add_diagonal_block(&m->mC, t2, D * f->idx, bs);
    memset(t1, 0, sizeof(double) * bs);
    for (int i = 0; i < bs; ++i) {
        t1[i] = 1;
        fiber_add_projection_diff(m, f, t1, t2 + bs * i);
        t1[i] = 0;
    }
    for (int i = 0; i < bs; ++i) {
        fiber_speeds(m, f, t2 + (size_t)bs * i, blk + (size_t)bs * i, -m->time_step, 0);
        blk[i + (size_t)bs * i] += 1.0;
    }
    free(t1);
    free(t2);
}

void orc_get_block(const orc_meca* m, int i, double* blk) { get_block(m, &m->fib[i], blk); }

int orc_get_inverse_block(const orc_meca* m, int i, double* blk)
{
    const orc_fiber* f = &m->fib[i];
    const uint32_t bs = m->dim * f->np;
    if (!f->use_blk || !f->blk) return 0;
    memcpy(blk, f->blk, sizeof(double) * bs * bs);
    return 1;
}

/* Meca::computePreconditionner (meca.cc:323-444) */
int orc_compute_preconditioner(orc_meca* m)
{
    const int D = m->dim;
    int* ipiv = (int*)malloc(sizeof(int) * D * (m->largest ? m->largest : 1));
    for (int i = 0; i < m->nfib; ++i) {
        orc_fiber* f = &m->fib[i];
        const uint32_t bs = D * f->np;
        if (f->blk_alloc < bs * bs) {
            free(f->blk);
            f->blk = (double*)malloc(sizeof(double) * bs * bs);
            f->blk_alloc = bs * bs;
        }
        get_block(m, f, f->blk);
        int res = 0;
        if (orc_dgetrf((int)bs, f->blk, (int)bs, ipiv)) res = 2;
        else if (orc_dgetri((int)bs, f->blk, (int)bs, ipiv)) res = 3;
        f->use_blk = (res == 0);
    }
    free(ipiv);
    return 0;
}

/* Meca::precondition (meca.cc:448-466): per-block dgemv, or copy */
void orc_precondition(const orc_meca* m, const double* X, double* Y)
{
    const int D = m->dim;
    for (int i = 0; i < m->nfib; ++i) {
        const orc_fiber* f = &m->fib[i];
        const uint32_t bs = D * f->np, o = D * f->idx;
        if (f->use_blk) {
            for (uint32_t r = 0; r < bs; ++r) Y[o + r] = 0.0;
            for (uint32_t c = 0; c < bs; ++c) {
                double x = X[o + c];
                const double* col = f->blk + (size_t)bs * c;
                for (uint32_t r = 0; r < bs; ++r) Y[o + r] += col[r] * x;
            }
        } else {
            memcpy(Y + o, X + o, sizeof(double) * bs);
        }
    }
}

void orc_project(const orc_meca* m, int i, const double* X, double* Y, double sc)
{
    const orc_fiber* f = &m->fib[i];
    fiber_project(m, f, X, Y, sc, m->ltmp + f->seg);
}

void orc_add_rigidity(const orc_meca* m, int i, const double* X, double* Y)
{
    fiber_add_rigidity(m, &m->fib[i], X, Y);
}

void orc_add_projection_diff(const orc_meca* m, int i, const double* X, double* Y)
{
    fiber_add_projection_diff(m, &m->fib[i], X, Y);
}

/* ------------------------------------------------------------------------ */
/* solver (bicgstab.{h,cc})                                                  */

typedef struct {
    int flag;
    unsigned iter, iter_max;
    double resid, resid_max;
} monitor;

static void mon_reset(monitor* mo) { mo->flag = 0; mo->iter = 0; mo->resid = INFINITY; }

/* infinity norm with NaN propagation (blas_xnrm8, cblas.h:283-300) */
static double nrm_inf(uint32_t n, const double* x)
{
    double r = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        double a = fabs(x[i]);
        if (a > r || a != a) r = a;
        if (r != r) return r;
    }
    return r;
}

/* Monitor::finished (bicgstab.cc:9-49) */
static int mon_finished(monitor* mo, uint32_t n, const double* x)
{
    mo->resid = nrm_inf(n, x);
    if (mo->resid != mo->resid) { mo->resid = INFINITY; return 1; }
    if (mo->iter > mo->iter_max) return 1;
    return mo->resid < mo->resid_max;
}

static int mon_finished_flag(monitor* mo, int f, uint32_t n, const double* x)
{
    mo->flag = f;
    return mon_finished(mo, n, x);
}

static double dot(uint32_t n, const double* a, const double* b)
{
    double s = 0.0;
    for (uint32_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static void axpy(uint32_t n, double a, const double* x, double* y)
{
    for (uint32_t i = 0; i < n; ++i) y[i] += a * x[i];
}

/* Solver::BCGSP (bicgstab.h:247-338); precond=0 gives Solver::BCGS (163-243) */
static void bicgstab(const orc_meca* m, int precond, const double* rhs, double* x, monitor* mo)
{
    const uint32_t n = m->dim * m->npts;
    double rho1 = 1, rho2, alpha = 0, beta = 0, omega = 1.0, delta;
    double* mem = (double*)malloc(sizeof(double) * 7 * (n ? n : 1));
    double *r = mem, *rt = mem + n, *p = mem + 2 * n, *t = mem + 3 * n, *v = mem + 4 * n;
    double *ph = mem + 5 * n, *sh = mem + 6 * n;
    memcpy(r, rhs, sizeof(double) * n);
    orc_multiply(m, x, rt);
    axpy(n, -1.0, rt, r);
    memcpy(rt, r, sizeof(double) * n);
    while (!mon_finished(mo, n, r)) {
        rho2 = rho1;
        rho1 = dot(n, rt, r);
        if (rho1 == 0.0) { mon_finished_flag(mo, 2, n, r); break; }
        beta = (rho1 / rho2) * (alpha / omega);
        if (beta == 0.0) {
            memcpy(p, r, sizeof(double) * n);
        } else {
            axpy(n, -omega, v, p);
            for (uint32_t i = 0; i < n; ++i) p[i] *= beta;
            axpy(n, 1.0, r, p);
        }
        const double* pp = p;
        if (precond) { orc_precondition(m, p, ph); pp = ph; }
        orc_multiply(m, pp, v);
        delta = dot(n, rt, v);
        if (precond && delta == 0.0) { mon_finished_flag(mo, 4, n, r); break; }
        alpha = rho1 / delta;
        axpy(n, -alpha, v, r);
        axpy(n, alpha, pp, x);
        const double* ss = r;
        if (precond) { orc_precondition(m, r, sh); ss = sh; }
        orc_multiply(m, ss, t);
        double tdt = dot(n, t, t);
        if (tdt == 0.0) { mon_finished_flag(mo, 0, n, r); break; }
        omega = dot(n, t, r) / tdt;
        if (omega == 0.0) { mon_finished_flag(mo, 3, n, r); break; }
        axpy(n, omega, ss, x);
        axpy(n, -omega, t, r);
        ++mo->iter;
    }
    free(mem);
}

/* ------------------------------------------------------------------------ */
/* Meca::solve phases                                                        */

int orc_setup(orc_meca* m, const double* gauss)
{
    if (!m->prepared) return -1;
    const int D = m->dim;
    const uint32_t n = D * m->npts;
    /* meca.cc:605-619 */
    m->use_mB = sp_nonzero(&m->mB);
    if (m->use_mB) sp_prepare(&m->mB);
    m->use_mC = sp_nonzero(&m->mC);
    if (m->use_mC) sp_prepare(&m->mC);
    /* meca.cc:622 */
    orc_forces_linear(m, m->vPTS, m->vFOR, 1);
    /* Brownian terms, RigidFiber::addBrownianForces (rigid_fiber.cc:90-98), meca.cc:633-643 */
    double noise = INFINITY;
    const double sc = m->prm.kT / m->time_step;
    for (int i = 0; i < m->nfib; ++i) {
        const orc_fiber* f = &m->fib[i];
        /* Solid::addBrownianForces (solid.cc:714-725) returns b / soDrag */
        const double b = f->kind == 1 ? sqrt(2 * sc * f->drag / f->np) : sqrt(2 * sc / f->mobility);
        const uint32_t o = D * f->idx;
        if (gauss)
            for (uint32_t j = 0; j < D * f->np; ++j) m->vFOR[o + j] += b * gauss[o + j];
        double th = f->kind == 1 ? b / f->drag : f->mobility * b;
        if (th < noise) noise = th;
    }
    m->noise_level = noise * m->time_step;
    /* meca.cc:646-651 */
    for (int i = 0; i < m->nfib; ++i) {
        const orc_fiber* f = &m->fib[i];
        const uint32_t o = D * f->idx;
        fiber_speeds(m, f, m->vFOR + o, m->vRHS + o, m->time_step, 1);
    }
    /* meca.cc:687-693 */
    for (int i = 0; i < m->nfib; ++i) fiber_make_projection_diff(m, &m->fib[i]);
    (void)n;
    m->setup_done = 1;
    return 0;
}

int orc_iterate(orc_meca* m, const orc_solve_params* sp, const double* guess, orc_report* rep)
{
    if (!m->setup_done) return -1;
    const uint32_t n = m->dim * m->npts;
    memset(rep, 0, sizeof(*rep));
    if (guess) memcpy(m->vSOL, guess, sizeof(double) * n);
    else memset(m->vSOL, 0, sizeof(double) * n);              /* meca.cc:704 */
    monitor mo;
    mon_reset(&mo);
    mo.iter_max = sp->max_iter > 0 ? (unsigned)sp->max_iter : n;   /* meca.cc:731 */
    mo.resid_max = sp->tol_mode == 1 ? sp->tolerance * nrm_inf(n, m->vRHS)
                                     : sp->tolerance * m->noise_level;
    int precond = sp->precondition;
    int pc_ok = 0;
    if (precond) { orc_compute_preconditioner(m); pc_ok = 1; }
    bicgstab(m, precond && pc_ok, m->vRHS, m->vSOL, &mo);          /* meca.cc:737-740 */
    rep->total_iterations = (int)mo.iter;
    rep->precond_used = precond && pc_ok;
    if (!(mo.resid < mo.resid_max)) {                               /* meca.cc:749-799 */
        memcpy(m->vSOL, m->vRHS, sizeof(double) * n);
        mon_reset(&mo);
        bicgstab(m, precond, m->vRHS, m->vSOL, &mo);
        rep->total_iterations += (int)mo.iter;
        if (mo.resid < mo.resid_max) {
            rep->fallback = 1;
        } else {
            memset(m->vSOL, 0, sizeof(double) * n);
            mon_reset(&mo);
            if (precond) {
                bicgstab(m, 0, m->vRHS, m->vSOL, &mo);
                rep->precond_used = 0;
            } else {
                orc_compute_preconditioner(m);
                bicgstab(m, 1, m->vRHS, m->vSOL, &mo);
                rep->precond_used = 1;
            }
            rep->total_iterations += (int)mo.iter;
            rep->fallback = (mo.resid < mo.resid_max) ? 2 : 3;
        }
    }
    rep->iterations = (int)mo.iter;
    rep->flag = mo.flag;
    rep->residual = mo.resid;
    rep->threshold = mo.resid_max;
    rep->converged = mo.resid < mo.resid_max;
    rep->noise_level = m->noise_level;
    return rep->fallback == 3 ? 1 : 0;   /* 1: "convergence failure in solver" */
}

/* FiberNaked::reshape_two (fiber_naked.cc:172-193) */
static void reshape_two(int D, const double* src, double* dst, double cut)
{
    double dif[3] = { 0, 0, 0 }, nn = 0.0;
    for (int d = 0; d < D; ++d) { dif[d] = src[D + d] - src[d]; nn += dif[d] * dif[d]; }
    double s = 0.5 * (1 - cut / sqrt(nn));
    for (int d = 0; d < D; ++d) {
        double v = dif[d] * s;
        dst[d] = src[d] + v;
        dst[D + d] = src[D + d] - v;
    }
}

/* FiberNaked::reshape_it (fiber_naked.cc:221-354): Newton on the segment
   lengths with a tridiagonal (dgtsv) linear solve; returns info != 0 on failure. */
static int reshape_it(int D, uint32_t ns, const double* src, double* dst, double cut)
{
    const double a2 = cut * cut;
    double* mem = (double*)malloc(sizeof(double) * ns * (2 * 3 + 5));
    double *dif = mem, *vec = mem + 3 * ns, *sca = mem + 6 * ns, *val = sca + ns;
    double *dia = val + ns, *low = dia + ns, *upe = low + ns;
    int info = 0;
    for (uint32_t p = 0; p < ns; ++p) {
        for (int d = 0; d < 3; ++d) dif[3 * p + d] = d < D ? src[D * p + D + d] - src[D * p + d] : 0.0;
        sca[p] = 0;
    }
    double err = 0;
    unsigned it = 0;
#define DF(p, d) dif[3 * (p) + (d)]
#define VC(p, d) vec[3 * (p) + (d)]
    do {
        for (int d = 0; d < 3; ++d) {
            VC(0, d) = (1 - 2 * sca[0]) * DF(0, d) + sca[1] * DF(1, d);
            for (uint32_t p = 1; p + 1 < ns; ++p)
                VC(p, d) = sca[p - 1] * DF(p - 1, d) + (1 - 2 * sca[p]) * DF(p, d) + sca[p + 1] * DF(p + 1, d);
            VC(ns - 1, d) = sca[ns - 2] * DF(ns - 2, d) + (1 - 2 * sca[ns - 1]) * DF(ns - 1, d);
        }
        {
            double vv = 0, vd = 0;
            for (int d = 0; d < 3; ++d) { vv += VC(0, d) * VC(0, d); vd += VC(0, d) * DF(0, d); }
            val[0] = vv - a2;
            dia[0] = -2 * vd;
        }
        for (uint32_t p = 1; p < ns; ++p) {
            double vv = 0, vl = 0, vd = 0, vu = 0;
            for (int d = 0; d < 3; ++d) {
                vv += VC(p, d) * VC(p, d);
                vl += VC(p, d) * DF(p - 1, d);
                vd += VC(p, d) * DF(p, d);
                vu += VC(p - 1, d) * DF(p, d);
            }
            val[p] = vv - a2;
            low[p] = vl;
            dia[p] = -2 * vd;
            upe[p - 1] = vu;
        }
        info = orc_dgtsv((int)ns, low + 1, dia, upe, val);
        if (info) goto finish;
        err = 0;
        for (uint32_t p = 0; p < ns; ++p) {
            sca[p] += -0.5 * val[p];
            err += fabs(val[p]);
        }
        if (++it > 32) { info = 1; goto finish; }
    } while (err > 0.001);
    {
        double e[3], dd[3] = { 0, 0, 0 };
        for (int d = 0; d < 3; ++d) e[d] = sca[0] * DF(0, d);
        for (int d = 0; d < D; ++d) dst[d] = src[d] + e[d];
        for (uint32_t p = 1; p < ns; ++p) {
            for (int d = 0; d < 3; ++d) dd[d] = sca[p] * DF(p, d);
            for (int d = 0; d < D; ++d) dst[D * p + d] = src[D * p + d] + dd[d] - e[d];
            for (int d = 0; d < 3; ++d) e[d] = dd[d];
        }
        for (int d = 0; d < D; ++d) dst[D * ns + d] = src[D * ns + d] - dd[d];
    }
#undef DF
#undef VC
finish:
    free(mem);
    return info;
}

/* FiberNaked::reshape_sure (fiber_naked.cc:367-401) */
static void reshape_sure(int D, uint32_t ns, double* x, double cut)
{
    double dp[3] = { 0, 0, 0 }, sum[3] = { 0, 0, 0 }, seg[3];
    double dis = 0;
    for (int d = 0; d < D; ++d) { seg[d] = x[D + d] - x[d]; dis += seg[d] * seg[d]; }
    dis = sqrt(dis);
    if (dis > REAL_EPSILON)
        for (int d = 0; d < D; ++d) dp[d] = (cut / dis - 1.0) * seg[d];
    for (uint32_t p = 1; p < ns; ++p) {
        dis = 0;
        for (int d = 0; d < D; ++d) { seg[d] = x[D * p + D + d] - x[D * p + d]; dis += seg[d] * seg[d]; }
        dis = sqrt(dis);
        for (int d = 0; d < D; ++d) { x[D * p + d] += dp[d]; sum[d] += dp[d]; }
        if (dis > REAL_EPSILON)
            for (int d = 0; d < D; ++d) dp[d] += (cut / dis - 1.0) * seg[d];
    }
    for (int d = 0; d < D; ++d) x[D * ns + d] += dp[d];
    for (int d = 0; d < D; ++d) sum[d] = (sum[d] + dp[d]) * (-1.0 / (ns + 1));
    for (uint32_t p = 0; p <= ns; ++p)
        for (int d = 0; d < D; ++d) x[D * p + d] += sum[d];
}

/* FiberNaked::getPoints (fiber_naked.cc:414-425) */
int orc_reshape(int D, uint32_t ns, const double* src, double* dst, double cut)
{
    if (ns == 1) { reshape_two(D, src, dst, cut); return 0; }
    if (reshape_it(D, ns, src, dst, cut)) {
        memcpy(dst, src, sizeof(double) * D * (ns + 1));
        reshape_sure(D, ns, dst, cut);
        return 1;
    }
    return 0;
}

/* Jacobi eigenvalue iteration on a symmetric 4x4 (column-major, full):
   eigenvector of the largest eigenvalue, unit 2-norm (the result LAPACK
   dsyevx('V','I') returns for il = iu = 4, up to its sign) */
static void sym4_top_eigenvector(const double* A0, double* v)
{
    double A[16], V[16];
    memcpy(A, A0, sizeof(A));
    for (int i = 0; i < 16; ++i) V[i] = (i % 5 == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0;
        for (int q = 1; q < 4; ++q)
            for (int p = 0; p < q; ++p) off += A[p + 4 * q] * A[p + 4 * q];
        if (off == 0.0) break;
        for (int p = 0; p < 3; ++p)
            for (int q = p + 1; q < 4; ++q) {
                const double apq = A[p + 4 * q];
                if (apq == 0.0) continue;
                const double th = (A[q + 4 * q] - A[p + 4 * p]) / (2 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
                const double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < 4; ++k) {      /* A <- A J */
                    const double akp = A[k + 4 * p], akq = A[k + 4 * q];
                    A[k + 4 * p] = c * akp - s * akq;
                    A[k + 4 * q] = s * akp + c * akq;
                }
                for (int k = 0; k < 4; ++k) {      /* A <- J' A */
                    const double apk = A[p + 4 * k], aqk = A[q + 4 * k];
                    A[p + 4 * k] = c * apk - s * aqk;
                    A[q + 4 * k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 4; ++k) {      /* V <- V J */
                    const double vkp = V[k + 4 * p], vkq = V[k + 4 * q];
                    V[k + 4 * p] = c * vkp - s * vkq;
                    V[k + 4 * q] = s * vkp + c * vkq;
                }
            }
    }
    int best = 0;
    for (int i = 1; i < 4; ++i)
        if (A[i + 4 * i] > A[best + 4 * best]) best = i;
    for (int k = 0; k < 4; ++k) v[k] = V[k + 4 * best];
}

/* Solid::getPoints (solid.cc:635-650): copy, then reshape() every 8th call
   (solid.cc:547-620 in 3D after Horn 1987; 503-540 in 2D), rescale()
   (solid.cc:434-467) otherwise */
static int solid_get_points(orc_meca* m, orc_fiber* f)
{
    const int D = m->dim;
    double* P = m->pos + D * f->idx;
    memcpy(P, m->vPTS + D * f->idx, sizeof(double) * D * f->np);
    if (f->np < 2) return 0;
    const double N = (double)f->np;
    double cc[3] = { 0, 0, 0 }, pp[3] = { 0, 0, 0 };
    for (int d = 0; d < D; ++d) {
        cc[d] = P[d];
        pp[d] = P[d] * P[d];
    }
    for (uint32_t p = 1; p < f->np; ++p)
        for (int d = 0; d < D; ++d) {
            cc[d] += P[D * p + d];
            pp[d] += P[D * p + d] * P[D * p + d];
        }
    for (int d = 0; d < D; ++d) cc[d] *= 1.0 / N;
    if (++f->timer > 7) {
        f->timer = 0;
        const double* Sh = f->shape;
        if (D == 2) {
            double a = 0, b = 0;
            for (uint32_t p = 0; p < f->np; ++p) {
                a += P[2 * p] * Sh[2 * p] + P[2 * p + 1] * Sh[2 * p + 1];
                b += Sh[2 * p] * P[2 * p + 1] - Sh[2 * p + 1] * P[2 * p];
            }
            const double n = sqrt(a * a + b * b);
            double c = 1, s = 0;
            if (n > REAL_EPSILON) { c = a / n; s = b / n; }
            for (uint32_t p = 0; p < f->np; ++p) {
                P[2 * p] = c * Sh[2 * p] - s * Sh[2 * p + 1] + cc[0];
                P[2 * p + 1] = s * Sh[2 * p] + c * Sh[2 * p + 1] + cc[1];
            }
            return 0;
        }
        double S[9] = { 0 };
        for (uint32_t p = 0; p < f->np; ++p)
            for (int dd = 0; dd < 3; ++dd)
                for (int ee = 0; ee < 3; ++ee) S[dd + 3 * ee] += Sh[3 * p + dd] * P[3 * p + ee];
        double mx = fabs(S[0]);
        if (fabs(S[4]) > mx) mx = fabs(S[4]);
        if (fabs(S[8]) > mx) mx = fabs(S[8]);
        const double sc = 1.0 / mx;
        double Nm[16];
        Nm[0 + 4 * 0] = sc * (S[0] + S[4] + S[8]);
        Nm[0 + 4 * 1] = sc * (S[1 + 3 * 2] - S[2 + 3 * 1]);
        Nm[0 + 4 * 2] = sc * (S[2 + 3 * 0] - S[0 + 3 * 2]);
        Nm[0 + 4 * 3] = sc * (S[0 + 3 * 1] - S[1 + 3 * 0]);
        Nm[1 + 4 * 1] = sc * (S[0] - S[4] - S[8]);
        Nm[1 + 4 * 2] = sc * (S[0 + 3 * 1] + S[1 + 3 * 0]);
        Nm[1 + 4 * 3] = sc * (S[2 + 3 * 0] + S[0 + 3 * 2]);
        Nm[2 + 4 * 2] = sc * (S[4] - S[0] - S[8]);
        Nm[2 + 4 * 3] = sc * (S[1 + 3 * 2] + S[2 + 3 * 1]);
        Nm[3 + 4 * 3] = sc * (S[8] - S[4] - S[0]);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < i; ++j) Nm[i + 4 * j] = Nm[j + 4 * i];
        double q[4];
        sym4_top_eigenvector(Nm, q);
        /* Quaternion::setMatrix3 (quaternion.h:376-400) */
        const double x2 = q[1] + q[1], y2 = q[2] + q[2], z2 = q[3] + q[3];
        const double rx = q[0] * x2, ry = q[0] * y2, rz = q[0] * z2;
        const double xx = q[1] * x2, xy = q[1] * y2, xz = q[1] * z2;
        const double yy = q[2] * y2, yz = q[2] * z2, zz = q[3] * z2;
        const double Rm[9] = { 1.0 - (yy + zz), xy + rz, xz - ry,
                               xy - rz, 1.0 - (xx + zz), yz + rx,
                               xz + ry, yz - rx, 1.0 - (xx + yy) };
        for (uint32_t p = 0; p < f->np; ++p) {
            const double* h = Sh + 3 * p;
            P[3 * p] = cc[0] + Rm[0] * h[0] + Rm[3] * h[1] + Rm[6] * h[2];
            P[3 * p + 1] = cc[1] + Rm[1] * h[0] + Rm[4] * h[1] + Rm[7] * h[2];
            P[3 * p + 2] = cc[2] + Rm[2] * h[0] + Rm[5] * h[1] + Rm[8] * h[2];
        }
        return 0;
    }
    double sz = 0.0;
    for (int d = 0; d < D; ++d) sz += pp[d] - N * cc[d] * cc[d];
    if (sz <= 0) return -1;
    const double scale = sqrt(f->shape_sqr / sz);
    for (uint32_t p = 0; p < f->np; ++p)
        for (int d = 0; d < D; ++d) P[D * p + d] = scale * (P[D * p + d] - cc[d]) + cc[d];
    return 0;
}

int orc_apply(orc_meca* m)
{
    const int D = m->dim;
    const uint32_t n = D * m->npts;
    axpy(n, 1.0, m->vSOL, m->vPTS);                 /* meca.cc:802 */
    orc_forces_linear(m, m->vPTS, m->vFOR, 0);      /* meca.cc:826 */
    for (int i = 0; i < m->nfib; ++i) {             /* meca.cc:830-835 */
        orc_fiber* f = &m->fib[i];
        if (f->kind == 1) {
            if (solid_get_points(m, f)) return -4;   /* "Cannot rescale the Solid" */
            continue;
        }
        orc_reshape(D, f->np - 1, m->vPTS + D * f->idx, m->pos + D * f->idx, f->cut);
    }
    return 0;
}

int orc_solve(orc_meca* m, const orc_solve_params* sp, const double* gauss, orc_report* rep)
{
    int r = orc_setup(m, gauss);
    if (r) return r;
    r = orc_iterate(m, sp, NULL, rep);
    if (r) return r;
    return orc_apply(m);
}

/* Meca::computeForces() (meca.cc:550-580) with RigidFiber::computeTensions
   (rigid_fiber_project.cc:153-161) */
void orc_compute_forces(orc_meca* m)
{
    const int D = m->dim;
    m->use_mB = sp_nonzero(&m->mB);
    if (m->use_mB) sp_prepare(&m->mB);
    m->use_mC = sp_nonzero(&m->mC);
    if (m->use_mC) sp_prepare(&m->mC);
    orc_forces_linear(m, m->vPTS, m->vFOR, 0);
    for (int i = 0; i < m->nfib; ++i) {
        const orc_fiber* f = &m->fib[i];
        if (f->kind != 0) continue;
        const uint32_t S = f->np - 1;
        const double* dif = m->diff + D * f->seg;
        const double* F = m->vFOR + D * f->idx;
        double* lag = m->lag + f->seg;
        for (uint32_t s = 0; s < S; ++s) {
            double t = 0.0;
            for (int d = 0; d < D; ++d) t += dif[D * s + d] * (F[D * s + D + d] - F[D * s + d]);
            lag[s] = t;
        }
        orc_dptts2((int)S, m->jd + f->seg, m->je + f->seg, lag);
    }
}

/* ------------------------------------------------------------------------ */
/* export                                                                    */

void orc_get_vector(const orc_meca* m, int which, double* out)
{
    const double* src = NULL;
    switch (which) {
        case ORC_PTS: src = m->vPTS; break;
        case ORC_SOL: src = m->vSOL; break;
        case ORC_BAS: src = m->vBAS; break;
        case ORC_RHS: src = m->vRHS; break;
        case ORC_FOR: src = m->vFOR; break;
        default: return;
    }
    memcpy(out, src, sizeof(double) * m->dim * m->npts);
}

void orc_get_fiber_points(const orc_meca* m, double* out)
{
    memcpy(out, m->pos, sizeof(double) * m->dim * m->npts);
}

void orc_get_tensions(const orc_meca* m, double* out)
{
    memcpy(out, m->lag, sizeof(double) * m->nseg);
}

double orc_get_mobility(const orc_meca* m, int i) { return m->fib[i].mobility; }

/* soDrag, soDragRot, soMom2D, soCenter[3], soMom (upper Cholesky factor, 9) */
void orc_get_solid(const orc_meca* m, int i, double* out)
{
    const orc_fiber* f = &m->fib[i];
    out[0] = f->drag;
    out[1] = f->drag_rot;
    out[2] = f->mom2d;
    for (int d = 0; d < 3; ++d) out[3 + d] = f->center[d];
    for (int k = 0; k < 9; ++k) out[6 + k] = f->mom[k];
}

int orc_solid_timer(const orc_meca* m, int i) { return m->fib[i].timer; }

void orc_get_jjt(const orc_meca* m, int i, double* diag, double* offd)
{
    const orc_fiber* f = &m->fib[i];
    memcpy(diag, m->jd + f->seg, sizeof(double) * (f->np - 1));
    if (f->np > 2) memcpy(offd, m->je + f->seg, sizeof(double) * (f->np - 2));
}

uint32_t orc_get_matrix_nr(orc_meca* m, int which, uint32_t* ija, double* sa)
{
    spsym* s = which ? &m->mC : &m->mB;
    sp_prepare(s);
    uint32_t nbe = s->ija[s->size];
    if (ija) memcpy(ija, s->ija, sizeof(uint32_t) * nbe);
    if (sa) memcpy(sa, s->sa, sizeof(double) * nbe);
    return nbe;
}

void orc_get_matrix_dense(const orc_meca* m, int which, double* out)
{
    const spsym* s = which ? &m->mC : &m->mB;
    const size_t n = s->size;
    memset(out, 0, sizeof(double) * n * n);
    for (uint32_t j = 0; j < s->size; ++j)
        for (uint32_t k = 0; k < s->colSize[j]; ++k) {
            uint32_t i = s->col[j][k].line;
            out[i + n * j] = s->col[j][k].val;
            out[j + n * i] = s->col[j][k].val;
        }
}

/* the storage on its own: thin wrappers of the sp_* routines above */
struct orc_sparse { spsym m; };

orc_sparse* orc_sparse_create(uint32_t n)
{
    orc_sparse* s = (orc_sparse*)calloc(1, sizeof(orc_sparse));
    sp_allocate(&s->m, n);
    return s;
}

void orc_sparse_destroy(orc_sparse* s)
{
    if (!s) return;
    sp_free(&s->m);
    free(s);
}

void orc_sparse_zero(orc_sparse* s) { sp_zero(&s->m); }

void orc_sparse_add(orc_sparse* s, uint32_t i, uint32_t j, double v) { *sp_ref(&s->m, i, j) += v; }

uint32_t orc_sparse_nnz(const orc_sparse* s)
{
    uint32_t n = 0;
    for (uint32_t j = 0; j < s->m.size; ++j) n += s->m.colSize[j];
    return n;
}

int orc_sparse_nonzero(const orc_sparse* s) { return sp_nonzero(&s->m); }

void orc_sparse_prepare(orc_sparse* s) { sp_prepare(&s->m); }

void orc_sparse_mul_add(const orc_sparse* s, int dim, const double* X, double* Y)
{
    if (dim > 1)
        sp_mul_add_iso(&s->m, dim, X, Y);
    else
        sp_mul_add(&s->m, X, Y);
}

void orc_sparse_block(const orc_sparse* s, int full, uint32_t x, uint32_t sx, double* M)
{
    if (full)
        sp_add_diagonal_block(&s->m, M, x, sx);
    else
        sp_add_triangular_block(&s->m, M, x, sx);
}

uint32_t orc_sparse_nr(const orc_sparse* s, uint32_t* ija, double* sa)
{
    const uint32_t nbe = s->m.ija ? s->m.ija[s->m.size] : 0;
    if (ija && nbe) memcpy(ija, s->m.ija, sizeof(uint32_t) * nbe);
    if (sa && nbe) memcpy(sa, s->m.sa, sizeof(double) * nbe);
    return nbe;
}

/* ------------------------------------------------------------------------ */
/* segments near a place (the FiberGrid query)                               */

/* FiberGrid::nearbySegments (fiber_grid.cc:302-337) by brute force: every
   segment of every fiber (Solids have none), in list order, is tested with
   FiberLocus::projectPoint (fiber_locus.cc:62-97): abs = (w - P1).(P2 - P1) /
   segmentation; inside the segment dis = |w - P1|^2 - abs^2; before the first
   segment of a fiber or after its last one, the squared distance to that end;
   otherwise dis stays infinite (the neighbouring segment covers the place).
   A segment is kept when dis < range^2 (inclusive: dis <= range^2, the test
   of tryToAttach, fiber_grid.cc:273-274), skipping fiber `exclude` (-1:
   none).  The kept segments are written up to `cap`: the fiber, the local
   point of the segment's first end, the abscissa from the minus end
   (segmentation * point + abs, FiberBinder site of fiber_grid.cc:278) and
   dis.  Returns the number kept. */
uint32_t orc_nearby_segments(const orc_meca* m, const double* w, double range, int inclusive, int exclude,
                             uint32_t cap, uint32_t* fiber, uint32_t* point, double* abscissa, double* dis)
{
    const int D = m->dim;
    const double rr = range * range;
    uint32_t n = 0;
    for (int f = 0; f < m->nfib; ++f) {
        const orc_fiber* F = &m->fib[f];
        if (F->kind != 0 || f == exclude) continue;
        const double ls = F->cut;
        for (uint32_t p = 0; p + 1 < F->np; ++p) {
            const double* P1 = m->pos + (size_t)D * (F->idx + p);
            const double* P2 = P1 + D;
            double dx[3] = { 0, 0, 0 }, aw[3] = { 0, 0, 0 };
            for (int d = 0; d < D; ++d) {
                dx[d] = P2[d] - P1[d];
                aw[d] = w[d] - P1[d];
            }
            double dot = aw[0] * dx[0];
            for (int d = 1; d < D; ++d) dot = dot + aw[d] * dx[d];
            const double ab = dot / ls;
            double ds = INFINITY;
            if (ab < 0) {
                if (p == 0) {
                    ds = (P1[0] - w[0]) * (P1[0] - w[0]);
                    for (int d = 1; d < D; ++d) ds = ds + (P1[d] - w[d]) * (P1[d] - w[d]);
                }
            } else if (ab > ls) {
                if (p + 2 == F->np) {
                    ds = (P2[0] - w[0]) * (P2[0] - w[0]);
                    for (int d = 1; d < D; ++d) ds = ds + (P2[d] - w[d]) * (P2[d] - w[d]);
                }
            } else {
                double nn = aw[0] * aw[0];
                for (int d = 1; d < D; ++d) nn = nn + aw[d] * aw[d];
                ds = nn - ab * ab;
            }
            const int keep = inclusive ? !(ds > rr) : (ds < rr);
            if (!keep) continue;
            if (n < cap) {
                fiber[n] = (uint32_t)f;
                point[n] = p;
                abscissa[n] = ls * (double)p + ab;
                dis[n] = ds;
            }
            ++n;
        }
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* motor couples between solves                                             */

/* position and dirFiber of a hand at abscissa ab: PointInterpolated::pos
   (point_set.h:191-200) and FiberNaked::dirPoint (fiber_naked.h:205) at the
   first interpolation point */
static void hand_geometry(const orc_meca* m, int fi, double ab, double* pos, double* dir)
{
    uint32_t m1, m2;
    double co;
    orc_interpolate(m, fi, ab, &m1, &m2, &co);
    const double cut = m->fib[fi].cut;
    for (int d = 0; d < m->dim; ++d) {
        const double p1 = m->pos[m->dim * m1 + d], p2 = m->pos[m->dim * m2 + d];
        pos[d] = p1 + co * (p2 - p1);
        dir[d] = (p2 - p1) / cut;
    }
}

/* FiberBinder::moveBy + checkFiberRange (fiber_binder.cc:59-65, 111-118),
   Hand::handleOutOfRange for an end that neither grows nor shrinks
   (hand.cc:102-122): hold the end (moveToEnd) or detach */
static void move_hand(double* ab, uint8_t* st, uint8_t bit, double dabs, double len, int hold)
{
    *ab += dabs;
    if (*ab < 0.0) {
        if (hold) *ab = 0.0;
        else *st &= (uint8_t)~bit;
    } else if (*ab > len) {
        if (hold) *ab = len;
        else *st &= (uint8_t)~bit;
    }
}

/* Motor::stepLoaded (motor.cc:32-57) without the Kramers detachment test */
static double loaded_dabs(int dim, const double* f, const double* dir, const orc_motor* mp, double dt)
{
    /* MotorProp::complete (motor_prop.cc:44-66) */
    const double max_speed_dt = dt * mp->max_speed;
    const double var_speed_dt = fabs(max_speed_dt) / mp->stall_force;
    double min_dabs, max_dabs;
    if (mp->max_speed > 0) { min_dabs = 0; max_dabs = 2 * max_speed_dt; }
    else { min_dabs = 2 * max_speed_dt; max_dabs = 0; }
    double load = f[0] * dir[0];
    for (int d = 1; d < dim; ++d) load += f[d] * dir[d];
    double dabs = max_speed_dt + load * var_speed_dt;
    if (mp->limit_speed) {
        if (dabs < min_dabs) dabs = min_dabs;
        else if (dabs > max_dabs) dabs = max_dabs;
    }
    return dabs;
}

/* Couple::stepAA / stepAF / stepFA (couple.cc:125-150) with Couple::force1
   (couple.cc:274-282) on the points held by the Mecables */
void orc_step_couples(const orc_meca* m, uint32_t n, const uint32_t* fa, double* aa, const uint32_t* fb, double* ab,
                      const double* w, uint8_t* state, const orc_motor* mp)
{
    const double dt = m->prm.time_step;
    const int D = m->dim;
    for (uint32_t k = 0; k < n; ++k) {
        uint8_t st = state[k];
        if (!(st & 12)) continue;
        const orc_fiber* f1 = &m->fib[fa[k]];
        const orc_fiber* f2 = &m->fib[fb[k]];
        const double len1 = f1->cut * (double)(f1->np - 1), len2 = f2->cut * (double)(f2->np - 1);
        if ((st & 3) == 3) {
            double p1[3], d1[3], p2[3], d2[3], f[3], g[3];
            hand_geometry(m, (int)fa[k], aa[k], p1, d1);
            hand_geometry(m, (int)fb[k], ab[k], p2, d2);
            for (int d = 0; d < D; ++d) {
                f[d] = w[k] * (p2[d] - p1[d]);
                g[d] = -f[d];
            }
            if (st & 4) move_hand(&aa[k], &st, 1, loaded_dabs(D, f, d1, mp, dt), len1, mp->hold_growing_end);
            if (st & 8) move_hand(&ab[k], &st, 2, loaded_dabs(D, g, d2, mp, dt), len2, mp->hold_growing_end);
        } else if (st & 1) {
            /* Motor::stepUnloaded (motor.cc:19-28) without the detachment test */
            if (st & 4) move_hand(&aa[k], &st, 1, dt * mp->max_speed, len1, mp->hold_growing_end);
        } else if (st & 2) {
            if (st & 8) move_hand(&ab[k], &st, 2, dt * mp->max_speed, len2, mp->hold_growing_end);
        }
        state[k] = st;
    }
}

/* ------------------------------------------------------------------------ */
/* trajectory frame (Simul::writeObjects, binary mode)                       */

typedef struct {
    unsigned char* out;
    size_t cap, n;
} fr_out;

static void fr_put(fr_out* o, const void* p, size_t len)
{
    if (o->n + len <= o->cap) memcpy(o->out + o->n, p, len);
    o->n += len;
}

static void fr_str(fr_out* o, const char* s) { fr_put(o, s, strlen(s)); }

/* OutputWrapper::writeUInt8/16/32, writeInt32 (iowrapper.cc:452-533): native
   (little-endian) bytes of the value in binary mode */
static void fr_u8(fr_out* o, uint32_t v) { const uint8_t x = (uint8_t)v; fr_put(o, &x, 1); }
static void fr_u16(fr_out* o, uint32_t v) { const uint16_t x = (uint16_t)v; fr_put(o, &x, 2); }
static void fr_u32(fr_out* o, uint32_t v) { fr_put(o, &v, 4); }
static void fr_i32(fr_out* o, int32_t v) { fr_put(o, &v, 4); }

/* writeFloat (iowrapper.cc:535-547: the double is converted to float) or
   writeDouble (iowrapper.cc:570-582) */
static void fr_real(fr_out* o, double v, int rb)
{
    if (rb == 8) {
        fr_put(o, &v, 8);
    } else {
        const float f = (float)v;
        fr_put(o, &f, 4);
    }
}

/* Object::writeReference (object.cc:76-97) */
static void fr_reference(fr_out* o, char tag, const orc_frame_record* r)
{
    if (r->number <= 65535 && r->prop <= 255 && r->mark == 0) {
        fr_put(o, &tag, 1);
        fr_u8(o, r->prop);
        fr_u16(o, r->number);
    } else {
        fr_put(o, "$", 1);
        fr_put(o, &tag, 1);
        fr_u16(o, r->prop);
        fr_u32(o, r->number);
        fr_i32(o, r->mark);
    }
}

size_t orc_write_frame(const orc_meca* m, const orc_frame_record* rec, double time, const char* date,
                       int rb, const void* before, size_t nbefore, const void* after, size_t nafter,
                       unsigned char* out, size_t cap)
{
    const int D = m->dim;
    fr_out o = { out, out ? cap : 0, 0 };
    char line[256];
    if (!date) date = "";
    /* frame tag, time line, binary signature (simul_file.cc:415-422; FRAME_TAG simul.h:31;
       OutputWrapper::writeBinarySignature, iowrapper.cc:378-388) */
    snprintf(line, sizeof line, "\n\n%s %s", "#Cytosim ", date);
    fr_str(&o, line);
    if (rb == 8)
        snprintf(line, sizeof line, "\n#time %.17g, dim %i, format %i, real 8", time, D, 41);
    else
        snprintf(line, sizeof line, "\n#time %.6f, dim %i, format %i", time, D, 41);
    fr_str(&o, line);
    fr_str(&o, "\n#binary ");
    fr_u16(&o, 12592U);
    fr_str(&o, "\n");
    if (nbefore) fr_put(&o, before, nbefore);
    int nf = 0, ns = 0;
    for (int f = 0; f < m->nfib; ++f) {
        if (m->fib[f].kind == 1) ++ns;
        else ++nf;
    }
    /* fibers (simul_file.cc:443-447): ObjectSet::write (object_set.cc:380-389),
       FiberNaked::write (fiber_naked.cc:1195-1202), PointSet::write (point_set.cc:337-342) */
    if (nf) {
        fr_str(&o, "\n#section fiber");
        for (int f = 0; f < m->nfib; ++f) {
            const orc_fiber* F = &m->fib[f];
            if (F->kind == 1) continue;
            fr_put(&o, "\n", 1);
            fr_reference(&o, 'f', &rec[f]);
            fr_u32(&o, rec[f].signature);
            fr_real(&o, (double)(F->np - 1) * F->cut, rb);   /* length() = nbSegments() * fnCut */
            fr_real(&o, rec[f].cut_wished, rb);
            fr_real(&o, rec[f].abscissa, rb);
            fr_u16(&o, F->np);
            for (uint32_t p = 0; p < F->np; ++p)
                for (int d = 0; d < D; ++d) fr_real(&o, m->pos[D * (F->idx + p) + d], rb);
        }
    }
    /* solids (simul_file.cc:449-453): Solid::write (solid.cc:896-905) */
    if (ns) {
        fr_str(&o, "\n#section solid");
        for (int f = 0; f < m->nfib; ++f) {
            const orc_fiber* F = &m->fib[f];
            if (F->kind != 1) continue;
            fr_put(&o, "\n", 1);
            fr_reference(&o, 'd', &rec[f]);
            fr_u16(&o, F->np);
            for (uint32_t p = 0; p < F->np; ++p) {
                for (int d = 0; d < D; ++d) fr_real(&o, m->pos[D * (F->idx + p) + d], rb);
                fr_real(&o, F->radius[p], rb);
            }
        }
    }
    if (nafter) fr_put(&o, after, nafter);
    fr_str(&o, "\n#section end");
    snprintf(line, sizeof line, "\n#end cytosim %s\n\n", date);
    fr_str(&o, line);
    return o.n;
}
