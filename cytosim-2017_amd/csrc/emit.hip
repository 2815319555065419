// Position-dependent interactions (meca_inter.cc), evaluated on the device.
//
// The host records each call (DeferredCall) in call order; when the streams
// are built, one thread per call reads the points from d_pos and writes the
// call's contributions into the entries reserved for it, in the order the
// reference makes them, and marks the entries it does not make (a link of
// zero length, an arm that cannot be placed) with an invalid key.  The stable
// sort of the streams then sums every matrix entry in call order, as
// MatrixSparseSymmetric1 does.  Every product and sum is rounded as written
// (contraction off): the values are those of the reference's arithmetic.
//
// Geometry of a PointInterpolated (point_interpolated.h:97-124): pos = P1 +
// c (P2 - P1), diff = P2 - P1, coef1 = c, coef2 = 1 - c; vector norms and dot
// products sum their terms in order (vector2.h, vector3.h).  At an exactly
// zero cross product the reference draws a random side (random.cc:73-82);
// here the side is +1.  No periodic space.  The BLAS products of the 2D side
// sliding links and of the 3D side clamp follow the reference BLAS order.
#include "context.h"

namespace cymg {

#pragma clang fp contract(off)

namespace {

constexpr double kEps = 128.0 * 2.220446049250313e-16;   // REAL_EPSILON, real.h

__device__ __forceinline__ uint64_t lkey(uint32_t i, uint32_t j)
{
    const uint32_t r = i > j ? i : j, c = i > j ? j : i;
    return ((uint64_t)c << 32) | r;
}

// the entries reserved for one call in one stream
struct Slots {
    uint64_t* k;
    double* v;
    uint32_t n, cap;
    __device__ void put(uint64_t key, double val)
    {
        if (n < cap) { k[n] = key; v[n] = val; }
        ++n;
    }
    __device__ void close()
    {
        for (uint32_t i = n; i < cap; ++i) { k[i] = kInvalidKey; v[i] = 0.0; }
    }
};

// the three streams of one call: mB(i,j) += v, mC(i,j) += v, vBAS[i] += v
struct Out {
    Slots b, c, s;
    __device__ void B(uint32_t i, uint32_t j, double v) { b.put(lkey(i, j), v); }
    __device__ void C(uint32_t i, uint32_t j, double v) { c.put(lkey(i, j), v); }
    __device__ void S(uint32_t i, double v) { s.put(i, v); }
};

template <int D>
struct Geo {
    const double* X;
    __device__ const double* pt(uint32_t p) const { return X + (size_t)D * p; }
    __device__ void pos(uint32_t p1, uint32_t p2, double c, double* o) const
    {
        const double *x = pt(p1), *y = pt(p2);
#pragma unroll
        for (int d = 0; d < D; ++d) o[d] = x[d] + c * (y[d] - x[d]);
    }
    __device__ void diff(uint32_t p1, uint32_t p2, double* o) const
    {
        const double *x = pt(p1), *y = pt(p2);
#pragma unroll
        for (int d = 0; d < D; ++d) o[d] = y[d] - x[d];
    }
};

template <int D>
__device__ __forceinline__ double dot(const double* a, const double* b)
{
    double s = a[0] * b[0];
#pragma unroll
    for (int d = 1; d < D; ++d) s = s + a[d] * b[d];
    return s;
}

template <int D>
__device__ __forceinline__ double norm(const double* a) { return sqrt(dot<D>(a, a)); }

template <int D>
__device__ __forceinline__ void unit(double* a)
{
    const double sn = 1.0 / norm<D>(a);
#pragma unroll
    for (int d = 0; d < D; ++d) a[d] = a[d] * sn;
}

__device__ __forceinline__ double cross2(const double* a, const double* b) { return a[0] * b[1] - a[1] * b[0]; }

__device__ __forceinline__ double side_of(double x) { return x < 0 ? -1.0 : 1.0; }

// interLongLink PE-PE (meca_inter.cc:334-415)
template <int D>
__device__ void long_link_ee(const Geo<D>& G, const DeferredCall& q, Out& o)
{
    const uint32_t a = q.p[0], b = q.p[1];
    const double w = q.w;
    double len = q.len, ab[D];
    G.diff(a, b, ab);
    const double abn = norm<D>(ab);
    if (abn < kEps) return;
#pragma unroll
    for (int d = 0; d < D; ++d) ab[d] /= abn;
    const uint32_t iA = D * a, iB = D * b;
    for (int d = 0; d < D; ++d) {
        o.S(iA + d, -(w * ab[d] * len));
        o.S(iB + d, w * ab[d] * len);
    }
    len /= abn;
    const bool cooked = len > 1.0;
    for (int i = 0; i < D; ++i) {
        double v = cooked ? w * ab[i] * ab[i] : w * (1.0 + len * (ab[i] * ab[i] - 1.0));
        o.C(iA + i, iA + i, -v);
        o.C(iA + i, iB + i, v);
        o.C(iB + i, iB + i, -v);
        for (int j = i + 1; j < D; ++j) {
            v = cooked ? w * ab[i] * ab[j] : w * len * ab[i] * ab[j];
            o.C(iA + i, iA + j, -v);
            o.C(iA + i, iB + j, v);
            o.C(iA + j, iB + i, v);
            o.C(iB + i, iB + j, -v);
        }
    }
}

// interLongClamp (meca_inter.cc:1977-2022)
template <int D>
__device__ void long_clamp(const Geo<D>& G, const DeferredCall& q, Out& o)
{
    const double w = q.w;
    const double* center = q.g;
    double len = q.len;
    const uint32_t inx = D * q.p[0];
    const double* x = G.pt(q.p[0]);
    double axis[D];
#pragma unroll
    for (int d = 0; d < D; ++d) axis[d] = x[d] - center[d];
    const double an = norm<D>(axis);
    if (an < kEps) return;
#pragma unroll
    for (int d = 0; d < D; ++d) axis[d] /= an;
    const double ac = dot<D>(axis, center);
    if (len < an) {
        len /= an;
        const double facX = w * len * (an + ac);
        const double facC = w * (1.0 - len);
        for (int i = 0; i < D; ++i) {
            o.C(inx + i, inx + i, w * (len * (1.0 - axis[i] * axis[i]) - 1.0));
            for (int j = i + 1; j < D; ++j) o.C(inx + i, inx + j, -(w * len * axis[i] * axis[j]));
            o.S(inx + i, facX * axis[i] + facC * center[i]);
        }
    } else {
        const double facX = w * (len + ac);
        for (int i = 0; i < D; ++i) {
            for (int j = i; j < D; ++j) o.C(inx + i, inx + j, -(w * axis[i] * axis[j]));
            o.S(inx + i, facX * axis[i]);
        }
    }
}

// interLongLink PI-PE (meca_inter.cc:425-505) and PI-PI (516-596) over the
// N = 3 or 4 points with coefficients c, ab = target - pta.pos()
template <int D, int N>
__device__ void long_link_n(const uint32_t* pts, const double* c, double* ab, double len, double w, Out& o)
{
    double cw[N];
    uint32_t inx[N];
#pragma unroll
    for (int k = 0; k < N; ++k) { cw[k] = w * c[k]; inx[k] = D * pts[k]; }
    const double abn = norm<D>(ab);
    if (abn < kEps) return;
#pragma unroll
    for (int d = 0; d < D; ++d) ab[d] /= abn;
    for (int i = 0; i < N; ++i)
        for (int d = 0; d < D; ++d) o.S(inx[i] + d, -(cw[i] * ab[d] * len));
    len /= abn;
    const bool cooked = len > 1.0;
    for (int i = 0; i < D; ++i) {
        double v = cooked ? ab[i] * ab[i] : 1.0 + len * (ab[i] * ab[i] - 1.0);
        for (int k = 0; k < N; ++k)
            for (int l = k; l < N; ++l) o.C(inx[k] + i, inx[l] + i, -(c[k] * cw[l] * v));
        for (int j = i + 1; j < D; ++j) {
            v = cooked ? ab[i] * ab[j] : len * ab[i] * ab[j];
            for (int k = 0; k < N; ++k)
                for (int l = 0; l < N; ++l) o.C(inx[k] + i, inx[l] + j, -(c[k] * cw[l] * v));
        }
    }
}

// 3D arm of the side links: the part of as orthogonal to a, scaled to len
// (meca_inter.cc:748-756, 921-929, 1489-1497, 1741-1749)
__device__ bool side_arm3(const double* a, const double* as, double len, double* arm)
{
    const double s = dot<3>(as, a) / dot<3>(a, a);
#pragma unroll
    for (int d = 0; d < 3; ++d) arm[d] = as[d] - s * a[d];
    const double n = norm<3>(arm);
    if (!(n > kEps)) return false;
    const double f = len / n;
#pragma unroll
    for (int d = 0; d < 3; ++d) arm[d] = arm[d] * f;
    return true;
}

// interSideLinkS (meca_inter.cc:671-726, 834-886) and interSideSlidingLinkS
// (1374-1453, 1626-1706): T = a a' + b b' or b b', RB = arm or len b
template <int N>
__device__ void side_s(const uint32_t* pts, const double* c, const double* a, const double* arm, double len,
                       double w, bool sliding, Out& o)
{
    double b[3];
#pragma unroll
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
#pragma unroll
    for (int d = 0; d < 3; ++d) RB[d] = sliding ? len * b[d] : arm[d];
    double cw[N];
    uint32_t inx[N];
#pragma unroll
    for (int k = 0; k < N; ++k) { cw[k] = -w * c[k]; inx[k] = 3 * pts[k]; }
    for (int i = 0; i < N; ++i) {
        for (int x = 0; x < 3; ++x) o.S(inx[i] + x, cw[i] * RB[x]);
        const double g = c[i] * cw[i];
        for (int x = 0; x < 3; ++x)
            for (int y = x; y < 3; ++y) o.C(inx[i] + x, inx[i] + y, g * T[x + 3 * y]);
        for (int j = i + 1; j < N; ++j) {
            const double h = c[i] * cw[j];
            for (int x = 0; x < 3; ++x)
                for (int y = 0; y < 3; ++y) o.C(inx[i] + x, inx[j] + y, h * T[x + 3 * y]);
        }
    }
}

// interSideLink2D PI-PE (meca_inter.cc:617-667)
__device__ void side_link2d_ie(const DeferredCall& q, double L, double arm, Out& o)
{
    const double w = q.w;
    const double aa = q.ca, bb = 1.0 - q.ca;
    const double ee = arm / L;
    uint32_t i1 = q.p[0], i2 = q.p[1], iE = q.p[2];
    const double aaw = aa * w, bbw = bb * w;
    const double eew = w * ee, eeeew = ee * eew;
    o.B(i1, i1, -(bbw * bb + eeeew));
    o.B(i1, i2, -(aaw * bb - eeeew));
    o.B(i2, i2, -(aaw * aa + eeeew));
    o.B(iE, iE, -w);
    o.B(i1, iE, bbw);
    o.B(i2, iE, aaw);
    i1 *= 2; i2 *= 2; iE *= 2;
    o.C(i1, i2 + 1, eew);
    o.C(i1 + 1, i2, -eew);
    o.C(i1, iE + 1, -eew);
    o.C(i1 + 1, iE, eew);
    o.C(i2, iE + 1, eew);
    o.C(i2 + 1, iE, -eew);
}

// interSideLink2D PI-PI (meca_inter.cc:764-830)
__device__ void side_link2d_ii(const DeferredCall& q, double L, double arm, Out& o)
{
    const double w = q.w;
    const double aa1 = q.ca, bb1 = 1.0 - q.ca;
    const double aa2 = q.cb, bb2 = 1.0 - q.cb;
    const double ee1 = arm / L;
    uint32_t i11 = q.p[0], i12 = q.p[1], i21 = q.p[2], i22 = q.p[3];
    const double aa1w = aa1 * w, aa2w = aa2 * w;
    const double bb1w = bb1 * w, bb2w = bb2 * w;
    const double ee1w = ee1 * w, ee1ee1w = ee1 * ee1w;
    o.B(i11, i11, -(bb1w * bb1 + ee1ee1w));
    o.B(i11, i12, -aa1w * bb1 + ee1ee1w);
    o.B(i12, i12, -(aa1w * aa1 + ee1ee1w));
    o.B(i21, i21, -(bb2w * bb2));
    o.B(i21, i22, -aa2w * bb2);
    o.B(i22, i22, -(aa2w * aa2));
    o.B(i11, i21, bb1w * bb2);
    o.B(i11, i22, bb1w * aa2);
    o.B(i12, i22, aa1w * aa2);
    o.B(i12, i21, aa1w * bb2);
    i11 *= 2; i12 *= 2; i21 *= 2; i22 *= 2;
    o.C(i11, i12 + 1, ee1w);
    o.C(i11 + 1, i12, -ee1w);
    const double ee1bb2w = ee1w * bb2;
    const double ee1aa2w = ee1w * aa2;
    o.C(i11, i21 + 1, -ee1bb2w);
    o.C(i11, i22 + 1, -ee1aa2w);
    o.C(i11 + 1, i21, ee1bb2w);
    o.C(i11 + 1, i22, ee1aa2w);
    o.C(i12, i21 + 1, ee1bb2w);
    o.C(i12, i22 + 1, ee1aa2w);
    o.C(i12 + 1, i21, -ee1bb2w);
    o.C(i12 + 1, i22, -ee1aa2w);
}

// interSideLink PI-PE (meca_inter.cc:730-759) and PI-PI (903-935)
template <int D, bool PAIR>
__device__ void side_link(const Geo<D>& G, const DeferredCall& q, Out& o)
{
    double pa[D], df[D], tg[D], as[D];
    G.pos(q.p[0], q.p[1], q.ca, pa);
    G.diff(q.p[0], q.p[1], df);
    if (PAIR) G.pos(q.p[2], q.p[3], q.cb, tg);
    else for (int d = 0; d < D; ++d) tg[d] = G.pt(q.p[2])[d];
#pragma unroll
    for (int d = 0; d < D; ++d) as[d] = tg[d] - pa[d];
    if constexpr (D == 2) {
        const double arm = q.len * side_of(cross2(df, as));
        if (PAIR) side_link2d_ii(q, norm<2>(df), arm, o);
        else side_link2d_ie(q, norm<2>(df), arm, o);
    } else {
        double arm[3];
        if (!side_arm3(df, as, q.len, arm)) return;
        unit<3>(df);
        if (PAIR) {
            const double c[4] = { 1.0 - q.ca, q.ca, -(1.0 - q.cb), -q.cb };
            side_s<4>(q.p, c, df, arm, q.len, q.w, false, o);
        } else {
            const uint32_t pts[3] = { q.p[0], q.p[1], q.p[2] };
            const double c[3] = { 1.0 - q.ca, q.ca, -1.0 };
            side_s<3>(pts, c, df, arm, q.len, q.w, false, o);
        }
    }
}

// interSideSideLink (meca_inter.cc:1038-1059) with interSideSideLink2D (943-1021)
__device__ void side_side_link(const Geo<2>& G, const DeferredCall& q, Out& o)
{
    double pa[2], pb[2], da[2], db[2], dir[2];
    G.pos(q.p[0], q.p[1], q.ca, pa);
    G.pos(q.p[2], q.p[3], q.cb, pb);
    G.diff(q.p[0], q.p[1], da);
    G.diff(q.p[2], q.p[3], db);
    for (int d = 0; d < 2; ++d) dir[d] = pb[d] - pa[d];
    const double side1 = side_of(cross2(da, dir));
    const double side2 = side_of(cross2(dir, db));
    const double w = q.w, len = q.len;
    const double aa1 = q.ca, bb1 = 1.0 - q.ca;
    const double aa2 = q.cb, bb2 = 1.0 - q.cb;
    const double ee1 = side1 * len / (2 * norm<2>(da));
    const double ee2 = side2 * len / (2 * norm<2>(db));
    uint32_t i11 = q.p[0], i12 = q.p[1], i21 = q.p[2], i22 = q.p[3];
    const double aa1w = aa1 * w, aa2w = aa2 * w;
    const double bb1w = bb1 * w, bb2w = bb2 * w;
    const double ee1w = ee1 * w, ee1ee1w = ee1 * ee1w;
    const double ee2w = ee2 * w, ee2ee2w = ee2 * ee2w;
    const double ee1ee2w = ee1 * ee2w;
    o.B(i11, i11, -(bb1w * bb1 + ee1ee1w));
    o.B(i11, i12, -(aa1w * bb1 - ee1ee1w));
    o.B(i12, i12, -(aa1w * aa1 + ee1ee1w));
    o.B(i21, i21, -(bb2w * bb2 + ee2ee2w));
    o.B(i21, i22, -(aa2w * bb2 - ee2ee2w));
    o.B(i22, i22, -(aa2w * aa2 + ee2ee2w));
    o.B(i11, i21, bb1w * bb2 + ee1ee2w);
    o.B(i11, i22, bb1w * aa2 - ee1ee2w);
    o.B(i12, i22, aa1w * aa2 + ee1ee2w);
    o.B(i12, i21, aa1w * bb2 - ee1ee2w);
    i11 *= 2; i12 *= 2; i21 *= 2; i22 *= 2;
    o.C(i11, i12 + 1, ee1w);
    o.C(i11 + 1, i12, -ee1w);
    o.C(i21, i22 + 1, ee2w);
    o.C(i21 + 1, i22, -ee2w);
    const double ee1aa2w = ee1w * aa2;
    const double ee1bb2w = ee1w * bb2;
    const double ee2aa1w = ee2w * aa1;
    const double ee2bb1w = ee2w * bb1;
    o.C(i11, i21 + 1, ee2bb1w - ee1bb2w);
    o.C(i11, i22 + 1, -(ee2bb1w + ee1aa2w));
    o.C(i11 + 1, i21, ee1bb2w - ee2bb1w);
    o.C(i11 + 1, i22, ee1aa2w + ee2bb1w);
    o.C(i12, i21 + 1, ee2aa1w + ee1bb2w);
    o.C(i12, i22 + 1, ee1aa2w - ee2aa1w);
    o.C(i12 + 1, i21, -(ee1bb2w + ee2aa1w));
    o.C(i12 + 1, i22, -(ee1aa2w - ee2aa1w));
}

// interSlidingLink PI-PE (meca_inter.cc:1075-1140)
template <int D>
__device__ void sliding_link_ie(const Geo<D>& G, const DeferredCall& q, Out& o)
{
    const double w = q.w;
    const double aa = q.ca, bb = 1.0 - q.ca;
    const double bbbb = bb * bb, aaaa = aa * aa, aabb = aa * bb;
    double dir[D];
    G.diff(q.p[0], q.p[1], dir);
    unit<D>(dir);
    const uint32_t i1 = D * q.p[0], i2 = D * q.p[1], iE = D * q.p[2];
    for (int x = 0; x < D; ++x) {
        for (int y = x; y < D; ++y) {
            const double P = w * ((x == y ? 1.0 : 0.0) - dir[x] * dir[y]);
            o.C(i1 + x, i1 + y, -(bbbb * P));
            o.C(i1 + x, i2 + y, -(aabb * P));
            o.C(i2 + x, i2 + y, -(aaaa * P));
            o.C(i1 + x, iE + y, bb * P);
            o.C(i2 + x, iE + y, aa * P);
            o.C(iE + x, iE + y, -P);
            if (x != y) {
                o.C(i1 + y, i2 + x, -(aabb * P));
                o.C(i1 + y, iE + x, bb * P);
                o.C(i2 + y, iE + x, aa * P);
            }
        }
    }
}

// interSlidingLink PI-PI (meca_inter.cc:1153-1258)
template <int D>
__device__ void sliding_link_ii(const Geo<D>& G, const DeferredCall& q, Out& o)
{
    const double c[4] = { 1.0 - q.ca, q.ca, -(1.0 - q.cb), -q.cb };
    double cw[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) cw[k] = -q.w * c[k];
    const uint32_t inx[4] = { D * q.p[0], D * q.p[1], D * q.p[2], D * q.p[3] };
    double dir[D];
    G.diff(q.p[0], q.p[1], dir);
    unit<D>(dir);
    if constexpr (D == 2) {
        const double dd[3] = { dir[0] * dir[0], dir[0] * dir[1], dir[1] * dir[1] };
        for (int j = 0; j < 4; ++j) {
            const double g = c[j] * cw[j];
            o.C(inx[j], inx[j], (1.0 - dd[0]) * g);
            o.C(inx[j] + 1, inx[j] + 1, (1.0 - dd[2]) * g);
            o.C(inx[j], inx[j] + 1, (-dd[1]) * g);
            for (int i = j + 1; i < 4; ++i) {
                const double h = c[j] * cw[i];
                o.C(inx[i], inx[j], (1.0 - dd[0]) * h);
                o.C(inx[i] + 1, inx[j] + 1, (1.0 - dd[2]) * h);
                o.C(inx[i], inx[j] + 1, (-dd[1]) * h);
                o.C(inx[i] + 1, inx[j], (-dd[1]) * h);
            }
        }
    } else {
        const double dd[6] = { dir[0] * dir[0], dir[0] * dir[1], dir[0] * dir[2],
                               dir[1] * dir[1], dir[1] * dir[2], dir[2] * dir[2] };
        for (int j = 0; j < 4; ++j) {
            const double g = c[j] * cw[j];
            o.C(inx[j], inx[j], (1.0 - dd[0]) * g);
            o.C(inx[j] + 1, inx[j] + 1, (1.0 - dd[3]) * g);
            o.C(inx[j] + 2, inx[j] + 2, (1.0 - dd[5]) * g);
            o.C(inx[j], inx[j] + 1, (-dd[1]) * g);
            o.C(inx[j], inx[j] + 2, (-dd[2]) * g);
            o.C(inx[j] + 1, inx[j] + 2, (-dd[4]) * g);
            for (int i = j + 1; i < 4; ++i) {
                const double h = c[j] * cw[i];
                o.C(inx[i], inx[j], (1.0 - dd[0]) * h);
                o.C(inx[i] + 1, inx[j] + 1, (1.0 - dd[3]) * h);
                o.C(inx[i] + 2, inx[j] + 2, (1.0 - dd[5]) * h);
                o.C(inx[i], inx[j] + 1, (-dd[1]) * h);
                o.C(inx[i], inx[j] + 2, (-dd[2]) * h);
                o.C(inx[i] + 1, inx[j] + 2, (-dd[4]) * h);
                o.C(inx[i] + 1, inx[j], (-dd[1]) * h);
                o.C(inx[i] + 2, inx[j], (-dd[2]) * h);
                o.C(inx[i] + 2, inx[j] + 1, (-dd[4]) * h);
            }
        }
    }
}

// interSideSlidingLink2D PI-PE (meca_inter.cc:1266-1366) and PI-PI
// (1508-1618): TPT = T' (-w P) T with netlib's dgemm order, N = 6 or 8
template <int N>
__device__ void side_sliding2d(const Geo<2>& G, const DeferredCall& q, double arm, Out& o)
{
    double dir[2], df[2];
    G.diff(q.p[0], q.p[1], df);
    G.diff(q.p[0], q.p[1], dir);
    unit<2>(dir);
    const double aa = q.ca, bb = 1.0 - q.ca;
    const double ee = arm / norm<2>(df);
    double T[2 * N];
    T[0] = bb; T[1] = -ee; T[2] = ee; T[3] = bb; T[4] = aa; T[5] = ee; T[6] = -ee; T[7] = aa;
    if constexpr (N == 6) {
        T[8] = -1; T[9] = 0; T[10] = 0; T[11] = -1;
    } else {
        const double cc = q.cb, dd = 1.0 - q.cb;
        T[8] = -dd; T[9] = 0; T[10] = 0; T[11] = -dd; T[12] = -cc; T[13] = 0; T[14] = 0; T[15] = -cc;
    }
    const double P[4] = { 1 - dir[0] * dir[0], -dir[0] * dir[1], -dir[0] * dir[1], 1 - dir[1] * dir[1] };
    const double alpha = -q.w;
    double PT[2 * N];
    for (int j = 0; j < N; ++j) {
        PT[2 * j] = 0.0;
        PT[2 * j + 1] = 0.0;
        for (int l = 0; l < 2; ++l) {
            const double t = alpha * T[l + 2 * j];
            PT[2 * j] = PT[2 * j] + t * P[2 * l];
            PT[2 * j + 1] = PT[2 * j + 1] + t * P[1 + 2 * l];
        }
    }
    uint32_t inx[N];
    for (int k = 0; k < N / 2; ++k) { inx[2 * k] = 2 * q.p[k]; inx[2 * k + 1] = 2 * q.p[k] + 1; }
    for (int i = 0; i < N; ++i)
        for (int j = i; j < N; ++j) {
            double t = 0.0;
            t = t + T[2 * i] * PT[2 * j];
            t = t + T[1 + 2 * i] * PT[1 + 2 * j];
            o.C(inx[i], inx[j], 1.0 * t);
        }
}

// interSideSlidingLink PI-PE (meca_inter.cc:1469-1499) and PI-PI (1724-1751)
template <int D, bool PAIR>
__device__ void side_sliding(const Geo<D>& G, const DeferredCall& q, Out& o)
{
    double pa[D], df[D], tg[D], as[D];
    G.pos(q.p[0], q.p[1], q.ca, pa);
    G.diff(q.p[0], q.p[1], df);
    if (PAIR) G.pos(q.p[2], q.p[3], q.cb, tg);
    else for (int d = 0; d < D; ++d) tg[d] = G.pt(q.p[2])[d];
#pragma unroll
    for (int d = 0; d < D; ++d) as[d] = tg[d] - pa[d];
    if constexpr (D == 2) {
        const double arm = q.len * side_of(cross2(df, as));
        if (PAIR) side_sliding2d<8>(G, q, arm, o);
        else side_sliding2d<6>(G, q, arm, o);
    } else {
        double arm[3];
        if (!side_arm3(df, as, q.len, arm)) return;
        unit<3>(df);
        if (PAIR) {
            const double c[4] = { 1.0 - q.ca, q.ca, -(1.0 - q.cb), -q.cb };
            side_s<4>(q.p, c, df, arm, q.len, q.w, true, o);
        } else {
            const uint32_t pts[3] = { q.p[0], q.p[1], q.p[2] };
            const double c[3] = { 1.0 - q.ca, q.ca, -1.0 };
            side_s<3>(pts, c, df, arm, q.len, q.w, true, o);
        }
    }
}

// interTorque2D (meca_inter.cc:1772-1869)
__device__ void torque(const Geo<2>& G, const DeferredCall& q, Out& o)
{
    const uint32_t a1 = q.p[0], a2 = q.p[1], b1 = q.p[2], b2 = q.p[3];
    const uint32_t index[8] = { 2 * a1, 2 * a1 + 1, 2 * a2, 2 * a2 + 1, 2 * b1, 2 * b1 + 1, 2 * b2, 2 * b2 + 1 };
    double ab[2], cd[2];
    G.diff(a1, a2, ab);
    G.diff(b1, b2, cd);
    const double *a = G.pt(a1), *b = G.pt(a2), *c = G.pt(b1), *d = G.pt(b2);
    const double coord[8] = { a[0], a[1], b[0], b[1], c[0], c[1], d[0], d[1] };
    const double cosinus = q.g[0], sinus = q.g[1];
    double ce[2];
    ce[0] = cd[0] * cosinus + cd[1] * sinus;
    ce[1] = -cd[0] * sinus + cd[1] * cosinus;
    const double abn = norm<2>(ab), abnS = dot<2>(ab, ab);
    const double cdn = norm<2>(cd), cdnS = dot<2>(cd, cd);
    if (abn < kEps || cdn < kEps) return;
    for (int k = 0; k < 2; ++k) { ab[k] /= abn; cd[k] /= cdn; ce[k] /= cdn; }
    const double h[8] = { ab[1] / abn, -ab[0] / abn, -ab[1] / abn, ab[0] / abn,
                          -cd[1] / cdn, cd[0] / cdn, cd[1] / cdn, -cd[0] / cdn };
    const double dangle = atan2(ab[0] * ce[1] - ab[1] * ce[0], dot<2>(ab, ce));
    const double w1 = q.w;
    const double w2 = q.w * dangle;
    // lower packed 8x8: -w1 h h' (dspr), then the derivative of the normals
    double m[36];
    for (int k = 0; k < 36; ++k) m[k] = 0.0;
    for (int j = 0, kk = 0; j < 8; kk += 8 - j, ++j) {
        if (h[j] != 0.0) {
            const double t = -w1 * h[j];
            for (int i = j, k = kk; i < 8; ++i, ++k) m[k] = m[k] + h[i] * t;
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
            m[i + sa] += sg * ea[i - j];
            m[i + sc] += sg * ec[i - j];
        }
        sa += 7 - j;
        sc += 3 - j;
    }
    for (int i = 0; i < 8; ++i) {
        o.S(index[i], w2 * h[i]);
        for (int j = 0; j < 8; ++j) {
            int k;
            if (j < i) {
                k = (int)(j * (7.5 - 0.5 * j) + i);
            } else {
                k = (int)(i * (7.5 - 0.5 * i) + j);
                o.C(index[i], index[j], m[k]);
            }
            o.S(index[i], -(m[k] * coord[j]));
        }
    }
}

// interSideClamp (meca_inter.cc:2200-2226), 2D (2032-2074) and 3D (2089-2186)
template <int D>
__device__ void side_clamp(const Geo<D>& G, const DeferredCall& q, Out& o)
{
    const double* g = q.g;
    const double w = q.w, len = q.len;
    double pa[D], df[D];
    G.pos(q.p[0], q.p[1], q.ca, pa);
    G.diff(q.p[0], q.p[1], df);
    if constexpr (D == 2) {
        const double gs[2] = { g[0] - pa[0], g[1] - pa[1] };
        const double arm = len * side_of(cross2(df, gs));
        const double aa = q.ca, aaw = w * aa;
        const double bb = 1.0 - q.ca, bbw = w * bb;
        const double ee = arm / norm<2>(df);
        const double eew = w * ee;
        const double ee2w = w * ee * ee;
        uint32_t i1 = q.p[0], i2 = q.p[1];
        o.B(i1, i1, -(bb * bbw + ee2w));
        o.B(i1, i2, -(aa * bbw - ee2w));
        o.B(i2, i2, -(aa * aaw + ee2w));
        i1 *= 2; i2 *= 2;
        o.C(i1, i2 + 1, eew);
        o.C(i1 + 1, i2, -eew);
        o.S(i1, bbw * g[0] - eew * g[1]);
        o.S(i1 + 1, bbw * g[1] + eew * g[0]);
        o.S(i2, aaw * g[0] + eew * g[1]);
        o.S(i2 + 1, aaw * g[1] - eew * g[0]);
    } else {
        const double ag[3] = { pa[0] - g[0], pa[1] - g[1], pa[2] - g[2] };
        double arm[3] = { ag[1] * df[2] - ag[2] * df[1], ag[2] * df[0] - ag[0] * df[2], ag[0] * df[1] - ag[1] * df[0] };
        const double n = norm<3>(arm);
        if (!(n > kEps)) return;
        const double f = len / n;
        for (int d = 0; d < 3; ++d) arm[d] = arm[d] * f;
        const double aa = 1.0 - q.ca, bb = q.ca;
        const double s = 1.0 / norm<3>(df);
        const double ex = s * arm[0], ey = s * arm[1], ez = s * arm[2];
        const uint32_t i1 = 3 * q.p[0], i2 = 3 * q.p[1];
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
        const double sca = 1.0 / norm<3>(arm);
        const double aan = aa * sca, bbn = bb * sca;
        const double TP[6] = { aan * ex, aan * ey, aan * ez, bbn * ex, bbn * ey, bbn * ez };
        // dsyrk('U','N', 6, 1, w, TP, 6, -w, TT, 6)
        for (int j = 0; j < 6; ++j) {
            for (int i = 0; i <= j; ++i) TT[i + 6 * j] = -w * TT[i + 6 * j];
            if (TP[j] != 0.0) {
                const double t = w * TP[j];
                for (int i = 0; i <= j; ++i) TT[i + 6 * j] = TT[i + 6 * j] + t * TP[i];
            }
        }
        for (int i = 0; i < 6; ++i)
            for (int j = i; j < 6; ++j) o.C(inx[i], inx[j], TT[i + 6 * j]);
        const double ws = dot<3>(arm, g) * sca * sca;
        const double gx = w * (g[0] - ws * arm[0]);
        const double gy = w * (g[1] - ws * arm[1]);
        const double gz = w * (g[2] - ws * arm[2]);
        for (int i = 0; i < 6; ++i) o.S(inx[i], T[i] * gx + T[i + 6] * gy + T[i + 12] * gz);
    }
}

// interSlidingClamp (meca_inter.cc:2241-2283)
template <int D>
__device__ void sliding_clamp(const Geo<D>& G, const DeferredCall& q, Out& o)
{
    const double w = q.w;
    const double* g = q.g;
    const double aa = q.ca, bb = 1.0 - q.ca;
    const double bbbb = bb * bb, aaaa = aa * aa, aabb = aa * bb;
    double dir[D];
    G.diff(q.p[0], q.p[1], dir);
    unit<D>(dir);
    const uint32_t i1 = D * q.p[0], i2 = D * q.p[1];
    for (int x = 0; x < D; ++x) {
        for (int y = x; y < D; ++y) {
            const double P = w * ((x == y ? 1.0 : 0.0) - dir[x] * dir[y]);
            o.C(i1 + x, i1 + y, -(bbbb * P));
            o.C(i2 + x, i2 + y, -(aaaa * P));
            o.C(i1 + x, i2 + y, -(aabb * P));
            if (x != y) o.C(i1 + y, i2 + x, -(aabb * P));
        }
    }
    const double pr = dot<D>(g, dir);
    for (int x = 0; x < D; ++x) {
        o.S(i1 + x, w * bb * (g[x] - pr * dir[x]));
        o.S(i2 + x, w * aa * (g[x] - pr * dir[x]));
    }
}

template <int D>
__device__ void evaluate(const Geo<D>& G, const DeferredCall& q, Out& o)
{
    switch (q.kind) {
    case kDefLongLinkEE: long_link_ee<D>(G, q, o); break;
    case kDefLongClamp: long_clamp<D>(G, q, o); break;
    case kDefLongLinkIE: {
        double pa[D], ab[D];
        G.pos(q.p[0], q.p[1], q.ca, pa);
        for (int d = 0; d < D; ++d) ab[d] = G.pt(q.p[2])[d] - pa[d];
        const uint32_t pts[3] = { q.p[0], q.p[1], q.p[2] };
        const double c[3] = { 1.0 - q.ca, q.ca, -1.0 };
        long_link_n<D, 3>(pts, c, ab, q.len, q.w, o);
        break;
    }
    case kDefLongLinkII: {
        double pa[D], pb[D], ab[D];
        G.pos(q.p[0], q.p[1], q.ca, pa);
        G.pos(q.p[2], q.p[3], q.cb, pb);
        for (int d = 0; d < D; ++d) ab[d] = pb[d] - pa[d];
        const double c[4] = { 1.0 - q.ca, q.ca, -(1.0 - q.cb), -q.cb };
        long_link_n<D, 4>(q.p, c, ab, q.len, q.w, o);
        break;
    }
    case kDefSideLinkIE: side_link<D, false>(G, q, o); break;
    case kDefSideLinkII: side_link<D, true>(G, q, o); break;
    case kDefSlidingLinkIE: sliding_link_ie<D>(G, q, o); break;
    case kDefSlidingLinkII: sliding_link_ii<D>(G, q, o); break;
    case kDefSideSlidingIE: side_sliding<D, false>(G, q, o); break;
    case kDefSideSlidingII: side_sliding<D, true>(G, q, o); break;
    case kDefSideClamp: side_clamp<D>(G, q, o); break;
    case kDefSlidingClamp: sliding_clamp<D>(G, q, o); break;
    case kDefSideSideLink:
        if constexpr (D == 2) side_side_link(G, q, o);
        break;
    case kDefTorque:
        if constexpr (D == 2) torque(G, q, o);
        break;
    default: break;
    }
}

template <int D>
__global__ void k_emit_deferred(uint32_t n, const DeferredCall* __restrict__ calls, const double* __restrict__ pos,
                                uint64_t* keyB, double* valB, uint64_t* keyC, double* valC, uint64_t* keyBas,
                                double* valBas)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const DeferredCall q = calls[t];
    uint32_t cap[3];
    deferred_slots(q.kind, D, cap);
    Out o{ { keyB + q.off[0], valB + q.off[0], 0, cap[0] },
           { keyC + q.off[1], valC + q.off[1], 0, cap[1] },
           { keyBas + q.off[2], valBas + q.off[2], 0, cap[2] } };
    const Geo<D> G{ pos };
    evaluate<D>(G, q, o);
    o.b.close();
    o.c.close();
    o.s.close();
}

}  // namespace

void launch_emit_deferred(Context& c, const EmitterMark& mk, uint64_t* keyB, double* valB, uint64_t* keyC,
                          double* valC, uint64_t* keyBas, double* valBas)
{
    if (!mk.count) return;
    // kind 4: calls recorded by the host; kind 5: made on the device by the steric grid
    const DeferredCall* calls = (mk.kind == 5 ? c.d_st_calls.get() : c.d_def.get()) + mk.first;
    const dim3 grid((mk.count + 127) / 128);
    if (c.dim == 2)
        hipLaunchKernelGGL(k_emit_deferred<2>, grid, dim3(128), 0, c.stream, mk.count, calls, c.d_pos.get(), keyB,
                           valB, keyC, valC, keyBas, valBas);
    else
        hipLaunchKernelGGL(k_emit_deferred<3>, grid, dim3(128), 0, c.stream, mk.count, calls, c.d_pos.get(), keyB,
                           valB, keyC, valC, keyBas, valBas);
    CYMG_HIP(hipGetLastError());
}

}  // namespace cymg
