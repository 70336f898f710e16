// Device-side numerics shared by every pgeof kernel on gfx950.
//
// Per neighbourhood: shifted first and second moments in FP64, a closed-form
// symmetric 3x3 eigensolve in FP64, and the feature formulas in FP32.
//
// Why FP64 for the moments and the eigensolve: on CDNA3/CDNA4 a wave64 FP64
// FMA issues at the same rate as a non-packed FP32 FMA, so carrying the
// covariance in FP64 costs register space but no extra VALU cycles, and it
// removes the FP32 cancellation that otherwise dominates lambda_min on flat or
// far-from-origin neighbourhoods.
//
// Reference semantics followed (pgeof repository):
//   include/pca.hpp:71-98   mean, centred covariance / k, ascending eigenpairs,
//                           clamp >= 0, flip v2 when v2.z < 0
//   include/pca.hpp:140-150 eigentropy (FP32)
//   include/pca.hpp:160-200 feature formulas (FP32, no FMA contraction)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pgeof_amd
{

constexpr float kEps = 1e-3f;  // pca.hpp:32-33 epsilon<float>

// --------------------------------------------------------------------------
// Shifted moments.  Coordinates enter as exact FP64 differences to a shift
// point (the row's first neighbour), so S1/k stays of the order of the
// neighbourhood extent and C = S2/k - m m^T loses nothing to cancellation.
// --------------------------------------------------------------------------
struct Moments
{
    double sx, sy, sz;
    double xx, xy, xz, yy, yz, zz;
};

__host__ __device__ __forceinline__ void mom_zero(Moments& m)
{
    m.sx = m.sy = m.sz = 0.0;
    m.xx = m.xy = m.xz = m.yy = m.yz = m.zz = 0.0;
}

__host__ __device__ __forceinline__ void mom_add(Moments& m, double dx, double dy, double dz)
{
    m.sx += dx;
    m.sy += dy;
    m.sz += dz;
    m.xx = fma(dx, dx, m.xx);
    m.xy = fma(dx, dy, m.xy);
    m.xz = fma(dx, dz, m.xz);
    m.yy = fma(dy, dy, m.yy);
    m.yz = fma(dy, dz, m.yz);
    m.zz = fma(dz, dz, m.zz);
}

// covariance packed as c00 c01 c02 c11 c12 c22
__host__ __device__ __forceinline__ void mom_cov(const Moments& m, double k, double c[6])
{
    const double inv = 1.0 / k;
    const double mx = m.sx * inv, my = m.sy * inv, mz = m.sz * inv;
    c[0] = fma(-mx, mx, m.xx * inv);
    c[1] = fma(-mx, my, m.xy * inv);
    c[2] = fma(-mx, mz, m.xz * inv);
    c[3] = fma(-my, my, m.yy * inv);
    c[4] = fma(-my, mz, m.yz * inv);
    c[5] = fma(-mz, mz, m.zz * inv);
    // a covariance has a non-negative diagonal; rounding may leave -0 ulp
    c[0] = fmax(c[0], 0.0);
    c[3] = fmax(c[3], 0.0);
    c[5] = fmax(c[5], 0.0);
}

// --------------------------------------------------------------------------
// Closed-form eigenvalues (Smith 1961, "Eigenvalues of a symmetric 3x3
// matrix", CACM 4(4):168).  Returns ascending l0 <= l1 <= l2 of the SCALED
// matrix a = c / scale.  iso = true when a is a multiple of the identity.
// --------------------------------------------------------------------------
struct EigVals
{
    double l0, l1, l2;
    double scale;
    bool iso;
};

__host__ __device__ __forceinline__ double max_abs6(const double c[6])
{
    double s = fabs(c[0]);
    s = fmax(s, fabs(c[1]));
    s = fmax(s, fabs(c[2]));
    s = fmax(s, fabs(c[3]));
    s = fmax(s, fabs(c[4]));
    s = fmax(s, fabs(c[5]));
    return s;
}

__host__ __device__ __forceinline__ EigVals eig_values(const double c[6], double a[6])
{
    EigVals ev;
    double s = max_abs6(c);
    ev.scale = s;
    if (!(s > 0.0))
    {
        // zero (or NaN) covariance
        for (int i = 0; i < 6; ++i) a[i] = 0.0;
        ev.l0 = ev.l1 = ev.l2 = 0.0;
        ev.scale = 0.0;
        ev.iso = true;
        return ev;
    }
    const double inv = 1.0 / s;
    for (int i = 0; i < 6; ++i) a[i] = c[i] * inv;
    const double q = (a[0] + a[3] + a[5]) * (1.0 / 3.0);
    const double b00 = a[0] - q, b11 = a[3] - q, b22 = a[5] - q;
    const double p1 = a[1] * a[1] + a[2] * a[2] + a[4] * a[4];
    const double p2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * p1;
    if (p2 < 1e-28)
    {
        ev.l0 = ev.l1 = ev.l2 = q;
        ev.iso = true;
        return ev;
    }
    ev.iso = false;
    const double p = sqrt(p2 * (1.0 / 6.0));
    const double ip = 1.0 / p;
    const double B00 = b00 * ip, B11 = b11 * ip, B22 = b22 * ip;
    const double B01 = a[1] * ip, B02 = a[2] * ip, B12 = a[4] * ip;
    const double det = B00 * (B11 * B22 - B12 * B12) - B01 * (B01 * B22 - B12 * B02) + B02 * (B01 * B12 - B11 * B02);
    double r = 0.5 * det;
    r = fmin(fmax(r, -1.0), 1.0);
    const double phi = acos(r) * (1.0 / 3.0);
    double sphi, cphi;
    sincos(phi, &sphi, &cphi);
    // cos(phi + 2 pi / 3) = -cos(phi)/2 - sqrt(3)/2 sin(phi)
    const double c2 = cphi;
    const double c0 = -0.5 * cphi - 0.86602540378443864676 * sphi;
    ev.l2 = q + 2.0 * p * c2;
    ev.l0 = q + 2.0 * p * c0;
    ev.l1 = 3.0 * q - ev.l0 - ev.l2;
    // keep the order exact under rounding
    if (ev.l1 > ev.l2) ev.l1 = ev.l2;
    if (ev.l1 < ev.l0) ev.l1 = ev.l0;
    return ev;
}

// Eigenvector of a (scaled) symmetric matrix for a simple eigenvalue lam:
// largest cross product of two rows of (a - lam I).
__host__ __device__ __forceinline__ void eigvec_cross(const double a[6], double lam, double v[3])
{
    const double r0x = a[0] - lam, r0y = a[1], r0z = a[2];
    const double r1x = a[1], r1y = a[3] - lam, r1z = a[4];
    const double r2x = a[2], r2y = a[4], r2z = a[5] - lam;
    const double ax = r0y * r1z - r0z * r1y, ay = r0z * r1x - r0x * r1z, az = r0x * r1y - r0y * r1x;
    const double bx = r0y * r2z - r0z * r2y, by = r0z * r2x - r0x * r2z, bz = r0x * r2y - r0y * r2x;
    const double cx = r1y * r2z - r1z * r2y, cy = r1z * r2x - r1x * r2z, cz = r1x * r2y - r1y * r2x;
    const double na = ax * ax + ay * ay + az * az;
    const double nb = bx * bx + by * by + bz * bz;
    const double nc = cx * cx + cy * cy + cz * cz;
    double x = ax, y = ay, z = az, n = na;
    if (nb > n) x = bx, y = by, z = bz, n = nb;
    if (nc > n) x = cx, y = cy, z = cz, n = nc;
    if (!(n > 0.0))
    {
        v[0] = 0.0, v[1] = 0.0, v[2] = 1.0;
        return;
    }
    const double in = 1.0 / sqrt(n);
    v[0] = x * in, v[1] = y * in, v[2] = z * in;
}

// Given a unit eigenvector u, the other two eigenvectors lie in the plane
// orthogonal to u: project a onto an orthonormal basis (e, f) of that plane
// and diagonalise the 2x2 result.  big/small receive the eigenvectors of the
// larger/smaller of the two remaining eigenvalues, lbig/lsmall the
// eigenvalues themselves (backward stable, also when they coincide).
__host__ __device__ __forceinline__ void eig_complete(const double a[6], const double u[3], double big[3], double small[3],
                                                      double& lbig, double& lsmall)
{
    // u x e_x, u x e_y or u x e_z for the smallest |u_i| (selects, not
    // branches, so that e stays in registers)
    const double ux = fabs(u[0]), uy = fabs(u[1]), uz = fabs(u[2]);
    const bool bx = ux <= uy && ux <= uz;
    const bool by = !bx && uy <= uz;
    double e[3];
    e[0] = bx ? 0.0 : (by ? -u[2] : u[1]);
    e[1] = bx ? u[2] : (by ? 0.0 : -u[0]);
    e[2] = bx ? -u[1] : (by ? u[0] : 0.0);
    const double ie = 1.0 / sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    e[0] *= ie, e[1] *= ie, e[2] *= ie;
    const double f[3] = {u[1] * e[2] - u[2] * e[1], u[2] * e[0] - u[0] * e[2], u[0] * e[1] - u[1] * e[0]};
    const double ae[3] = {a[0] * e[0] + a[1] * e[1] + a[2] * e[2], a[1] * e[0] + a[3] * e[1] + a[4] * e[2],
                          a[2] * e[0] + a[4] * e[1] + a[5] * e[2]};
    const double af[3] = {a[0] * f[0] + a[1] * f[1] + a[2] * f[2], a[1] * f[0] + a[3] * f[1] + a[4] * f[2],
                          a[2] * f[0] + a[4] * f[1] + a[5] * f[2]};
    const double m00 = e[0] * ae[0] + e[1] * ae[1] + e[2] * ae[2];
    const double m01 = e[0] * af[0] + e[1] * af[1] + e[2] * af[2];
    const double m11 = f[0] * af[0] + f[1] * af[1] + f[2] * af[2];
    const double t = 0.5 * (m00 - m11);
    const double root = sqrt(t * t + m01 * m01);
    const double mean = 0.5 * (m00 + m11);
    lbig = mean + root;
    lsmall = mean - root;
    double x, y;
    if (t >= 0.0)
        x = t + root, y = m01;
    else
        x = m01, y = root - t;
    const double nn = x * x + y * y;
    if (nn > 0.0)
    {
        const double in = 1.0 / sqrt(nn);
        x *= in, y *= in;
    }
    else
        x = 1.0, y = 0.0;
    for (int i = 0; i < 3; ++i)
    {
        big[i] = x * e[i] + y * f[i];
        small[i] = -y * e[i] + x * f[i];
    }
}

struct Eig
{
    double l[3];     // ascending, unscaled, NOT clamped
    double v[3][3];  // v[i] = unit eigenvector of l[i]
};

// Full symmetric eigendecomposition, eigenpairs ascending (Eigen's order).
// The trigonometric eigenvalues only select the best separated eigenvalue
// (gap >= half the spread) and its eigenvector; all three eigenvalues are then
// taken from that vector's Rayleigh quotient and from the 2x2 projection on
// its orthogonal plane.  The trigonometric values alone lose ~sqrt(eps) next
// to a repeated eigenvalue (acos near +-1), e.g. on collinear or planar
// neighbourhoods where lambda_min must come out as 0.
__host__ __device__ __forceinline__ Eig eig_full(const double c[6])
{
    double a[6];
    const EigVals ev = eig_values(c, a);
    // The separated eigenvalue is the largest (hi) or the smallest one. The
    // code below is branch-free (selects only) so that every array stays in
    // registers.
    const bool hi = ev.l2 - ev.l1 >= ev.l1 - ev.l0;
    double u[3], big[3], small[3], lb, ls;
    eigvec_cross(a, hi ? ev.l2 : ev.l0, u);
    eig_complete(a, u, big, small, lb, ls);
    // Rayleigh quotient u'Au of the separated eigenvector
    const double au0 = a[0] * u[0] + a[1] * u[1] + a[2] * u[2];
    const double au1 = a[1] * u[0] + a[3] * u[1] + a[4] * u[2];
    const double au2 = a[2] * u[0] + a[4] * u[1] + a[5] * u[2];
    const double lu = u[0] * au0 + u[1] * au1 + u[2] * au2;
    // hi: (l0, l1, l2) = (small, big, u); otherwise (u, small, big)
    double l[3] = {hi ? ls : lu, hi ? lb : ls, hi ? lu : lb};
    double v[3][3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
    {
        v[0][j] = hi ? small[j] : u[j];
        v[1][j] = hi ? big[j] : small[j];
        v[2][j] = hi ? u[j] : big[j];
    }
    // ascending order (a swap can only happen within rounding)
#pragma unroll
    for (int pass = 0; pass < 2; ++pass)
#pragma unroll
        for (int i = 0; i < 2; ++i)
        {
            const bool sw = l[i] > l[i + 1];
            const double tl = l[i];
            l[i] = sw ? l[i + 1] : l[i];
            l[i + 1] = sw ? tl : l[i + 1];
#pragma unroll
            for (int j = 0; j < 3; ++j)
            {
                const double tv = v[i][j];
                v[i][j] = sw ? v[i + 1][j] : v[i][j];
                v[i + 1][j] = sw ? tv : v[i + 1][j];
            }
        }
    // Eigen returns the identity basis for a multiple of the identity
    const double li[3] = {ev.l0, ev.l1, ev.l2};
    Eig E;
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        E.l[i] = (ev.iso ? li[i] : l[i]) * ev.scale;
#pragma unroll
        for (int j = 0; j < 3; ++j) E.v[i][j] = ev.iso ? (i == j ? 1.0 : 0.0) : v[i][j];
    }
    return E;
}

// Eigenvalues only (for eigentropy), unscaled, ascending.
__host__ __device__ __forceinline__ void eig_vals_only(const double c[6], double l[3])
{
    double a[6];
    const EigVals ev = eig_values(c, a);
    l[0] = ev.l0 * ev.scale;
    l[1] = ev.l1 * ev.scale;
    l[2] = ev.l2 * ev.scale;
}

// --------------------------------------------------------------------------
// FP32 epilogue: pca.hpp:85-96 clamp + orientation, pca.hpp:160-200 features.
// No FMA contraction, as the reference's x86-64 baseline build.
// --------------------------------------------------------------------------
struct Pca32
{
    float val[3];
    float v[3][3];
};

__host__ __device__ __forceinline__ Pca32 to_pca32(const Eig& E)
{
    Pca32 P;
    for (int i = 0; i < 3; ++i)
    {
        P.val[i] = fmaxf((float)E.l[i], 0.0f);
        for (int j = 0; j < 3; ++j) P.v[i][j] = (float)E.v[i][j];
    }
    // pca.hpp:96; for v2.z == 0 (solver sign in the reference, not pinned)
    // the first non-zero of (y, x) is made positive.
    bool flip = P.v[2][2] < 0.0f;
    if (P.v[2][2] == 0.0f) flip = (P.v[2][1] < 0.0f) || (P.v[2][1] == 0.0f && P.v[2][0] < 0.0f);
    if (flip)
        for (int j = 0; j < 3; ++j) P.v[2][j] = -P.v[2][j];
    return P;
}

__host__ __device__ __forceinline__ void features11(const Pca32& P, float f[11])
{
#pragma clang fp contract(off)
    const float val0 = sqrtf(P.val[0]);
    const float val1 = sqrtf(P.val[1]);
    const float val2 = sqrtf(P.val[2]);
    const float fact = 1.0f / (val0 + kEps);
    f[4] = P.v[2][0];
    f[5] = P.v[2][1];
    f[6] = P.v[2][2];
    f[0] = (val0 - val1) * fact;
    f[1] = (val1 - val2) * fact;
    f[2] = val2 * fact;
    f[7] = val0;
    f[8] = sqrtf(val0 * val1 + 1e-6f);
    f[9] = powf(val0 * val1 * val2 + 1e-9f, 1.0f / 3.0f);
    f[10] = val2 / (val0 + val1 + val2 + kEps);
    f[3] = 0.0f;
    if (val0 > 0.0f)
    {
        float u[3];
        for (int c = 0; c < 3; ++c)
            u[c] = P.val[0] * fabsf(P.v[0][c]) + P.val[1] * fabsf(P.v[1][c]) + P.val[2] * fabsf(P.v[2][c]);
        f[3] = u[2] / sqrtf(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    }
}

// pca.hpp:140-150 on FP32-rounded clamped eigenvalues
__host__ __device__ __forceinline__ float eigentropy32(const double l[3])
{
#pragma clang fp contract(off)
    const float v0 = fmaxf((float)l[0], 0.0f), v1 = fmaxf((float)l[1], 0.0f), v2 = fmaxf((float)l[2], 0.0f);
    const float sum = v0 + v1 + v2 + kEps;
    const float e0 = v0 / sum, e1 = v1 / sum, e2 = v2 / sum;
    return -e0 * logf(e0 + kEps) - e1 * logf(e1 + kEps) - e2 * logf(e2 + kEps);
}

// nanoflann metric_L2_Simple in float: ((dx*dx + dy*dy) + dz*dz), dx = q - p,
// every product and sum rounded separately.
__host__ __device__ __forceinline__ float sqdist32(float qx, float qy, float qz, float px, float py, float pz)
{
#pragma clang fp contract(off)
    const float dx = qx - px, dy = qy - py, dz = qz - pz;
    float d = dx * dx;
    d = d + dy * dy;
    d = d + dz * dz;
    return d;
}

// The 11 features of a row in three or four stores instead of eleven: two
// 16-byte stores at the row's 16-byte boundaries, the entries before and
// after them in one 4-, 8- or 12-byte store each.  (Rows are 44 bytes, so a
// row starts 0, 4, 8 or 12 bytes past a boundary.  The rows of a wave's
// lanes are far apart: every store reaches one cache line per lane.)
__device__ __forceinline__ void store_row11(float* __restrict__ o, const float f[11])
{
    switch ((4u - ((uint32_t)(reinterpret_cast<uintptr_t>(o) >> 2) & 3u)) & 3u)
    {
        case 0:
            *reinterpret_cast<float4*>(o) = make_float4(f[0], f[1], f[2], f[3]);
            *reinterpret_cast<float4*>(o + 4) = make_float4(f[4], f[5], f[6], f[7]);
            *reinterpret_cast<float3*>(o + 8) = make_float3(f[8], f[9], f[10]);
            break;
        case 1:
            o[0] = f[0];
            *reinterpret_cast<float4*>(o + 1) = make_float4(f[1], f[2], f[3], f[4]);
            *reinterpret_cast<float4*>(o + 5) = make_float4(f[5], f[6], f[7], f[8]);
            *reinterpret_cast<float2*>(o + 9) = make_float2(f[9], f[10]);
            break;
        case 2:
            *reinterpret_cast<float2*>(o) = make_float2(f[0], f[1]);
            *reinterpret_cast<float4*>(o + 2) = make_float4(f[2], f[3], f[4], f[5]);
            *reinterpret_cast<float4*>(o + 6) = make_float4(f[6], f[7], f[8], f[9]);
            o[10] = f[10];
            break;
        default:
            *reinterpret_cast<float3*>(o) = make_float3(f[0], f[1], f[2]);
            *reinterpret_cast<float4*>(o + 3) = make_float4(f[3], f[4], f[5], f[6]);
            *reinterpret_cast<float4*>(o + 7) = make_float4(f[7], f[8], f[9], f[10]);
            break;
    }
}

// --------------------------------------------------------------------------
// Launch helpers
// --------------------------------------------------------------------------
// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X.  This
// maps hardware block b to a logical block so that consecutive logical blocks
// (spatially adjacent queries) run on the same XCD and share its L2.
__device__ __forceinline__ uint32_t xcd_logical_block(uint32_t b, uint32_t nblocks)
{
    constexpr uint32_t kXcd = 8;
    const uint32_t q = nblocks / kXcd, r = nblocks % kXcd;
    const uint32_t x = b % kXcd, s = b / kXcd;
    return x * q + (x < r ? x : r) + s;
}

}  // namespace pgeof_amd
