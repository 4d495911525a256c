// Robust relative orientation from putative matches (host code, double precision): the
// consumer of the match lists in the reference's SfM drivers
// (computeRobustOrientationMLE, V3D/Apps/SfM/reconstruction_common.cpp:139-340).
//
// Five-point solver: the ten cubic constraints on E = x X + y Y + z Z + W are reduced by
// Gauss-Jordan elimination over the twenty monomials of degree <= 3 in graded reverse
// lexicographic order; the ten cubic monomials are then linear in the ten lower ones
// {x^2, xy, xz, y^2, yz, z^2, x, y, z, 1}, which gives the 10x10 matrix of multiplication
// by x on that basis.  Its real eigenvalues are the x of the solutions and its eigenvectors
// the monomial vectors (Stewenius, Engels, Nister, "Recent developments on direct relative
// orientation", 2006).  Eigenvalues: Householder reduction to Hessenberg form and the
// Francis double-shift QR iteration; eigenvectors: inverse iteration.
#include "../../include/v3dm_ransac.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/v3dm.h"

namespace {

struct Mat3 {
    double m[9];
    double& operator()(int r, int c) { return m[3 * r + c]; }
    double operator()(int r, int c) const { return m[3 * r + c]; }
};

Mat3 zero3()
{
    Mat3 a;
    std::memset(a.m, 0, sizeof(a.m));
    return a;
}

Mat3 load3(const double* p)
{
    Mat3 a;
    std::memcpy(a.m, p, sizeof(a.m));
    return a;
}

Mat3 mul3(Mat3 const& a, Mat3 const& b)
{
    Mat3 c = zero3();
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k)
            for (int s = 0; s < 3; ++s) c(r, s) += a(r, k) * b(k, s);
    return c;
}

Mat3 transpose3(Mat3 const& a)
{
    Mat3 t;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) t(c, r) = a(r, c);
    return t;
}

// The adjugate scaled by 1 / det, det expanded as a e i + b f g + c d h - c e g - b d i -
// a f h (invertedMatrix, V3D/Math/v3d_linear.h:558-575, in its operation order)
bool invert3(Mat3 const& A, Mat3& out)
{
    double const a = A(0, 0), b = A(0, 1), c = A(0, 2);
    double const d = A(1, 0), e = A(1, 1), f = A(1, 2);
    double const g = A(2, 0), h = A(2, 1), i = A(2, 2);
    double const det = a * e * i + b * f * g + c * d * h - c * e * g - b * d * i - a * f * h;
    if (!(std::fabs(det) > 0.0)) return false;
    double const s = 1.0 / det;
    out(0, 0) = (e * i - f * h) * s, out(0, 1) = (c * h - b * i) * s, out(0, 2) = (b * f - c * e) * s;
    out(1, 0) = (f * g - d * i) * s, out(1, 1) = (a * i - c * g) * s, out(1, 2) = (c * d - a * f) * s;
    out(2, 0) = (d * h - e * g) * s, out(2, 1) = (b * g - a * h) * s, out(2, 2) = (a * e - b * d) * s;
    return true;
}

// K^-1 (u, v, 1), divided by its third component, each sum started from the last column
// (multiply_A_v_projective, V3D/Math/v3d_linearbase.h:1030-1046); Acc = double, or float
// where the reference stores the result in a float vector (its PointMeasurement positions)
template <typename Acc>
void normalise_point_t(Mat3 const& iK, double u, double v, double& x, double& y)
{
    Acc w = (Acc)iK(2, 2);
    w = (Acc)(w + iK(2, 0) * u);
    w = (Acc)(w + iK(2, 1) * v);
    Acc d0 = (Acc)iK(0, 2), d1 = (Acc)iK(1, 2);
    d0 = (Acc)(d0 + iK(0, 0) * u);
    d0 = (Acc)(d0 + iK(0, 1) * v);
    d1 = (Acc)(d1 + iK(1, 0) * u);
    d1 = (Acc)(d1 + iK(1, 1) * v);
    d0 /= w;
    d1 /= w;
    x = d0;
    y = d1;
}

void normalise_point(Mat3 const& iK, double u, double v, double& x, double& y)
{
    normalise_point_t<double>(iK, u, v, x, y);
}

// (A (x, y, 1))_r summed from the last column on (multiply_A_v_affine, v3d_linearbase.h:
// 1048-1061); At: the same with A transposed
double affine_row(Mat3 const& A, int r, double x, double y) { return A(r, 2) + A(r, 0) * x + A(r, 1) * y; }
double affine_col(Mat3 const& A, int c, double x, double y) { return A(2, c) + A(0, c) * x + A(1, c) * y; }

// sampsonEpipolarError (reconstruction_common.h:588-608), same operations in the same order
double sampson(Mat3 const& F, double px, double py, double qx, double qy)
{
    double const fp0 = affine_row(F, 0, px, py), fp1 = affine_row(F, 1, px, py), fp2 = affine_row(F, 2, px, py);
    double const ftq0 = affine_col(F, 0, qx, qy), ftq1 = affine_col(F, 1, qx, qy);
    double const tmp = qx * fp0 + qy * fp1 + fp2;
    double const num = tmp * tmp;
    double const denom = fp0 * fp0 + fp1 * fp1 + ftq0 * ftq0 + ftq1 * ftq1;
    return num / denom;
}

// [T]x R
Mat3 essential_from_pose(Mat3 const& R, double const T[3])
{
    Mat3 tx = zero3();
    tx(0, 1) = -T[2];
    tx(0, 2) = T[1];
    tx(1, 0) = T[2];
    tx(1, 2) = -T[0];
    tx(2, 0) = -T[1];
    tx(2, 1) = T[0];
    return mul3(tx, R);
}

// ---------------------------------------------------------------------------- polynomials
constexpr int kMon = 20;
// (x, y, z) exponents, graded reverse lexicographic: the ten cubics, then the basis
constexpr int kExp[kMon][3] = {{3, 0, 0}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {1, 1, 1}, {1, 0, 2}, {0, 3, 0},
                               {0, 2, 1}, {0, 1, 2}, {0, 0, 3}, {2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0},
                               {0, 1, 1}, {0, 0, 2}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, 0}};

int mon_index(int a, int b, int c)
{
    for (int i = 0; i < kMon; ++i)
        if (kExp[i][0] == a && kExp[i][1] == b && kExp[i][2] == c) return i;
    return -1;
}

struct Poly {
    double c[kMon];
};

Poly poly_zero()
{
    Poly p;
    std::memset(p.c, 0, sizeof(p.c));
    return p;
}

// monomial of the product of monomials i and j (-1 above degree 3)
struct ProductTable {
    int t[kMon][kMon];
    ProductTable()
    {
        for (int i = 0; i < kMon; ++i)
            for (int j = 0; j < kMon; ++j) {
                int const a = kExp[i][0] + kExp[j][0], b = kExp[i][1] + kExp[j][1], c = kExp[i][2] + kExp[j][2];
                t[i][j] = (a + b + c <= 3) ? mon_index(a, b, c) : -1;
            }
    }
};

// product of two polynomials whose degrees add up to at most 3
Poly poly_mul(Poly const& p, Poly const& q)
{
    static ProductTable const tab;  // initialised once, thread-safe
    auto const& table = tab.t;
    Poly r = poly_zero();
    for (int i = 0; i < kMon; ++i) {
        if (p.c[i] == 0.0) continue;
        for (int j = 0; j < kMon; ++j) {
            if (q.c[j] == 0.0) continue;
            int const k = table[i][j];
            if (k >= 0) r.c[k] += p.c[i] * q.c[j];
        }
    }
    return r;
}

Poly poly_add(Poly const& a, Poly const& b, double sb = 1.0)
{
    Poly r;
    for (int i = 0; i < kMon; ++i) r.c[i] = a.c[i] + sb * b.c[i];
    return r;
}

// ---------------------------------------------------------------------------- eigenvalues
constexpr int kN = 10;

// Householder reduction of a (n x n, row-major) to upper Hessenberg form, in place
void hessenberg(double (&a)[kN][kN])
{
    for (int k = 0; k < kN - 2; ++k) {
        double v[kN];
        double sigma = 0.0;
        for (int i = k + 1; i < kN; ++i) sigma += a[i][k] * a[i][k];
        sigma = std::sqrt(sigma);
        if (sigma == 0.0) continue;
        int const m = kN - k - 1;
        for (int i = 0; i < m; ++i) v[i] = a[k + 1 + i][k];
        v[0] += (v[0] >= 0.0 ? sigma : -sigma);
        double beta = 0.0;
        for (int i = 0; i < m; ++i) beta += v[i] * v[i];
        if (beta == 0.0) continue;
        for (int j = 0; j < kN; ++j) {  // left: (I - 2 v v^T / beta) a
            double s = 0.0;
            for (int i = 0; i < m; ++i) s += v[i] * a[k + 1 + i][j];
            s *= 2.0 / beta;
            for (int i = 0; i < m; ++i) a[k + 1 + i][j] -= s * v[i];
        }
        for (int i = 0; i < kN; ++i) {  // right: a (I - 2 v v^T / beta)
            double s = 0.0;
            for (int j = 0; j < m; ++j) s += a[i][k + 1 + j] * v[j];
            s *= 2.0 / beta;
            for (int j = 0; j < m; ++j) a[i][k + 1 + j] -= s * v[j];
        }
        for (int i = k + 2; i < kN; ++i) a[i][k] = 0.0;
    }
}

double sign_of(double a, double b) { return b >= 0.0 ? std::fabs(a) : -std::fabs(a); }

// Eigenvalues of an upper Hessenberg matrix (destroyed): Francis double-shift QR with
// deflation on small subdiagonal entries and exceptional shifts after 10 and 20 stalled
// iterations.  Returns false when an eigenvalue takes more than 60 iterations.
bool hessenberg_eigenvalues(double (&a)[kN][kN], double (&wr)[kN], double (&wi)[kN])
{
    double anorm = 0.0;
    for (int i = 0; i < kN; ++i)
        for (int j = std::max(i - 1, 0); j < kN; ++j) anorm += std::fabs(a[i][j]);
    int nn = kN - 1;
    double t = 0.0;
    while (nn >= 0) {
        int its = 0, l;
        do {
            for (l = nn; l >= 1; --l) {
                double s = std::fabs(a[l - 1][l - 1]) + std::fabs(a[l][l]);
                if (s == 0.0) s = anorm;
                if (std::fabs(a[l][l - 1]) + s == s) {
                    a[l][l - 1] = 0.0;
                    break;
                }
            }
            double x = a[nn][nn];
            if (l == nn) {  // one eigenvalue
                wr[nn] = x + t;
                wi[nn--] = 0.0;
            } else {
                double y = a[nn - 1][nn - 1];
                double w = a[nn][nn - 1] * a[nn - 1][nn];
                if (l == nn - 1) {  // a pair
                    double const p = 0.5 * (y - x);
                    double const q = p * p + w;
                    double z = std::sqrt(std::fabs(q));
                    x += t;
                    if (q >= 0.0) {
                        z = p + sign_of(z, p);
                        wr[nn - 1] = wr[nn] = x + z;
                        if (z != 0.0) wr[nn] = x - w / z;
                        wi[nn - 1] = wi[nn] = 0.0;
                    } else {
                        wr[nn - 1] = wr[nn] = x + p;
                        wi[nn - 1] = -z;
                        wi[nn] = z;
                    }
                    nn -= 2;
                } else {  // no eigenvalue yet: one double-shift step on rows / columns l .. nn
                    if (its == 60) return false;
                    if (its == 10 || its == 20) {
                        t += x;
                        for (int i = 0; i <= nn; ++i) a[i][i] -= x;
                        double const s = std::fabs(a[nn][nn - 1]) + std::fabs(a[nn - 1][nn - 2]);
                        y = x = 0.75 * s;
                        w = -0.4375 * s * s;
                    }
                    ++its;
                    int m;
                    double p = 0.0, q = 0.0, r = 0.0, z;
                    for (m = nn - 2; m >= l; --m) {
                        z = a[m][m];
                        r = x - z;
                        double s = y - z;
                        p = (r * s - w) / a[m + 1][m] + a[m][m + 1];
                        q = a[m + 1][m + 1] - z - r - s;
                        r = a[m + 2][m + 1];
                        s = std::fabs(p) + std::fabs(q) + std::fabs(r);
                        p /= s;
                        q /= s;
                        r /= s;
                        if (m == l) break;
                        double const u = std::fabs(a[m][m - 1]) * (std::fabs(q) + std::fabs(r));
                        double const v =
                            std::fabs(p) * (std::fabs(a[m - 1][m - 1]) + std::fabs(z) + std::fabs(a[m + 1][m + 1]));
                        if (u + v == v) break;
                    }
                    for (int i = m + 2; i <= nn; ++i) {
                        a[i][i - 2] = 0.0;
                        if (i != m + 2) a[i][i - 3] = 0.0;
                    }
                    for (int k = m; k <= nn - 1; ++k) {
                        if (k != m) {
                            p = a[k][k - 1];
                            q = a[k + 1][k - 1];
                            r = 0.0;
                            if (k != nn - 1) r = a[k + 2][k - 1];
                            if ((x = std::fabs(p) + std::fabs(q) + std::fabs(r)) != 0.0) {
                                p /= x;
                                q /= x;
                                r /= x;
                            }
                        }
                        double const s = sign_of(std::sqrt(p * p + q * q + r * r), p);
                        if (s != 0.0) {
                            if (k == m) {
                                if (l != m) a[k][k - 1] = -a[k][k - 1];
                            } else {
                                a[k][k - 1] = -s * x;
                            }
                            p += s;
                            x = p / s;
                            y = q / s;
                            z = r / s;
                            q /= p;
                            r /= p;
                            for (int j = k; j <= nn; ++j) {
                                double pp = a[k][j] + q * a[k + 1][j];
                                if (k != nn - 1) {
                                    pp += r * a[k + 2][j];
                                    a[k + 2][j] -= pp * z;
                                }
                                a[k + 1][j] -= pp * y;
                                a[k][j] -= pp * x;
                            }
                            int const mmin = nn < k + 3 ? nn : k + 3;
                            for (int i = l; i <= mmin; ++i) {
                                double pp = x * a[i][k] + y * a[i][k + 1];
                                if (k != nn - 1) {
                                    pp += z * a[i][k + 2];
                                    a[i][k + 2] -= pp * r;
                                }
                                a[i][k + 1] -= pp * q;
                                a[i][k] -= pp;
                            }
                        }
                    }
                }
            }
        } while (l < nn - 1);
    }
    return true;
}

// Eigenvector of m for the real eigenvalue lambda: inverse iteration with the shift
// slightly off lambda (Gaussian elimination with partial pivoting).
bool eigenvector(double const (&m)[kN][kN], double lambda, double (&v)[kN])
{
    double scale = 0.0;
    for (int i = 0; i < kN; ++i)
        for (int j = 0; j < kN; ++j) scale = std::max(scale, std::fabs(m[i][j]));
    double const mu = lambda + 1e-10 * (scale + std::fabs(lambda));
    double lu[kN][kN];
    int piv[kN];
    for (int i = 0; i < kN; ++i) {
        for (int j = 0; j < kN; ++j) lu[i][j] = m[i][j];
        lu[i][i] -= mu;
        piv[i] = i;
    }
    double const tiny = 1e-300 + 1e-15 * scale;
    for (int k = 0; k < kN; ++k) {
        int p = k;
        for (int i = k + 1; i < kN; ++i)
            if (std::fabs(lu[i][k]) > std::fabs(lu[p][k])) p = i;
        if (p != k) {
            for (int j = 0; j < kN; ++j) std::swap(lu[k][j], lu[p][j]);
            std::swap(piv[k], piv[p]);
        }
        if (std::fabs(lu[k][k]) < tiny) lu[k][k] = tiny;
        for (int i = k + 1; i < kN; ++i) {
            double const f = lu[i][k] / lu[k][k];
            lu[i][k] = f;
            for (int j = k + 1; j < kN; ++j) lu[i][j] -= f * lu[k][j];
        }
    }
    for (int i = 0; i < kN; ++i) v[i] = 1.0;
    for (int it = 0; it < 3; ++it) {
        double b[kN];
        for (int i = 0; i < kN; ++i) b[i] = v[piv[i]];
        for (int i = 0; i < kN; ++i)
            for (int j = 0; j < i; ++j) b[i] -= lu[i][j] * b[j];
        for (int i = kN - 1; i >= 0; --i) {
            for (int j = i + 1; j < kN; ++j) b[i] -= lu[i][j] * b[j];
            b[i] /= lu[i][i];
        }
        double nrm = 0.0;
        for (int i = 0; i < kN; ++i) nrm += b[i] * b[i];
        nrm = std::sqrt(nrm);
        if (!(nrm > 0.0) || !std::isfinite(nrm)) return false;
        for (int i = 0; i < kN; ++i) v[i] = b[i] / nrm;
    }
    return true;
}

// ---------------------------------------------------------------------------- five points
// Null space of the 5 x 9 epipolar constraint matrix: Householder QR of its transpose; the
// last four columns of the orthogonal factor.
void null_space(double const x1[5], double const y1[5], double const x2[5], double const y2[5], double (&basis)[4][9])
{
    double a[9][5];
    for (int i = 0; i < 5; ++i) {
        double const r[9] = {x2[i] * x1[i], x2[i] * y1[i], x2[i], y2[i] * x1[i], y2[i] * y1[i], y2[i],
                             x1[i],         y1[i],         1.0};
        for (int j = 0; j < 9; ++j) a[j][i] = r[j];
    }
    double q[9][9];
    for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) q[i][j] = (i == j) ? 1.0 : 0.0;
    for (int k = 0; k < 5; ++k) {
        double sigma = 0.0;
        for (int i = k; i < 9; ++i) sigma += a[i][k] * a[i][k];
        sigma = std::sqrt(sigma);
        if (sigma == 0.0) continue;
        double v[9] = {0};
        for (int i = k; i < 9; ++i) v[i] = a[i][k];
        v[k] += (v[k] >= 0.0 ? sigma : -sigma);
        double beta = 0.0;
        for (int i = k; i < 9; ++i) beta += v[i] * v[i];
        if (beta == 0.0) continue;
        for (int j = 0; j < 5; ++j) {  // a <- H a
            double s = 0.0;
            for (int i = k; i < 9; ++i) s += v[i] * a[i][j];
            s *= 2.0 / beta;
            for (int i = k; i < 9; ++i) a[i][j] -= s * v[i];
        }
        for (int i = 0; i < 9; ++i) {  // q <- q H
            double s = 0.0;
            for (int j = k; j < 9; ++j) s += q[i][j] * v[j];
            s *= 2.0 / beta;
            for (int j = k; j < 9; ++j) q[i][j] -= s * v[j];
        }
    }
    for (int b = 0; b < 4; ++b)
        for (int i = 0; i < 9; ++i) basis[b][i] = q[i][5 + b];
}

int five_point(double const x1[5], double const y1[5], double const x2[5], double const y2[5], double* Es)
{
    double basis[4][9];
    null_space(x1, y1, x2, y2, basis);
    // E entries as polynomials x X + y Y + z Z + W
    Poly E[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            Poly p = poly_zero();
            p.c[16] = basis[0][3 * r + c];
            p.c[17] = basis[1][3 * r + c];
            p.c[18] = basis[2][3 * r + c];
            p.c[19] = basis[3][3 * r + c];
            E[r][c] = p;
        }
    double A[10][kMon];
    // 2 E E^T E - tr(E E^T) E = 0
    Poly EEt[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Poly s = poly_zero();
            for (int k = 0; k < 3; ++k) s = poly_add(s, poly_mul(E[i][k], E[j][k]));
            EEt[i][j] = s;
        }
    Poly const tr = poly_add(poly_add(EEt[0][0], EEt[1][1]), EEt[2][2]);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Poly s = poly_zero();
            for (int k = 0; k < 3; ++k) s = poly_add(s, poly_mul(EEt[i][k], E[k][j]));
            Poly const row = poly_add(poly_add(s, s), poly_mul(tr, E[i][j]), -1.0);
            std::memcpy(A[3 * i + j], row.c, sizeof(row.c));
        }
    // det E = 0
    Poly const m0 = poly_add(poly_mul(E[1][1], E[2][2]), poly_mul(E[1][2], E[2][1]), -1.0);
    Poly const m1 = poly_add(poly_mul(E[1][0], E[2][2]), poly_mul(E[1][2], E[2][0]), -1.0);
    Poly const m2 = poly_add(poly_mul(E[1][0], E[2][1]), poly_mul(E[1][1], E[2][0]), -1.0);
    Poly const det = poly_add(poly_add(poly_mul(E[0][0], m0), poly_mul(E[0][1], m1), -1.0), poly_mul(E[0][2], m2));
    std::memcpy(A[9], det.c, sizeof(det.c));
    // Gauss-Jordan elimination of the ten cubic monomials (partial pivoting)
    double scale = 0.0;
    for (int i = 0; i < 10; ++i)
        for (int j = 0; j < kMon; ++j) scale = std::max(scale, std::fabs(A[i][j]));
    if (!(scale > 0.0)) return 0;
    for (int c = 0; c < 10; ++c) {
        int p = c;
        for (int r = c + 1; r < 10; ++r)
            if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
        if (std::fabs(A[p][c]) < 1e-13 * scale) return 0;  // degenerate sample
        if (p != c)
            for (int j = 0; j < kMon; ++j) std::swap(A[p][j], A[c][j]);
        double const inv = 1.0 / A[c][c];
        for (int j = 0; j < kMon; ++j) A[c][j] *= inv;
        for (int r = 0; r < 10; ++r) {
            if (r == c || A[r][c] == 0.0) continue;
            double const f = A[r][c];
            for (int j = 0; j < kMon; ++j) A[r][j] -= f * A[c][j];
        }
    }
    // action matrix of x on the basis (x^2, xy, xz, y^2, yz, z^2, x, y, z, 1): x times the
    // first six basis monomials are the cubics x^3, x^2y, x^2z, xy^2, xyz, xz^2 (rows 0-5 of
    // the eliminated system), x times x, y, z, 1 are basis monomials
    double M[kN][kN];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < kN; ++j) M[i][j] = -A[i][10 + j];
    for (int i = 6; i < kN; ++i)
        for (int j = 0; j < kN; ++j) M[i][j] = 0.0;
    M[6][0] = 1.0;  // x * x = x^2
    M[7][1] = 1.0;  // x * y = xy
    M[8][2] = 1.0;  // x * z = xz
    M[9][6] = 1.0;  // x * 1 = x
    double H[kN][kN];
    std::memcpy(H, M, sizeof(H));
    hessenberg(H);
    double wr[kN], wi[kN];
    if (!hessenberg_eigenvalues(H, wr, wi)) return 0;
    int n = 0;
    for (int k = 0; k < kN; ++k) {
        if (wi[k] != 0.0 || !std::isfinite(wr[k])) continue;
        double v[kN];
        if (!eigenvector(M, wr[k], v)) continue;
        if (std::fabs(v[9]) < 1e-12) continue;  // a solution at infinity
        double const x = v[6] / v[9], y = v[7] / v[9], z = v[8] / v[9];
        if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z)) continue;
        for (int j = 0; j < 9; ++j) Es[9 * n + j] = x * basis[0][j] + y * basis[1][j] + z * basis[2][j] + basis[3][j];
        ++n;
    }
    return n;
}

// ---------------------------------------------------------------------------- pose from E
// Nister's procedure ("An efficient solution to the five-point relative pose problem",
// PAMI 2004, section 3.3), in the reference's arithmetic order so that every per-point
// decision is the reference's (computeEssentialSVD, getExactConfiguration,
// relativePoseFromEssential, triangulatePointExact: v3d_poseutilities.cpp:1108-1291).

void cross3(double const a[3], double const b[3], double (&c)[3])
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

double norm3(double const a[3]) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

void unit3(double (&a)[3])
{
    double const n = norm3(a);
    for (int i = 0; i < 3; ++i) a[i] /= n;
}

void mul3v(Mat3 const& E, double const v[3], double (&out)[3])
{
    for (int r = 0; r < 3; ++r) out[r] = E(r, 0) * v[0] + E(r, 1) * v[1] + E(r, 2) * v[2];
}

// E = U diag(s, s, 0) V^T without an iterative SVD: the third right singular vector is the
// longest cross product of two rows of E (the null direction), the first is the first row
// of that pair, the second completes the frame; U's columns are E v / |E v| and their cross
// product.  false for a rank < 2 matrix (no cross product of non-zero length).
bool essential_svd(Mat3 const& E, Mat3& U, Mat3& V)
{
    double rows[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) rows[r][c] = E(r, c);
    double x01[3], x02[3], x12[3];
    cross3(rows[0], rows[1], x01);
    cross3(rows[0], rows[2], x02);
    cross3(rows[1], rows[2], x12);
    double const* x = x01;
    double const* y = rows[0];
    double maxLen = norm3(x01);
    double len = norm3(x02);
    if (len > maxLen) maxLen = len, x = x02, y = rows[0];
    len = norm3(x12);
    if (len > maxLen) maxLen = len, x = x12, y = rows[1];
    if (!(maxLen > 0.0)) return false;
    double const inv = 1.0 / maxLen;
    double vc[3] = {inv * x[0], inv * x[1], inv * x[2]};
    double va[3] = {y[0], y[1], y[2]};
    unit3(va);
    double vb[3];
    cross3(vc, va, vb);
    double ua[3], ub[3], uc[3];
    mul3v(E, va, ua);
    mul3v(E, vb, ub);
    unit3(ua);
    unit3(ub);
    cross3(ua, ub, uc);
    for (int i = 0; i < 3; ++i) {
        V(i, 0) = va[i], V(i, 1) = vb[i], V(i, 2) = vc[i];
        U(i, 0) = ua[i], U(i, 1) = ub[i], U(i, 2) = uc[i];
    }
    return true;
}

Mat3 twist_d()
{
    Mat3 D = zero3();
    D(0, 1) = 1.0;
    D(1, 0) = -1.0;
    D(2, 2) = 1.0;
    return D;
}

// the four decompositions: 0 (R1, t), 1 (R1, -t), 2 (R2, t), 3 (R2, -t), R1 = U D V^T,
// R2 = U D^T V^T, t = the third column of U, normalised
void pose_of(Mat3 const& U, Mat3 const& V, int conf, Mat3& R, double (&T)[3])
{
    Mat3 const D = twist_d();
    R = mul3(mul3(U, conf < 2 ? D : transpose3(D)), transpose3(V));
    for (int i = 0; i < 3; ++i) T[i] = (conf == 0 || conf == 2) ? U(i, 2) : -1.0 * U(i, 2);
    unit3(T);
}

// The decomposition a correspondence p = (x1, y1, 1) <-> q = (x2, y2, 1) selects: the point
// triangulated exactly under P_A = [R1 | t] (the plane through the camera-2 ray and the
// baseline, C = P_A^T (q x diag(1, 1, 0) E p), meets the camera-1 ray at (C_4 p, -p.C_123));
// in front of both cameras -> (R1, t), behind both -> (R1, -t); else the twisted pair
// H_t = [I 0; -2 v_c^T -1] decides between (R2, t) and (R2, -t).
int configuration(Mat3 const& E, Mat3 const& U, Mat3 const& V, double x1, double y1, double x2, double y2)
{
    double const p[3] = {x1, y1, 1.0}, q[3] = {x2, y2, 1.0};
    Mat3 const R1 = mul3(mul3(U, twist_d()), transpose3(V));
    double const t1[3] = {U(0, 2), U(1, 2), U(2, 2)};
    double Pa[3][4];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) Pa[r][c] = R1(r, c);
        Pa[r][3] = t1[r];
    }
    double const p1[3] = {E(0, 0) * p[0] + E(0, 1) * p[1] + E(0, 2) * p[2],
                          E(1, 0) * p[0] + E(1, 1) * p[1] + E(1, 2) * p[2], 0.0};
    double cpl[3];
    cross3(q, p1, cpl);
    double C[4];
    for (int j = 0; j < 4; ++j) C[j] = Pa[0][j] * cpl[0] + Pa[1][j] * cpl[1] + Pa[2][j] * cpl[2];
    double Q[4] = {p[0] * C[3], p[1] * C[3], p[2] * C[3], -(p[0] * C[0] + p[1] * C[1] + p[2] * C[2])};
    Q[0] /= Q[3];
    Q[1] /= Q[3];
    Q[2] /= Q[3];
    Q[3] = 1.0;
    double const c1 = Q[2] * Q[3];
    double const c = Pa[2][0] * Q[0] + Pa[2][1] * Q[1] + Pa[2][2] * Q[2] + Pa[2][3] * Q[3];
    double const c2 = c * Q[3];
    if (c1 >= 0 && c2 >= 0) return 0;
    if (c1 < 0 && c2 < 0) return 1;
    double const ct = -(2 * V(0, 2) * Q[0] + 2 * V(1, 2) * Q[1] + 2 * V(2, 2) * Q[2] + Q[3]);
    return Q[2] * ct >= 0 ? 2 : 3;
}

bool pose_from_essential(Mat3 const& E, int n, double const* x1, double const* y1, double const* x2,
                         double const* y2, Mat3& R, double (&T)[3])
{
    Mat3 U, V;
    if (!essential_svd(E, U, V)) return false;
    int const conf0 = configuration(E, U, V, x1[0], y1[0], x2[0], y2[0]);
    for (int i = 1; i < n; ++i)
        if (configuration(E, U, V, x1[i], y1[i], x2[i], y2[i]) != conf0) return false;
    pose_of(U, V, conf0, R, T);
    return true;
}

// ---------------------------------------------------------------------------- refinement
// exp([w]x) = I + sin(th)/th [w]x + (1 - cos(th))/th^2 [w]x^2, th = |w|; a rotation of at
// most 1e-6 rad is taken as the identity, as the reference's optimiser does
// (createRotationMatrixRodrigues, V3D/Math/v3d_mathutilities.h:170-190): its refinement
// drops such rotation steps and moves only T, and this refinement follows it
Mat3 rodrigues(double const w[3])
{
    double const th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    Mat3 R = zero3();
    R(0, 0) = R(1, 1) = R(2, 2) = 1.0;
    if (!(std::fabs(th) > 1e-6)) return R;
    double const a = std::sin(th) / th, b = (1.0 - std::cos(th)) / (th * th);
    Mat3 K = zero3();
    K(0, 1) = -w[2];
    K(0, 2) = w[1];
    K(1, 0) = w[2];
    K(1, 2) = -w[0];
    K(2, 0) = -w[1];
    K(2, 1) = w[0];
    Mat3 const K2 = mul3(K, K);
    for (int i = 0; i < 9; ++i) R.m[i] += a * K.m[i] + b * K2.m[i];
    return R;
}

struct PoseRefiner {
    std::vector<double> nl, nr, pl, pr;  // normalised and pixel points (x, y interleaved)
    Mat3 iK1, iK2t;
    double thr;

    // signed Sampson residuals num / sqrt(denom) of the normalised points under E = [T]x R
    void residuals(Mat3 const& R, double const T[3], std::vector<double>& f) const
    {
        Mat3 const E = essential_from_pose(R, T);
        size_t const n = nl.size() / 2;
        for (size_t k = 0; k < n; ++k) {
            double const px = nl[2 * k], py = nl[2 * k + 1], qx = nr[2 * k], qy = nr[2 * k + 1];
            double const ep0 = affine_row(E, 0, px, py), ep1 = affine_row(E, 1, px, py), ep2 = affine_row(E, 2, px, py);
            double const etq0 = affine_col(E, 0, qx, qy), etq1 = affine_col(E, 1, qx, qy);
            double const num = qx * ep0 + qy * ep1 + ep2;
            f[k] = num / std::sqrt(ep0 * ep0 + ep1 * ep1 + etq0 * etq0 + etq1 * etq1);
        }
    }

    // 1 for a pixel-space Sampson residual below the threshold, else sqrt(thr / |e|)
    void weights(Mat3 const& R, double const T[3], std::vector<double>& w) const
    {
        Mat3 const F = mul3(mul3(iK2t, essential_from_pose(R, T)), iK1);
        size_t const n = pl.size() / 2;
        for (size_t k = 0; k < n; ++k) {
            double const px = pl[2 * k], py = pl[2 * k + 1], qx = pr[2 * k], qy = pr[2 * k + 1];
            double const fp0 = affine_row(F, 0, px, py), fp1 = affine_row(F, 1, px, py), fp2 = affine_row(F, 2, px, py);
            double const ftq0 = affine_col(F, 0, qx, qy), ftq1 = affine_col(F, 1, qx, qy);
            double const e = (qx * fp0 + qy * fp1 + fp2) / std::sqrt(fp0 * fp0 + fp1 * fp1 + ftq0 * ftq0 + ftq1 * ftq1);
            w[k] = (std::fabs(e) < thr) ? 1.0 : std::sqrt(thr / std::fabs(e));
        }
    }

    // d residual / d (omega, T) for the update R <- R rodrigues(omega), T <- T + dT:
    // dE/domega_i = E [e_i]x, dE/dT_i = [e_i]x R, chained with d(num / sqrt(S)) / dE
    void jacobian(Mat3 const& R, double const T[3], std::vector<double>& J) const
    {
        Mat3 const E = essential_from_pose(R, T);
        double inner[9][6];
        for (int i = 0; i < 3; ++i) {
            Mat3 ex = zero3();  // [e_i]x
            int const a = (i + 1) % 3, b = (i + 2) % 3;
            ex(b, a) = 1.0;
            ex(a, b) = -1.0;
            Mat3 const dw = mul3(E, ex), dt = mul3(ex, R);
            for (int j = 0; j < 9; ++j) {
                inner[j][i] = dw.m[j];
                inner[j][3 + i] = dt.m[j];
            }
        }
        size_t const n = nl.size() / 2;
        for (size_t k = 0; k < n; ++k) {
            double const p[3] = {nl[2 * k], nl[2 * k + 1], 1.0}, q[3] = {nr[2 * k], nr[2 * k + 1], 1.0};
            double ep[3], etq[3];
            for (int r = 0; r < 3; ++r) {
                ep[r] = affine_row(E, r, p[0], p[1]);
                etq[r] = affine_col(E, r, q[0], q[1]);
            }
            double const num = q[0] * ep[0] + q[1] * ep[1] + ep[2];
            double const S = ep[0] * ep[0] + ep[1] * ep[1] + etq[0] * etq[0] + etq[1] * etq[1];
            double const sq = std::sqrt(S);
            double const invSq = 1.0 / sq, invS = 1.0 / S;
            // d num / dE = q p^T; d S / dE: 2 (Ep)_r p_c (rows 0, 1) + 2 (E^T q)_c q_r (columns
            // 0, 1), the entries with a factor 1 or 0 written as the reference writes them
            double const dS[9] = {2 * ep[0] * p[0] + 2 * etq[0] * q[0], 2 * ep[0] * p[1] + 2 * etq[1] * q[0], 2 * ep[0],
                                  2 * ep[1] * p[0] + 2 * etq[0] * q[1], 2 * ep[1] * p[1] + 2 * etq[1] * q[1], 2 * ep[1],
                                  2 * etq[0], 2 * etq[1], 0.0};
            double outer[9];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    double const dn = q[r] * p[c];
                    outer[3 * r + c] = (dn * sq - num * 0.5 * invSq * dS[3 * r + c]) * invS;
                }
            for (int c = 0; c < 6; ++c) {
                double s = 0.0;
                for (int j = 0; j < 9; ++j) s += outer[j] * inner[j][c];
                J[6 * k + c] = s;
            }
        }
    }
};

// L L^T x = b, the Cholesky-Banachiewicz order of the public-domain JAMA/TNT routine the
// reference's optimiser calls (V3D/Math/v3d_linear_tnt.h:93-171): L_jk = (N_jk - sum_i<k
// L_ki L_ji) / L_kk, L_jj = sqrt(N_jj - sum_k<j L_jk^2); false when N is not positive
// definite
bool cholesky_solve6(double (&N)[6][6], double const (&b)[6], double (&x)[6])
{
    double L[6][6] = {};
    bool spd = true;
    for (int j = 0; j < 6; ++j) {
        double d = 0.0;
        for (int k = 0; k < j; ++k) {
            double s = 0.0;
            for (int i = 0; i < k; ++i) s += L[k][i] * L[j][i];
            L[j][k] = s = (N[j][k] - s) / L[k][k];
            d = d + s * s;
        }
        d = N[j][j] - d;
        spd = spd && d > 0.0;
        L[j][j] = std::sqrt(d > 0.0 ? d : 0.0);
    }
    if (!spd) return false;
    for (int k = 0; k < 6; ++k) x[k] = b[k];
    for (int k = 0; k < 6; ++k) {
        for (int i = 0; i < k; ++i) x[k] -= x[i] * L[k][i];
        x[k] /= L[k][k];
    }
    for (int k = 5; k >= 0; --k) {
        for (int i = k + 1; i < 6; ++i) x[k] -= x[i] * L[i][k];
        x[k] /= L[k][k];
    }
    return true;
}

// Levenberg-Marquardt with the reference optimiser's schedule (v3d_optimization.h:24-277:
// at most 50 iterations, at least 10 before the small-update test, lambda = 1e-3 max
// diag(J^T J) at the start, x10 on a rejected step, x0.1 (>= 1e-10) on an accepted one,
// gradient and update thresholds 1e-8, parameter length 4); the residual vector is the
// weighted residual, the Jacobian the unweighted one, as there.
void refine_pose(PoseRefiner const& pr, Mat3& R, double (&T)[3])
{
    size_t const n = pr.nl.size() / 2;
    if (n == 0) return;
    std::vector<double> f(n), w(n), e(n), J(6 * n);
    double lambda = 1e-3;
    bool derivs = true;
    double JtJ[6][6], Jte[6];
    for (int it = 0; it < 50; ++it) {
        pr.weights(R, T, w);
        pr.residuals(R, T, f);
        double err = 0.0;
        for (size_t k = 0; k < n; ++k) {
            e[k] = -w[k] * f[k];
            err += e[k] * e[k];
        }
        if (derivs) {
            pr.jacobian(R, T, J);
            for (int a = 0; a < 6; ++a) {
                for (int b = 0; b < 6; ++b) {
                    double s = 0.0;
                    for (size_t k = 0; k < n; ++k) s += J[6 * k + a] * J[6 * k + b];
                    JtJ[a][b] = s;
                }
                double s = 0.0;
                for (size_t k = 0; k < n; ++k) s += J[6 * k + a] * e[k];
                Jte[a] = s;
            }
        }
        double gmax = 0.0;
        for (int a = 0; a < 6; ++a) gmax = std::max(gmax, std::fabs(Jte[a]));
        if (gmax < 1e-8) return;
        if (it == 0) {
            double mx = -1e30;
            for (int a = 0; a < 6; ++a) mx = std::max(mx, JtJ[a][a]);
            lambda = 1e-3 * mx;
        }
        double N[6][6];
        for (int a = 0; a < 6; ++a)
            for (int b = 0; b < 6; ++b) N[a][b] = JtJ[a][b] + (a == b ? lambda : 0.0);
        double delta[6];
        if (!cholesky_solve6(N, Jte, delta)) return;
        double dl2 = 0.0;
        for (int a = 0; a < 6; ++a) dl2 += delta[a] * delta[a];
        double const dl = std::sqrt(dl2);
        if (it >= 10 && dl < 1e-8 * (4.0 + 1e-8)) return;
        Mat3 const savedR = R;
        double const savedT[3] = {T[0], T[1], T[2]};
        double const om[3] = {delta[0], delta[1], delta[2]};
        R = mul3(R, rodrigues(om));
        for (int i = 0; i < 3; ++i) T[i] += delta[3 + i];
        double const tn = norm3(T);
        for (int i = 0; i < 3; ++i) T[i] /= tn;
        pr.residuals(R, T, f);
        double errNew = 0.0;
        for (size_t k = 0; k < n; ++k) errNew += (w[k] * f[k]) * (w[k] * f[k]);
        double dJ = 0.0;
        for (int a = 0; a < 6; ++a) dJ += delta[a] * Jte[a];
        double const rho = (err - errNew) / (dl2 * lambda + dJ);
        if (rho > 0.0) {
            lambda = std::max(1e-10, lambda * 0.1);
            derivs = true;
        } else {
            R = savedR;
            for (int i = 0; i < 3; ++i) T[i] = savedT[i];
            lambda *= 10.0;
            derivs = false;
        }
    }
}

PoseRefiner make_refiner(double const* corrs, std::vector<int> const& idx, Mat3 const& iK1, Mat3 const& iK2,
                         double thr)
{
    PoseRefiner pr;
    pr.iK1 = iK1;
    pr.iK2t = transpose3(iK2);
    pr.thr = thr;
    size_t const n = idx.size();
    pr.nl.resize(2 * n);
    pr.nr.resize(2 * n);
    pr.pl.resize(2 * n);
    pr.pr.resize(2 * n);
    for (size_t k = 0; k < n; ++k) {
        double const* c = corrs + 4 * (size_t)idx[k];
        pr.pl[2 * k] = c[0];
        pr.pl[2 * k + 1] = c[1];
        pr.pr[2 * k] = c[2];
        pr.pr[2 * k + 1] = c[3];
        normalise_point(iK1, c[0], c[1], pr.nl[2 * k], pr.nl[2 * k + 1]);
        normalise_point(iK2, c[2], c[3], pr.nr[2 * k], pr.nr[2 * k + 1]);
    }
    return pr;
}

// ---------------------------------------------------------------------------- sampling
// splitmix64: the library's own generator for the index shuffles
struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
};

// getShuffleCount (v3d_homography.cpp:14-18)
int shuffle_count(int N, int nRounds, int minSampleSize)
{
    return std::max(1, (int)((float)nRounds * (float)minSampleSize / (float)N + 0.5));
}

// ransacNSamples(nNeededSamples, outlierRatio, wantedPsucces) (reconstruction_common.h:642-650)
int ransac_samples(int nNeeded, double outlierRatio, double wanted)
{
    double const maxN = 1e8;
    if (outlierRatio >= 1.0) return (int)maxN;
    double N = std::log(1.0 - wanted) / std::log(1.0 - std::pow(1.0 - outlierRatio, nNeeded));
    N = std::max(0.0, std::min(maxN, N));
    return (int)N;
}

}  // namespace

extern "C" {

int v3dm_sampson_errors(const double F[9], const double* corrs, int n, double* out)
{
    if (!F || n < 0 || (n > 0 && (!corrs || !out))) return V3DM_ERR_INVALID;
    Mat3 const f = load3(F);
    for (int i = 0; i < n; ++i) {  // positions as the reference's PointMeasurement holds them: float
        double const* c = corrs + 4 * (size_t)i;
        out[i] = sampson(f, (float)c[0], (float)c[1], (float)c[2], (float)c[3]);
    }
    return V3DM_OK;
}

int v3dm_essentials_from_five(const double x1[5], const double y1[5], const double x2[5], const double y2[5],
                              double Es[90], int* nSolutions)
{
    if (!x1 || !y1 || !x2 || !y2 || !Es || !nSolutions) return V3DM_ERR_INVALID;
    *nSolutions = five_point(x1, y1, x2, y2, Es);
    return V3DM_OK;
}

int v3dm_relative_pose_from_essential(const double E[9], int n, const double* x1, const double* y1,
                                      const double* x2, const double* y2, double R[9], double T[3], int* ok)
{
    if (!E || n < 1 || !x1 || !y1 || !x2 || !y2 || !R || !T || !ok) return V3DM_ERR_INVALID;
    Mat3 r;
    double t[3];
    bool const good = pose_from_essential(load3(E), n, x1, y1, x2, y2, r, t);
    if (good) {
        std::memcpy(R, r.m, sizeof(r.m));
        for (int i = 0; i < 3; ++i) T[i] = t[i];
    } else {
        std::memset(R, 0, 9 * sizeof(double));
        std::memset(T, 0, 3 * sizeof(double));
    }
    *ok = good ? 1 : 0;
    return V3DM_OK;
}

int v3dm_refine_relative_pose(const double* corrs, int n, const double K1[9], const double K2[9], double R[9],
                              double T[3], double sampsonInlierThreshold)
{
    if (n < 0 || (n > 0 && !corrs) || !K1 || !K2 || !R || !T) return V3DM_ERR_INVALID;
    Mat3 iK1, iK2;
    if (!invert3(load3(K1), iK1) || !invert3(load3(K2), iK2)) return V3DM_ERR_INVALID;
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    PoseRefiner const pr = make_refiner(corrs, idx, iK1, iK2, sampsonInlierThreshold);
    Mat3 r = load3(R);
    double t[3] = {T[0], T[1], T[2]};
    refine_pose(pr, r, t);
    std::memcpy(R, r.m, sizeof(r.m));
    for (int i = 0; i < 3; ++i) T[i] = t[i];
    return V3DM_OK;
}

int v3dm_robust_orientation(const double* raw, int n, const double K1[9], const double K2[9],
                            double inlierThreshold, int nSamples, int reportInliers, int iterativeRefinement,
                            uint64_t seed, v3dm_orientation* res, int32_t* inliers)
{
    int const minSize = 5;
    double const* corrs = raw;
    if (!res || !K1 || !K2 || (n > 0 && !corrs)) return V3DM_ERR_INVALID;
    if (n < minSize) return V3DM_ERR_INVALID;  // "At least 5 point correspondences are required."
    if (reportInliers && !inliers) return V3DM_ERR_INVALID;
    std::memset(res, 0, sizeof(*res));
    res->error = INFINITY;
    // the reference's correspondences hold float positions (PointMeasurement), and it keeps
    // the normalised ones in the same type: the samples' coordinates are float-rounded
    std::vector<double> cf(4 * (size_t)n);
    for (size_t i = 0; i < cf.size(); ++i) cf[i] = (float)raw[i];
    corrs = cf.data();
    double const sqrThr = inlierThreshold * inlierThreshold;
    Mat3 const k1 = load3(K1), k2 = load3(K2);
    Mat3 iK1, iK2;
    if (!invert3(k1, iK1) || !invert3(k2, iK2)) return V3DM_ERR_INVALID;
    Mat3 const iK2t = transpose3(iK2);
    std::vector<double> nx1(n), ny1(n), nx2(n), ny2(n);
    for (int i = 0; i < n; ++i) {
        double const* c = corrs + 4 * (size_t)i;
        normalise_point_t<float>(iK1, c[0], c[1], nx1[i], ny1[i]);
        normalise_point_t<float>(iK2, c[2], c[3], nx2[i], ny2[i]);
    }
    unsigned const nShuffles = (unsigned)shuffle_count(n, nSamples, minSize);
    double minError = 1e30;
    std::vector<int> indices(n);
    for (int i = 0; i < n; ++i) indices[i] = i;
    Rng rng{seed};
    Mat3 bestF = zero3();
    bool have = false;
    std::vector<int> inl;
    int drawn = 0, hyps = 0;
    double Es[90];
    for (unsigned s = 0; s < nShuffles; ++s) {
        if (drawn > nSamples) break;
        for (int i = n - 1; i > 0; --i) std::swap(indices[i], indices[rng.below((uint32_t)i + 1u)]);
        for (int j = 0; j < n - minSize; j += minSize) {
            if (drawn > nSamples) break;
            double x1[5], y1[5], x2[5], y2[5];
            for (int k = 0; k < 5; ++k) {
                int const ix = indices[j + k];
                x1[k] = nx1[ix];
                y1[k] = ny1[ix];
                x2[k] = nx2[ix];
                y2[k] = ny2[ix];
            }
            int const nE = five_point(x1, y1, x2, y2, Es);
            ++drawn;
            for (int r = 0; r < nE; ++r) {
                ++hyps;
                Mat3 E = load3(Es + 9 * r);
                Mat3 fund = mul3(mul3(iK2t, E), iK1);
                int nInl = 0;
                double curError = 0.0;
                for (int i = 0; i < n; ++i) {
                    double const* c = corrs + 4 * (size_t)i;
                    double const dist = sampson(fund, c[0], c[1], c[2], c[3]);
                    curError += std::min(dist, sqrThr);
                    if (dist < sqrThr) ++nInl;
                }
                if (!(curError < minError)) continue;
                Mat3 R;
                double T[3];
                if (!pose_from_essential(E, 5, x1, y1, x2, y2, R, T)) continue;
                if (iterativeRefinement) {
                    inl.clear();
                    for (int i = 0; i < n; ++i) {
                        double const* c = corrs + 4 * (size_t)i;
                        if (sampson(fund, c[0], c[1], c[2], c[3]) < sqrThr) inl.push_back(i);
                    }
                    PoseRefiner const pr = make_refiner(corrs, inl, iK1, iK2, inlierThreshold);
                    refine_pose(pr, R, T);
                    Mat3 const Eref = essential_from_pose(R, T);
                    Mat3 const Fref = mul3(mul3(iK2t, Eref), iK1);
                    double errRef = 0.0;
                    nInl = 0;
                    for (int i = 0; i < n; ++i) {
                        double const* c = corrs + 4 * (size_t)i;
                        double const dist = sampson(Fref, c[0], c[1], c[2], c[3]);
                        errRef += std::min(dist, sqrThr);
                        if (dist < sqrThr) ++nInl;
                    }
                    fund = Fref;
                    E = Eref;
                    curError = std::min(curError, errRef);
                }
                minError = curError;
                res->error = minError / n;
                std::memcpy(res->essential, E.m, sizeof(E.m));
                std::memcpy(res->fundamental, fund.m, sizeof(fund.m));
                std::memcpy(res->rotation, R.m, sizeof(R.m));
                for (int i = 0; i < 3; ++i) res->translation[i] = T[i];
                bestF = fund;
                have = true;
                // adaptive number of samples (confidence 1 - 10e-10, as the reference)
                double const outlierFraction = 1.0 - (float)(nInl - minSize) / (float)(n - minSize);
                nSamples = std::min(nSamples, ransac_samples(minSize, outlierFraction, 1.0 - 10e-10));
            }
        }
    }
    res->nSamplesDrawn = drawn;
    res->nHypotheses = hyps;
    int cnt = 0;
    if (have) {
        for (int i = 0; i < n; ++i) {
            double const* c = corrs + 4 * (size_t)i;
            if (sampson(bestF, c[0], c[1], c[2], c[3]) < sqrThr) {
                if (reportInliers) inliers[cnt] = i;
                ++cnt;
            }
        }
    }
    res->nInliers = cnt;
    return V3DM_OK;
}

}  // extern "C"
