// Host-side FP64 setup: grid, Chebyshev operators, eigen-decomposition.
//
// Everything here runs once per context on the CPU (the reference does the
// same work once in initialize(), init.f90:36-106).  The per-step hot path
// only consumes the matrices built here.

#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "lsb_internal.h"

namespace lsb {

using ld = long double;
static const double kPi = 4.0 * std::atan(1.0);

// ---------------------------------------------------------------- grid
// init.f90:130-273 setup_grid (linstab = 0, one rank).
int build_grid(lsb_ctx* c) {
    const lsb_config& g = c->cfg;
    const int nx = c->nx, ny = c->ny, nz = c->nz, nzp = c->nzp;
    const double twopi = 8.0 * std::atan(1.0);
    c->dx = g.xlen / (nx - 1.0);  // Q4: the reference spaces xpts by xlen/(nx-1)
    c->xpts.assign(nx, 0.0);
    for (int i = 0; i < nx; ++i) c->xpts[i] = g.x0 + i * c->dx;
    c->dy = g.ylen / (ny - 1.0);
    c->alpha = twopi / g.xlen;
    c->wavx.assign(nx, 0.0);
    for (int i = 0; i < nx / 2; ++i) {
        c->wavx[2 * i] = (i + 0.0) * c->alpha;
        c->wavx[2 * i + 1] = c->wavx[2 * i];
    }
    c->beta = twopi / g.ylen;
    c->wavy.assign(ny, 0.0);
    for (int j = 1; j <= ny / 2; ++j) {
        c->wavy[ny - j] = -j * c->beta;
        c->wavy[j - 1] = (j - 1.0) * c->beta;
    }
    c->wavxx.resize(nx);
    c->wavyy.resize(ny);
    for (int i = 0; i < nx; ++i) c->wavxx[i] = c->wavx[i] * c->wavx[i];
    for (int j = 0; j < ny; ++j) c->wavyy[j] = c->wavy[j] * c->wavy[j];
    c->zpts.assign(nzp, 0.0);
    c->gf.assign(nzp, 0.0);
    std::vector<double> xi(nzp, 0.0);
    switch (g.mapping) {
        case 0:
            c->xl = g.zlen;
            for (int k = 0; k < nzp; ++k) {
                xi[k] = c->xl / 2.0 * std::cos(kPi * (k + 0.0) / nz);
                c->zpts[k] = xi[k];
                c->gf[k] = 1.0;
            }
            break;
        case 1:
            c->xl = 1.0;
            c->aa = 1.0;
            c->bb = 1.875 * std::sqrt(g.xnu * g.x0 / g.u0);
            for (int k = 1; k < nzp; ++k) {
                xi[k] = c->xl / 2.0 * (1.0 + std::cos(kPi * (k + 0.0) / nz));
                c->zpts[k] = c->bb * xi[k] / (c->aa - xi[k]);
                c->gf[k] = (c->aa - xi[k]) * (c->aa - xi[k]) / (c->bb * c->aa);
            }
            c->zpts[0] = 1.0e20;
            c->gf[0] = 0.0;
            break;
        case 2:
            c->xl = 1.0;
            c->bb = 1.875 * std::sqrt(g.xnu * (g.x0 + g.blfac * g.xlen) / g.u0);
            c->aa = 1.0 + c->bb / g.zlen;
            for (int k = 0; k < nzp; ++k) {
                xi[k] = c->xl / 2.0 * (1.0 + std::cos(kPi * (k + 0.0) / (double)nz));
                c->zpts[k] = c->bb * xi[k] / (c->aa - xi[k]);
                c->gf[k] = (c->aa - xi[k]) * (c->aa - xi[k]) / (c->aa * c->bb);
            }
            break;
        case 3:
            c->xl = g.zlen;
            for (int k = 0; k < nzp; ++k) {
                xi[k] = c->xl / 2.0 * std::cos(kPi * (k + 0.0) / (double)nz);
                c->zpts[k] = c->xl / 2.0 + xi[k];
                c->gf[k] = 1.0;
            }
            break;
        default:
            set_error("mapping must be 0..3");
            return LSB_E_ARG;
    }
    return LSB_OK;
}

// ------------------------------------------------- Chebyshev matrix entries
// chebyshev.f90:51-159 (del/den/xnum/ff), 0-based kr, jr.
static double cheb_ff(int i, int nr) {
    if (i == 0) return 0.0;
    double v = nr * 0.5 / std::tan(kPi * i / (2.0 * nr));
    if (i % 2 == 0) v = -v;
    return v;
}
static double cheb_xnum(int kr, int jr, int nr) {
    if (kr == 0) {
        if (jr == 0) return (1.0 / 3.0) * nr * nr + 1.0 / 6.0;
        if (jr == nr) return (nr % 2 == 1) ? -0.5 : 0.5;
        const double t = 1.0 / std::tan(kPi * jr / (2.0 * nr));
        const double s = std::sin(kPi * jr / (2.0 * nr));
        return 0.5 * (nr + 0.5) * ((nr + 0.5) + t * t) + 1.0 / 8.0 - 0.25 / (s * s) - 0.5 * nr * nr;
    }
    return cheb_ff(kr + jr, nr) + cheb_ff(kr - jr, nr);
}
static double cheb_den(int kr, int jr, int nr) {
    if (kr == 0) {
        double d = 0.5 * nr;
        if (jr % 2 == 1) d = -d;
        if (jr == 0 || jr == nr) d = 1.0;
        return d;
    }
    double d = nr * std::sin(kPi * kr / nr);
    if (jr == 0 || jr == nr) d *= 2.0;
    return d;
}

// chebyshev.f90:9-48: dor(i,j) = del(i,j,nzp), d = diag(gf) dor, d2 = d*d.
void cheb_matrix(int nzp, double xl, const double* gf, double* dor, double* d, double* d2) {
    const int nr = nzp - 1;
    for (int j = 0; j < nzp; ++j) {
        for (int i = 0; i < nzp; ++i) {
            int kr = i, jr = j;
            double fac = 1.0;
            if (kr == nr) {
                kr = 0;
                jr = nr - jr;
                fac = -1.0;
            }
            double v = fac * cheb_xnum(kr, jr, nr) / cheb_den(kr, jr, nr);
            v = v * (2.0 / xl);
            dor[i + (size_t)nzp * j] = v;
            d[i + (size_t)nzp * j] = v * gf[i];
        }
    }
    // d2 = matmul(d, d) with long double accumulation
    for (int j = 0; j < nzp; ++j)
        for (int i = 0; i < nzp; ++i) {
            ld s = 0;
            for (int k = 0; k < nzp; ++k) s += (ld)d[i + (size_t)nzp * k] * (ld)d[k + (size_t)nzp * j];
            d2[i + (size_t)nzp * j] = (double)s;
        }
}

// chebyshev.f90:204-270 dct_diff applied to the unit vectors: column j of the
// returned (column-major n x n) matrix is dct_diff(e_j).  The DCT-I
// (FFTW REDFT00: Y_k = x_0 + (-1)^k x_{n-1} + 2 sum x_j cos(pi j k/(n-1))) is
// evaluated directly in long double.  The reference scales only alpha(2:n-1) by
// the half interval (chebyshev.f90:242); the top coefficient needs 1/(2 half) =
// 1/xl (its DCT-I weight is half the interior one), so the reference operator is
// exact for xl = 1 and doubles the derivative of T_N for xl = 2.  exact = 1
// applies 1/xl to alpha(n): the exact Gauss-Lobatto derivative on every mapping.
void dct_diff_matrix(int n, double xl, const double* gf, int p, double* m, int exact) {
    const ld half = 0.5L * (ld)xl;
    const ld nlog = 2.0L * (n - 1);
    std::vector<ld> cs((size_t)n * n);
    const ld pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j) {
            long long q = ((long long)j * k) % (2LL * (n - 1));
            cs[(size_t)k * n + j] = cosl(pi * (ld)q / (ld)(n - 1));
        }
    auto redft00 = [&](const std::vector<ld>& x, std::vector<ld>& y) {
        for (int k = 0; k < n; ++k) {
            ld s = x[0] + ((k % 2) ? -x[n - 1] : x[n - 1]);
            for (int j = 1; j < n - 1; ++j) s += 2.0L * x[j] * cs[(size_t)k * n + j];
            y[k] = s;
        }
    };
    std::vector<ld> beta(n), alpha(n);
    for (int col = 0; col < n; ++col) {
        for (int i = 0; i < n; ++i) beta[i] = (i == col) ? 1.0L : 0.0L;
        for (int l = 0; l < p; ++l) {
            redft00(beta, alpha);
            for (int i = 1; i < n - 1; ++i) alpha[i] /= half;  // alpha(2:n-1) only (Q2)
            if (exact) alpha[n - 1] /= (ld)xl;
            for (int i = 0; i < n; ++i) beta[i] = 0;
            beta[n - 2] = 2.0L * (n - 1.0L) * alpha[n - 1];
            for (int k = n - 1; k >= 2; --k) beta[k - 2] = beta[k] + 2.0L * (k - 1.0L) * alpha[k - 1];
            redft00(beta, alpha);
            for (int i = 0; i < n; ++i) beta[i] = alpha[i] * (ld)gf[i] / nlog;
        }
        for (int i = 0; i < n; ++i) m[i + (size_t)n * col] = (double)beta[i];
    }
}

// ------------------------------------------------------- eigen-decomposition
// Replaces helmholtz.f90:13-126 (LAPACK dggev with B = I, dgetrf, dgetri).
// 1. Householder reduction to Hessenberg form H = Q^T A Q (Q accumulated);
// 2. Francis double-shift QR on H for the eigenvalues (all must be real, as
//    the reference assumes at helmholtz.f90:88);
// 3. inverse iteration with the Hessenberg LU for each eigenvector, v = Q y;
// 4. E^-1 by Gauss-Jordan with partial pivoting; eigenvalues refined as the
//    diagonal of E^-1 A E.  All in long double.
static void hess_reduce(std::vector<ld>& a, std::vector<ld>& q, int n) {
    // a, q row-major n x n; q starts as identity
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) q[(size_t)i * n + j] = (i == j) ? 1.0L : 0.0L;
    std::vector<ld> v(n);
    for (int k = 0; k < n - 2; ++k) {
        ld alpha = 0;
        for (int i = k + 1; i < n; ++i) alpha += a[(size_t)i * n + k] * a[(size_t)i * n + k];
        alpha = sqrtl(alpha);
        if (alpha == 0) continue;
        if (a[(size_t)(k + 1) * n + k] > 0) alpha = -alpha;
        for (int i = 0; i < n; ++i) v[i] = 0;
        v[k + 1] = a[(size_t)(k + 1) * n + k] - alpha;
        for (int i = k + 2; i < n; ++i) v[i] = a[(size_t)i * n + k];
        ld vn = 0;
        for (int i = k + 1; i < n; ++i) vn += v[i] * v[i];
        if (vn == 0) continue;
        const ld f = 2.0L / vn;
        for (int j = 0; j < n; ++j) {
            ld s = 0;
            for (int i = k + 1; i < n; ++i) s += v[i] * a[(size_t)i * n + j];
            s *= f;
            for (int i = k + 1; i < n; ++i) a[(size_t)i * n + j] -= s * v[i];
        }
        for (int i = 0; i < n; ++i) {
            ld s = 0;
            for (int j = k + 1; j < n; ++j) s += a[(size_t)i * n + j] * v[j];
            s *= f;
            for (int j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= s * v[j];
        }
        // Q = Q P
        for (int i = 0; i < n; ++i) {
            ld s = 0;
            for (int j = k + 1; j < n; ++j) s += q[(size_t)i * n + j] * v[j];
            s *= f;
            for (int j = k + 1; j < n; ++j) q[(size_t)i * n + j] -= s * v[j];
        }
        for (int i = k + 2; i < n; ++i) a[(size_t)i * n + k] = 0;
    }
}

static void house3(ld x, ld y, ld z, ld v[3], ld* beta) {
    ld nrm = sqrtl(x * x + y * y + z * z);
    if (nrm == 0) {
        v[0] = 1;
        v[1] = v[2] = 0;
        *beta = 0;
        return;
    }
    ld alpha = (x > 0) ? -nrm : nrm;
    v[0] = x - alpha;
    v[1] = y;
    v[2] = z;
    ld vn = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    *beta = (vn == 0) ? 0 : 2.0L / vn;
}

static int francis_eigenvalues(std::vector<ld>& h, int n, std::vector<ld>& wr, std::vector<ld>& wi) {
    auto H = [&](int i, int j) -> ld& { return h[(size_t)i * n + j]; };
    ld anorm = 0;
    for (size_t i = 0; i < h.size(); ++i) anorm = fmaxl(anorm, fabsl(h[i]));
    const ld eps = LDBL_EPSILON;
    int p = n - 1, iter = 0, total = 0;
    while (p >= 0) {
        int l;
        for (l = p; l > 0; --l) {
            ld s = fabsl(H(l - 1, l - 1)) + fabsl(H(l, l));
            if (s == 0) s = anorm;
            if (fabsl(H(l, l - 1)) <= eps * s) {
                H(l, l - 1) = 0;
                break;
            }
        }
        if (l == p) {
            wr[p] = H(p, p);
            wi[p] = 0;
            --p;
            iter = 0;
            continue;
        }
        if (l == p - 1) {
            ld a = H(p - 1, p - 1), b = H(p - 1, p), cc = H(p, p - 1), d = H(p, p);
            ld m = 0.5L * (a + d);
            ld disc = 0.25L * (a - d) * (a - d) + b * cc;
            if (disc >= 0) {
                ld s = sqrtl(disc);
                ld r1 = (m >= 0) ? m + s : m - s;
                ld det = a * d - b * cc;
                ld r2 = (r1 != 0) ? det / r1 : m - s;
                wr[p - 1] = r1;
                wr[p] = r2;
                wi[p - 1] = wi[p] = 0;
            } else {
                ld s = sqrtl(-disc);
                wr[p - 1] = wr[p] = m;
                wi[p - 1] = s;
                wi[p] = -s;
            }
            p -= 2;
            iter = 0;
            continue;
        }
        if (++iter > 200 || ++total > 100 * n) return LSB_E_NUMERIC;
        ld s, t;
        if (iter % 11 == 10) {
            ld x = fabsl(H(p, p - 1)) + fabsl(H(p - 1, p - 2));
            s = 1.5L * x;
            t = x * x;
        } else {
            s = H(p - 1, p - 1) + H(p, p);
            t = H(p - 1, p - 1) * H(p, p) - H(p - 1, p) * H(p, p - 1);
        }
        ld x = H(l, l) * H(l, l) + H(l, l + 1) * H(l + 1, l) - s * H(l, l) + t;
        ld y = H(l + 1, l) * (H(l, l) + H(l + 1, l + 1) - s);
        ld z = H(l + 1, l) * H(l + 2, l + 1);
        for (int k = l; k <= p - 2; ++k) {
            ld v[3], beta;
            house3(x, y, z, v, &beta);
            if (beta != 0) {
                const int r = (k > l) ? k - 1 : l;
                for (int j = r; j <= p; ++j) {
                    ld sum = v[0] * H(k, j) + v[1] * H(k + 1, j) + v[2] * H(k + 2, j);
                    sum *= beta;
                    H(k, j) -= sum * v[0];
                    H(k + 1, j) -= sum * v[1];
                    H(k + 2, j) -= sum * v[2];
                }
                const int rmax = (k + 3 < p) ? k + 3 : p;
                for (int i = l; i <= rmax; ++i) {
                    ld sum = H(i, k) * v[0] + H(i, k + 1) * v[1] + H(i, k + 2) * v[2];
                    sum *= beta;
                    H(i, k) -= sum * v[0];
                    H(i, k + 1) -= sum * v[1];
                    H(i, k + 2) -= sum * v[2];
                }
            }
            x = H(k + 1, k);
            y = H(k + 2, k);
            if (k < p - 2) z = H(k + 3, k);
        }
        // final 2x2 Givens on rows p-1, p
        {
            ld r = sqrtl(x * x + y * y);
            if (r != 0) {
                ld cg = x / r, sg = y / r;
                for (int j = p - 2; j <= p; ++j) {
                    ld t1 = H(p - 1, j), t2 = H(p, j);
                    H(p - 1, j) = cg * t1 + sg * t2;
                    H(p, j) = -sg * t1 + cg * t2;
                }
                for (int i = l; i <= p; ++i) {
                    ld t1 = H(i, p - 1), t2 = H(i, p);
                    H(i, p - 1) = cg * t1 + sg * t2;
                    H(i, p) = -sg * t1 + cg * t2;
                }
            }
        }
    }
    return LSB_OK;
}

// Solve (Hs - lam I) y = b for upper Hessenberg Hs (row-major) by LU with
// adjacent-row partial pivoting; b overwritten by y.
static void hess_solve(const std::vector<ld>& hs, int n, ld lam, std::vector<ld>& b, ld tiny) {
    std::vector<ld> u(hs);
    for (int i = 0; i < n; ++i) u[(size_t)i * n + i] -= lam;
    for (int k = 0; k < n - 1; ++k) {
        ld* rk = &u[(size_t)k * n];
        ld* rk1 = &u[(size_t)(k + 1) * n];
        if (fabsl(rk1[k]) > fabsl(rk[k])) {
            for (int j = k; j < n; ++j) std::swap(rk[j], rk1[j]);
            std::swap(b[k], b[k + 1]);
        }
        if (rk[k] == 0) rk[k] = tiny;
        ld m = rk1[k] / rk[k];
        if (m != 0) {
            for (int j = k; j < n; ++j) rk1[j] -= m * rk[j];
            b[k + 1] -= m * b[k];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        ld s = b[i];
        for (int j = i + 1; j < n; ++j) s -= u[(size_t)i * n + j] * b[j];
        ld piv = u[(size_t)i * n + i];
        if (piv == 0) piv = tiny;
        b[i] = s / piv;
    }
}

static int invert(std::vector<ld>& a, int n, std::vector<ld>& inv) {
    inv.assign((size_t)n * n, 0);
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1;
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int i = k + 1; i < n; ++i)
            if (fabsl(a[(size_t)i * n + k]) > fabsl(a[(size_t)piv * n + k])) piv = i;
        if (a[(size_t)piv * n + k] == 0) return LSB_E_NUMERIC;
        if (piv != k)
            for (int j = 0; j < n; ++j) {
                std::swap(a[(size_t)k * n + j], a[(size_t)piv * n + j]);
                std::swap(inv[(size_t)k * n + j], inv[(size_t)piv * n + j]);
            }
        ld d = a[(size_t)k * n + k];
        for (int j = 0; j < n; ++j) {
            a[(size_t)k * n + j] /= d;
            inv[(size_t)k * n + j] /= d;
        }
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            ld f = a[(size_t)i * n + k];
            if (f == 0) continue;
            for (int j = 0; j < n; ++j) {
                a[(size_t)i * n + j] -= f * a[(size_t)k * n + j];
                inv[(size_t)i * n + j] -= f * inv[(size_t)k * n + j];
            }
        }
    }
    return LSB_OK;
}

// a: column-major n x n (Fortran D2).  Outputs: lam (n), e and ei column-major.
int eigen_decompose(const double* a_cm, int n, double* lam_out, double* e_out, double* ei_out) {
    std::vector<ld> a((size_t)n * n), q((size_t)n * n);
    ld anorm = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            a[(size_t)i * n + j] = a_cm[i + (size_t)n * j];
            anorm = fmaxl(anorm, fabsl(a[(size_t)i * n + j]));
        }
    std::vector<ld> a0(a);
    hess_reduce(a, q, n);
    std::vector<ld> hs(a);  // Hessenberg before QR
    std::vector<ld> wr(n), wi(n);
    int rc = francis_eigenvalues(a, n, wr, wi);
    if (rc != LSB_OK) {
        set_error("diagonalize_matrix: QR iteration did not converge");
        return rc;
    }
    for (int i = 0; i < n; ++i)
        if (wi[i] != 0) {
            set_error("diagonalize_matrix: complex eigenvalue (reference assumes a real spectrum)");
            return LSB_E_NUMERIC;
        }
    // eigenvectors by inverse iteration
    std::vector<ld> E((size_t)n * n);  // row-major, column i = eigenvector i
    std::vector<ld> y(n), v(n);
    const ld tiny = LDBL_EPSILON * anorm;
    for (int c = 0; c < n; ++c) {
        ld lam = wr[c];
        ld shift = lam + fmaxl(fabsl(lam), 1.0L) * 64 * LDBL_EPSILON;
        for (int i = 0; i < n; ++i) y[i] = 1.0L + 0.1L * (ld)((i * 7919 + c * 104729) % 97) / 97.0L;
        for (int it = 0; it < 4; ++it) {
            hess_solve(hs, n, shift, y, tiny);
            ld m = 0;
            for (int i = 0; i < n; ++i) m = fmaxl(m, fabsl(y[i]));
            if (m == 0 || !std::isfinite((double)m)) {
                set_error("diagonalize_matrix: inverse iteration failed");
                return LSB_E_NUMERIC;
            }
            for (int i = 0; i < n; ++i) y[i] /= m;
        }
        for (int i = 0; i < n; ++i) {
            ld s = 0;
            for (int j = 0; j < n; ++j) s += q[(size_t)i * n + j] * y[j];
            v[i] = s;
        }
        int imax = 0;
        for (int i = 1; i < n; ++i)
            if (fabsl(v[i]) > fabsl(v[imax])) imax = i;
        ld sc = v[imax];
        for (int i = 0; i < n; ++i) E[(size_t)i * n + c] = v[i] / sc;
    }
    std::vector<ld> Ecopy(E), Ei;
    rc = invert(Ecopy, n, Ei);
    if (rc != LSB_OK) {
        set_error("diagonalize_matrix: singular eigenvector matrix");
        return rc;
    }
    // refined eigenvalues: diag(E^-1 A E)
    std::vector<ld> AE((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            ld s = 0;
            for (int k = 0; k < n; ++k) s += a0[(size_t)i * n + k] * E[(size_t)k * n + j];
            AE[(size_t)i * n + j] = s;
        }
    for (int c = 0; c < n; ++c) {
        ld s = 0;
        for (int k = 0; k < n; ++k) s += Ei[(size_t)c * n + k] * AE[(size_t)k * n + c];
        lam_out[c] = (double)s;
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            e_out[i + (size_t)n * j] = (double)E[(size_t)i * n + j];
            ei_out[i + (size_t)n * j] = (double)Ei[(size_t)i * n + j];
        }
    return LSB_OK;
}

// ----------------------------------------------------- lifting functions
// helmholtz.f90:240-307 (per mapping); mapping 2 needs wbar/wbarpp.
void lifting_functions(lsb_ctx* c) {
    const int nzp = c->nzp;
    const std::vector<double>& z = c->zpts;
    c->hbot.assign(nzp, 0.0);
    c->htop.assign(nzp, 0.0);
    c->d2hbot.assign(nzp, 0.0);
    c->d2htop.assign(nzp, 0.0);
    switch (c->cfg.mapping) {
        case 0: {
            double width = z[0] - z[nzp - 1];
            for (int k = 0; k < nzp; ++k) {
                c->htop[k] = (z[k] - z[nzp - 1]) / width;
                c->hbot[k] = -(z[k] - z[0]) / width;
            }
            break;
        }
        case 1: {
            const double bbot = 50.0;
            for (int k = 0; k < nzp; ++k) c->hbot[k] = std::exp(-bbot * z[k]);
            c->hbot[0] = 0.0;
            for (int k = 0; k < nzp; ++k) c->d2hbot[k] = std::pow(bbot, 2.0) * c->hbot[k];
            break;
        }
        case 2: {
            const double h = z[0] - z[nzp - 1];
            const double bbot = 5.0;
            const double den = 1.0 - std::exp(-2.0 * bbot * h);
            for (int k = 0; k < nzp; ++k) {
                c->hbot[k] = (std::exp(-bbot * z[k]) - std::exp(bbot * (z[k] - 2.0 * h))) / den;
                c->d2hbot[k] = bbot * bbot * c->hbot[k];
            }
            if (!c->wbar.empty() && c->wbar[0] != 0.0) {
                for (int k = 0; k < nzp; ++k) {
                    c->htop[k] = c->wbar[k] / c->wbar[0];
                    c->d2htop[k] = c->wbarpp[k] / c->wbar[0];
                }
                c->htop[nzp - 1] = 0.0;
            }
            break;
        }
        default: {
            const double btop = 50.0, bbot = 50.0;
            const double h = z[0] - z[nzp - 1];
            const double denb = 1.0 - std::exp(-2.0 * bbot * h);
            const double dent = 1.0 - std::exp(-2.0 * btop * h);
            for (int k = 0; k < nzp; ++k) {
                c->hbot[k] = (std::exp(-bbot * z[k]) - std::exp(bbot * (z[k] - 2.0 * h))) / denb;
                c->d2hbot[k] = std::pow(bbot, 2.0) * c->hbot[k];
                c->htop[k] = std::exp(btop * (z[k] - h)) * (1.0 - std::exp(-2.0 * btop * z[k])) / dent;
                c->d2htop[k] = std::pow(btop, 2.0) * c->htop[k];
            }
            break;
        }
    }
}

// filters.f90:396-448 z weights (az, bz, cz) for filter_q_3pt.
static void filter_z_weights(const lsb_ctx* c, double a, double b, double* w) {
    const int nzp = c->nzp;
    const std::vector<double>& z = c->zpts;
    double* az = w;
    double* bz = w + nzp;
    double* cz = w + 2 * nzp;
    for (int k = 0; k < nzp; ++k) {
        az[k] = a;
        bz[k] = b;
        cz[k] = a;
    }
    const int m = c->cfg.mapping;
    if (m == 1 || m == 2) {
        for (int k = 2; k <= nzp - 1; ++k) {
            double dltm = z[k - 1] - z[k - 2], dltp = z[k] - z[k - 1];
            az[k - 1] = a * 2.0 * dltm / (dltp + dltm);
            bz[k - 1] = a * 2.0 * (1. - dltm / dltp) + b;
            cz[k - 1] = a * 2.0 * dltm * dltm / (dltp * (dltp + dltm));
        }
        if (m == 1) {
            az[1] = 0.0;
            bz[1] = 2.0 * a + b;
            cz[1] = 0.0;
        }
    } else {
        for (int k = 2; k <= (nzp + 1) / 2 - 1; ++k) {
            double dltm = z[k - 1] - z[k - 2], dltp = z[k] - z[k - 1];
            az[k - 1] = a * 2.0 * dltm / (dltp + dltm);
            bz[k - 1] = a * 2.0 * (1. - dltm / dltp) + b;
            cz[k - 1] = a * 2.0 * dltm * dltm / (dltp * (dltp + dltm));
        }
        for (int k = (nzp + 1) / 2 + 1; k <= nzp - 1; ++k) {
            double dltm = z[k - 1] - z[k - 2], dltp = z[k] - z[k - 1];
            az[k - 1] = a * 2.0 * dltp * dltp / (dltm * (dltm + dltp));
            bz[k - 1] = a * 2.0 * (1. - dltp / dltm) + b;
            cz[k - 1] = a * 2.0 * dltp / (dltm + dltp);
        }
        int mid = (nzp + 1) / 2 - 1;
        az[mid] = a;
        bz[mid] = b;
        cz[mid] = a;
    }
    az[0] = 0.0;
    bz[0] = 2.0 * a + b;
    cz[0] = 0.0;
    az[nzp - 1] = 0.0;
    bz[nzp - 1] = 2.0 * a + b;
    cz[nzp - 1] = 0.0;
}

// Operators that depend on the grid: Chebyshev matrices, dfdz/d2fdz2 operators,
// eigen-decomposition of d2(2:nz,2:nz) (init.f90:276-293).
int build_operators(lsb_ctx* c) {
    const int nzp = c->nzp, nz = c->nz, nzm = c->nzm;
    c->dor.assign((size_t)nzp * nzp, 0.0);
    c->d.assign((size_t)nzp * nzp, 0.0);
    c->d2.assign((size_t)nzp * nzp, 0.0);
    cheb_matrix(nzp, c->xl, c->gf.data(), c->dor.data(), c->d.data(), c->d2.data());
    c->diag.assign(nzm, 0.0);
    c->smat.assign((size_t)nzm * nzm, 0.0);
    c->simat.assign((size_t)nzm * nzm, 0.0);
    std::vector<double> d2i((size_t)nzm * nzm);
    for (int j = 0; j < nzm; ++j)
        for (int i = 0; i < nzm; ++i) d2i[i + (size_t)nzm * j] = c->d2[(i + 1) + (size_t)nzp * (j + 1)];
    int rc = eigen_decompose(d2i.data(), nzm, c->diag.data(), c->smat.data(), c->simat.data());
    if (rc != LSB_OK) return rc;
    (void)nz;
    return build_derived(c);
}

// Quantities derived from grid + mats (recomputed after lsb_set_grid/lsb_set_mats).
int build_derived(lsb_ctx* c) {
    const int nzp = c->nzp, nz = c->nz;
    c->mdz.assign((size_t)nzp * nzp, 0.0);
    c->md2.assign((size_t)nzp * nzp, 0.0);
    if (nz > 31) {  // derivatives.f90:122,175
        dct_diff_matrix(nzp, c->xl, c->gf.data(), 1, c->mdz.data(), c->cfg.dct_exact);
        dct_diff_matrix(nzp, c->xl, c->gf.data(), 2, c->md2.data(), c->cfg.dct_exact);
    } else {
        c->mdz = c->d;
        c->md2 = c->d2;
    }
    lifting_functions(c);
    c->fz_ab.assign(3 * nzp, 0.0);
    c->fz_11.assign(3 * nzp, 0.0);
    c->fz_16.assign(3 * nzp, 0.0);
    c->fz_12.assign(3 * nzp, 0.0);
    filter_z_weights(c, c->cfg.a, c->cfg.b, c->fz_ab.data());
    filter_z_weights(c, 1.0, 1.0, c->fz_11.data());
    // the two filters of the automatic TNS filtering (filters.f90:98-100)
    filter_z_weights(c, 1.0, 6.0, c->fz_16.data());
    filter_z_weights(c, 1.0, 2.0, c->fz_12.data());
    // SGS filter widths
    const std::vector<double>& z = c->zpts;
    const double dx = c->dx, dy = c->dy;
    c->delta_static.assign(nzp, 0.0);
    c->delta_dyn.assign(nzp, 0.0);
    c->delta_sigma.assign(nzp, 0.0);
    {   // sgsterm.f90:96-103 (Q8: delta(nzp) is never set; 0 here, multiplied by a zero damping)
        double dz = std::fabs(z[0] - z[1]);
        c->delta_static[0] = std::pow(dx * dy * dz, 1. / 3.);
        for (int k = 2; k <= nz; ++k) {
            dz = (z[k - 2] - z[k]) / 2.;
            c->delta_static[k - 1] = std::pow(dx * dy * dz, 1. / 3.);
        }
        dz = std::fabs(z[nzp - 1] - z[nzp - 2]);
        c->delta_static[0] = std::pow(dx * dy * dz, 1. / 3.);
        c->delta_static[nzp - 1] = 0.0;
    }
    {   // sgsterm.f90:159-164
        c->delta_dyn[0] = std::pow(dx * dy * (z[0] - z[1]), 1. / 3.);
        for (int k = 2; k <= nzp - 1; ++k) {
            double dz = 0.5 * (z[k - 2] - z[k]);
            c->delta_dyn[k - 1] = std::pow(dx * dy * dz, 1. / 3.);
        }
        c->delta_dyn[nzp - 1] = std::pow(dx * dy * (z[nzp - 2] - z[nzp - 1]), 1. / 3.);
    }
    {   // sgsterm.f90:368-376
        for (int k = 0; k < nzp; ++k) {
            double dz;
            if (k == 0)
                dz = z[0] - z[1];
            else if (k == nzp - 1)
                dz = z[nz - 1] - z[nzp - 1];
            else
                dz = 0.5 * (z[k - 1] - z[k + 1]);
            c->delta_sigma[k] = std::pow(dx * dy * dz, 1. / 3.);
        }
    }
    if (c->cfg.nut_ceil0) {   // no eddy viscosity on the ceiling row, as on the wall row
        c->delta_static[0] = 0.0;
        c->delta_sigma[0] = 0.0;
    }
    return LSB_OK;
}

// navierstokes.f90:119-187 force(), ftype = 1: the fringe's Rayleigh-damping coefficient.
// kappa = 2 u0/dxf (-ln 1e-5); r(i) = kappa (s0 + s): s0 = exp(-eta0^2) ramps up over dxs
// (1 on the plateau, 0 once eta >= 0), s = exp(-eta^2) decays over the left part dxl;
// the profile is then rotated right by xshift, the last i <= nx/4 (1-based) with
// xpts(i) < xpts(1) + dxl/1.5.
void sponge_lambda(int nx, const double* xpts, double xlen, double u0, double fsize, double lfsize, double sfsize,
                   double* lam) {
    const double epsx = 1.0e-5;
    const double sfs = sfsize * fsize, lfs = lfsize * fsize;
    const double rfsize = fsize - sfs - lfs;
    const double dxs = sfs * xlen, dxl = lfs * xlen, dxr = rfsize * xlen;
    const double dxf = dxs + dxl + dxr;
    const double kappa = 2.0 * u0 / dxf * (-std::log(epsx));
    std::vector<double> r(nx);
    for (int i = 0; i < nx; ++i) {
        const double eta0 = (xpts[i] - (xpts[nx - 1] - dxf + dxs)) / (dxs / 2.0);
        const double eta = (xpts[i] - (xpts[nx - 1] - dxl)) / (dxl / 2.5);
        double s0 = std::exp(-(eta0 * eta0));
        double s = std::exp(-(eta * eta));
        if (eta0 > 0.0 && eta < 0.0) s0 = 1.0;
        if (eta < 0.0) s = 0.0;
        if (eta >= 0.0) s0 = 0.0;
        r[i] = kappa * (s0 + s);
    }
    int xshift = 0;
    for (int i = 1; i <= nx / 4; ++i)
        if (xpts[i - 1] < xpts[0] + dxl / 1.5) xshift = i;
    if (xshift > 1) {
        for (int i = 0; i < nx - xshift; ++i) lam[xshift + i] = r[i];
        for (int i = 0; i < xshift; ++i) lam[i] = r[nx - xshift + i];
    } else {
        for (int i = 0; i < nx; ++i) lam[i] = r[i];
    }
}

}  // namespace lsb

// ============================================ Orr-Sommerfeld mode (orrsomm.f90:110-562)
#include <complex>

namespace lsb {
namespace {

using cplx = std::complex<double>;

// LU with partial pivoting of an n x n row-major complex matrix (in place); false if singular
bool clu(std::vector<cplx>& a, int n, std::vector<int>& piv) {
    piv.resize(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::abs(a[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (std::abs(a[(size_t)i * n + k]) > best) {
                best = std::abs(a[(size_t)i * n + k]);
                p = i;
            }
        piv[k] = p;
        if (best == 0.0) return false;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(a[(size_t)k * n + j], a[(size_t)p * n + j]);
        const cplx inv = 1.0 / a[(size_t)k * n + k];
        for (int i = k + 1; i < n; ++i) {
            const cplx f = a[(size_t)i * n + k] * inv;
            a[(size_t)i * n + k] = f;
            for (int j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= f * a[(size_t)k * n + j];
        }
    }
    return true;
}

// solve (L U) x = P b  (trans = false) or (P^T L U)^H x = b  (trans = true), x in b
void clu_solve(const std::vector<cplx>& lu, int n, const std::vector<int>& piv, std::vector<cplx>& b, bool herm) {
    if (!herm) {
        for (int k = 0; k < n; ++k) std::swap(b[k], b[piv[k]]);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j) b[i] -= lu[(size_t)i * n + j] * b[j];
        for (int i = n - 1; i >= 0; --i) {
            for (int j = i + 1; j < n; ++j) b[i] -= lu[(size_t)i * n + j] * b[j];
            b[i] /= lu[(size_t)i * n + i];
        }
    } else {
        // A = P^T L U  ->  A^H = U^H L^H P: solve U^H y = b, L^H z = y, x = P^T z
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j < i; ++j) b[i] -= std::conj(lu[(size_t)j * n + i]) * b[j];
            b[i] /= std::conj(lu[(size_t)i * n + i]);
        }
        for (int i = n - 1; i >= 0; --i)
            for (int j = i + 1; j < n; ++j) b[i] -= std::conj(lu[(size_t)j * n + i]) * b[j];
        for (int k = n - 1; k >= 0; --k) std::swap(b[k], b[piv[k]]);
    }
}

std::vector<cplx> cmatvec(const std::vector<cplx>& a, int n, const std::vector<cplx>& x) {
    std::vector<cplx> y(n, 0.0);
    for (int i = 0; i < n; ++i) {
        cplx s = 0.0;
        for (int j = 0; j < n; ++j) s += a[(size_t)i * n + j] * x[j];
        y[i] = s;
    }
    return y;
}

double l1(const std::vector<cplx>& x) {
    double s = 0.0;
    for (const cplx& v : x) s += std::abs(v);
    return s;
}

}  // namespace

// Column-major (Fortran) nzp x nzp d, d2; returns LSB_E_NUMERIC when the iteration fails the
// way the reference answers with a random restart.
int orrsomm_solve(int nzp, const double* d, const double* d2, const double* ubar, const double* ubarp,
                  const double* ubarpp, const double* zpts, double xnu, double eps, double alphax, double betax,
                  double* omega, double* uvec, double* vvec, double* wvec) {
    const int n = nzp + 2;
    const double c01 = 1.0e-40, c02 = 1.0e-40, c03 = 1.0e-40, c04 = 1.0e-40;
    const double rey = 1.0 / xnu;
    const double wav2 = alphax * alphax + betax * betax;
    const cplx I(0.0, 1.0);
    auto D = [&](int i, int j) { return d[(size_t)j * nzp + i]; };
    auto D2 = [&](int i, int j) { return d2[(size_t)j * nzp + i]; };
    std::vector<double> d4((size_t)nzp * nzp, 0.0);  // row-major d2 d2
    for (int i = 0; i < nzp; ++i)
        for (int j = 0; j < nzp; ++j) {
            double s = 0.0;
            for (int k = 0; k < nzp; ++k) s += D2(i, k) * D2(k, j);
            d4[(size_t)i * nzp + j] = s;
        }
    std::vector<cplx> a((size_t)n * n, 0.0), b((size_t)n * n, 0.0);
    for (int i = 0; i < nzp; ++i)
        for (int j = 0; j < nzp; ++j) {
            const double did = (i == j) ? 1.0 : 0.0;
            a[(size_t)i * n + j] = d4[(size_t)i * nzp + j] - 2.0 * wav2 * D2(i, j) + wav2 * wav2 * did -
                                   I * alphax * rey * (ubar[i] * (D2(i, j) - wav2 * did) - ubarpp[i] * did);
            b[(size_t)i * n + j] = -I * rey * (D2(i, j) - wav2 * did);
        }
    for (int j = 0; j < nzp; ++j) {  // membrane bottom-wall row (orrsomm.f90:189-194)
        cplx s = 0.0;
        for (int k = 0; k < nzp; ++k) s = s - I * c01 / (alphax * rey) * D2(nzp - 1, k) * D(k, j);
        a[(size_t)(n - 1) * n + j] = s;
    }
    for (int r : {0, 1, 2, n - 2, n - 1})
        for (int j = 0; j < n; ++j) b[(size_t)r * n + j] = 0.0;
    for (int r : {0, 2, n - 2})
        for (int j = 0; j < n; ++j) a[(size_t)r * n + j] = 0.0;
    for (int j = 0; j < nzp; ++j) a[(size_t)1 * n + j] = D(nzp - 1, j);
    a[0] = 1.0;
    a[(size_t)2 * n + nzp - 1] = 1.0;
    b[(size_t)2 * n + n - 2] = 1.0 / alphax;
    a[(size_t)1 * n + n - 2] = ubarp[nzp - 1];
    a[(size_t)1 * n + n - 1] = 0.0;
    a[(size_t)(n - 2) * n + n - 1] = 1.0;
    b[(size_t)(n - 2) * n + n - 2] = -I;
    a[(size_t)(n - 1) * n + n - 2] = c03 + c04 * wav2 - I * wav2 * (c02 / rey) * (ubarp[nzp - 1] / alphax);
    a[(size_t)(n - 1) * n + n - 1] = c01;
    b[(size_t)(n - 1) * n + n - 1] = I;
    // B^H
    std::vector<cplx> bh((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) bh[(size_t)i * n + j] = std::conj(b[(size_t)j * n + i]);
    // two-sided inverse iteration with Rayleigh-quotient shifts (orrsomm.f90:336-498)
    cplx omk(omega[0], omega[1]);
    std::vector<cplx> ph(n, 1.0), ga(n, 1.0), phk1 = ph, gak1 = ga;
    std::vector<int> piv;
    auto ck_of = [&](cplx om) {
        std::vector<cplx> ck((size_t)n * n);
        for (size_t q = 0; q < ck.size(); ++q) ck[q] = a[q] - om * b[q];
        return ck;
    };
    for (int outer = 0; outer < 10; ++outer) {
        const std::vector<cplx> ck = ck_of(omk);
        const double rk = l1(cmatvec(ck, n, ph));
        std::vector<cplx> lu = ck;
        if (!clu(lu, n, piv)) {
            set_error("solve_orrsomm: A - omega B is singular");
            return LSB_E_NUMERIC;
        }
        for (int ii = 1; ii <= 10; ++ii) {
            const double inner_rk = l1(cmatvec(ck, n, ph));
            std::vector<cplx> rphi = cmatvec(b, n, ph), rgam = cmatvec(bh, n, ga);
            clu_solve(lu, n, piv, rphi, false);
            clu_solve(lu, n, piv, rgam, true);
            const double nphi = l1(rphi), ngam = l1(rgam);
            for (int q = 0; q < n; ++q) {
                phk1[q] = rphi[q] / nphi;
                gak1[q] = rgam[q] / ngam;
            }
            const double inner_rk1 = l1(cmatvec(ck, n, phk1));
            double dphi = 0.0;
            for (int q = 0; q < n; ++q) dphi += std::abs(phk1[q] - ph[q]);
            if (dphi < 5.0e-5 || inner_rk1 < 5.0e-6) {
                if (inner_rk1 > inner_rk) {
                    phk1 = ph;
                    gak1 = ga;
                }
                ph = phk1;
                ga = gak1;
                break;
            }
            // (a grown residual makes the reference add random noise, orrsomm.f90:411-425;
            // the iteration goes on without it)
            ph = phk1;
            ga = gak1;
        }
        const std::vector<cplx> aph = cmatvec(a, n, phk1), bph = cmatvec(b, n, phk1);
        cplx num = 0.0, den = 0.0;
        for (int q = 0; q < n; ++q) {
            num += std::conj(gak1[q]) * aph[q];
            den += std::conj(gak1[q]) * bph[q];
        }
        cplx omk1 = num / den;
        const double rk1 = l1(cmatvec(ck_of(omk1), n, phk1));
        if (std::abs(omk1 - omk) < 1.0e-14 && rk1 < 1.5e-6) {
            if (rk1 >= rk) {
                omk1 = omk;
                phk1 = ph;
                gak1 = ga;
            }
            omk = omk1;
            break;
        }
        if (std::abs(omk1 - omk) < 1.0e-14) {
            // the Rayleigh quotient has stopped moving but the residual is above 1.5e-6, a
            // test near the rounding floor of (A - omega B) phi: the reference restarts from
            // a random omega here (orrsomm.f90:476-483); the restatement keeps the mode
            omk = omk1;
            break;
        }
        if (rk1 >= rk) {
            omk1 = omk;
            phk1 = ph;
            gak1 = ga;
        }
        omk = omk1;
    }
    // streamfunction -> (u, v, w) = (D phi, 0, -i alpha phi), energy eps (orrsomm.f90:537-548)
    std::vector<cplx> w(phk1.begin(), phk1.begin() + nzp), u(nzp, 0.0), v(nzp, 0.0);
    for (int i = 0; i < nzp; ++i) {
        cplx s = 0.0;
        for (int j = 0; j < nzp; ++j) s += D(i, j) * w[j];
        u[i] = s;
    }
    for (int i = 0; i < nzp; ++i) w[i] = -I * alphax * w[i];
    auto energy = [&](const std::vector<cplx>& q) {  // orrsomm.f90:669-682
        double e = 0.0;
        for (int k = 1; k < nzp - 1; ++k) {
            const double a1 = std::abs(q[k + 1]), a0 = std::abs(q[k]);
            e = e - 0.5 * (a1 * a1 + a0 * a0) * (zpts[k + 1] - zpts[k]);
        }
        return e * 2.0 / 7.;
    };
    const double etot = energy(u) + energy(v) + energy(w);
    const double xn = std::sqrt(eps / etot);
    for (int i = 0; i < nzp; ++i) {
        uvec[2 * i] = (u[i] * xn).real();
        uvec[2 * i + 1] = (u[i] * xn).imag();
        vvec[2 * i] = (v[i] * xn).real();
        vvec[2 * i + 1] = (v[i] * xn).imag();
        wvec[2 * i] = (w[i] * xn).real();
        wvec[2 * i + 1] = (w[i] * xn).imag();
    }
    omega[0] = omk.real();
    omega[1] = omk.imag();
    return LSB_OK;
}

}  // namespace lsb
