/*
 * FP64 CPU cube for the oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * Same arithmetic as oracle/martini_oracle.py (which restates the reference),
 * organised particle by particle so that million-particle cubes finish in
 * seconds.  Used only by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg; never linked into the product.
 *
 *   cube[i][j][c] += sum_p W_p(x_p - i, y_p - j) * S_p[c]
 *
 * over the square mask |i - x_p| <= r_p and |j - y_p| <= r_p
 * (martini/martini.py:271-281), with W the kernel integral of
 * martini/sph_kernels.py (cubic :791-858, WendlandC2 :400-441,
 * WendlandC6 :544-685, quartic :1486-1564, Gaussian :994-1044,
 * Dirac :1136-1165) and S the spectrum of spectral_models.py
 * (Gaussian :347-425, Dirac :510-570) times mHI/d^2/|dv|/2.36e5 (:119-145).
 *
 * Threads: each thread owns a private cube and a fixed contiguous slice of
 * the particles; the private cubes are summed in thread order, so the result
 * does not depend on scheduling.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { K_CUBIC = 0, K_WC2 = 1, K_WC6 = 2, K_QUARTIC = 3, K_GAUSS = 4, K_DIRAC = 5 };

static const double GSIG = 0.42466090014400953; /* 1/(2 sqrt(2 ln 2)) */

static double w_cubic(double dx, double dy, double h) {
    double R2 = 4.0 * (dx * dx + dy * dy) / (h * h), F;
    if (R2 == 0.0) {
        F = 11.0 / 16.0 + 0.25 * 0.25;
    } else if (R2 <= 1.0) {
        double A = sqrt(1.0 - R2), B = sqrt(4.0 - R2);
        double I1 = A - 0.5 * A * A * A - 1.5 * R2 * A + 3.0 / 32.0 * A * (3.0 * R2 + 2.0) +
                    9.0 / 32.0 * R2 * R2 * (log(1.0 + A) - log(sqrt(R2)));
        double I3 = -B * (3.0 * R2 + 56.0) / 4.0 + A * (4.0 * R2 + 50.0) / 8.0 -
                    3.0 / 8.0 * R2 * (R2 + 16.0) * log((2.0 + B) / (1.0 + A)) +
                    2.0 * (3.0 * R2 + 4.0) * (B - A) + 2.0 * (B * B * B - A * A * A);
        F = I1 + 0.25 * I3;
    } else if (R2 <= 4.0) {
        double B = sqrt(4.0 - R2);
        double I2 = -B * (3.0 * R2 + 56.0) / 4.0 -
                    3.0 / 8.0 * R2 * (R2 + 16.0) * log((2.0 + B) / sqrt(R2)) +
                    2.0 * (3.0 * R2 + 4.0) * B + 2.0 * B * B * B;
        F = 0.25 * I2;
    } else {
        F = 0.0;
    }
    return F / 1.59689476201133 / (h * h) * 4.0;
}

static double w_wc2(double dx, double dy, double h) {
    double R2 = (dx * dx + dy * dy) / (h * h), G;
    if (R2 == 0.0) {
        G = 2.0 / 3.0;
    } else if (R2 < 1.0) {
        double A = sqrt(1.0 - R2);
        G = 5.0 * R2 * R2 * (0.5 * R2 + 3.0) * log((1.0 + A) / sqrt(R2)) +
            A * (-27.0 / 2.0 * R2 * R2 - 14.0 / 3.0 * R2 + 2.0 / 3.0);
    } else {
        G = 0.0;
    }
    return G * (21.0 / 2.0 / M_PI) / (h * h);
}

/* WendlandC6: coefficient table of the reference's indefinite integral
 * (sph_kernels.py:576-678); columns: coefficient, power of R, power of z,
 * factor (0 plain, 1 sqrt(R^2+z^2), 2 log(sqrt(R^2+z^2)+z)). */
static const double WC6_TERMS[][4] = {
    {-231, 10, 1, 0}, {-385, 8, 3, 0}, {-1155, 8, 1, 0}, {-462, 6, 5, 0},
    {-1540, 6, 3, 0}, {-462, 6, 1, 0}, {-330, 4, 7, 0}, {-1386, 4, 5, 0},
    {-462, 4, 3, 0}, {66, 4, 1, 0}, {-(128 + 1.0 / 3.0), 2, 9, 0}, {-660, 2, 7, 0},
    {-277.2, 2, 5, 0}, {44, 2, 3, 0}, {8.0 / 3.0, 0, 11, 1}, {16 + 4.0 / 15.0, 2, 9, 1},
    {70.4, 0, 9, 1}, {360.8, 2, 7, 1}, {132, 0, 7, 1}, {550, 2, 5, 1},
    {-11, 2, 1, 0}, {7.21875, 12, 0, 2}, {24.7813, 10, 1, 1}, {173.25, 10, 0, 2},
    {530.75, 8, 1, 1}, {288.75, 8, 0, 2}, {47.4792, 8, 3, 1}, {767.25, 6, 1, 1},
    {58.0167, 6, 5, 1}, {819.5, 6, 3, 1}, {41.7, 4, 7, 1}, {752.4, 4, 5, 1},
    {896.5, 4, 3, 1}, {-21, 0, 11, 0}, {-(128 + 1.0 / 3.0), 0, 9, 0}, {-66, 0, 7, 0},
    {13.2, 0, 5, 0}, {-11.0 / 3.0, 0, 3, 0}, {1, 0, 1, 0},
};

static double wc6_indef(double R, double z) {
    double S = sqrt(R * R + z * z), L = log(S + z), tot = 0.0;
    for (size_t k = 0; k < sizeof(WC6_TERMS) / sizeof(WC6_TERMS[0]); ++k) {
        double t = WC6_TERMS[k][0] * pow(R, WC6_TERMS[k][1]) * pow(z, WC6_TERMS[k][2]);
        if (WC6_TERMS[k][3] == 1) t *= S;
        else if (WC6_TERMS[k][3] == 2) t *= L;
        tot += t;
    }
    return tot;
}

static double w_wc6(double dx, double dy, double h) {
    double R = sqrt(dx * dx + dy * dy) / h, norm = 1365.0 / 64.0 / M_PI, v;
    if (R == 0.0) v = norm * 2.0 * (4.0 / 15.0);
    else if (R < 1.0) v = norm * 2.0 * (wc6_indef(R, sqrt(1.0 - R * R)) - wc6_indef(R, 0.0));
    else v = 0.0;
    return v / (h * h);
}

static double quartic_IA(double R, double z, double A) {
    double q = sqrt(z * z + R * R), R2 = R * R;
    return pow(A, 4) * z - 2 * pow(A, 3) * z * q + 2 * A * A * z * (3 * R2 + z * z) -
           A * R2 * (4 * A * A + 3 * R2) * asinh(z / R) / 2 - A * z * q * (5 * R2 + 2 * z * z) / 2 +
           R2 * R2 * z + 2 * R2 * z * z * z / 3 + pow(z, 5) / 5;
}

static double w_quartic(double dx, double dy, double h) {
    double R = sqrt(dx * dx + dy * dy) / h, v = 0.0;
    if (R == 0.0) {
        v = 384.0 / 3125.0;
    } else {
        if (R < 0.2) v += 10.0 * quartic_IA(R, sqrt(0.04 - R * R), 0.2);
        if (R < 0.6) v -= 5.0 * quartic_IA(R, sqrt(0.36 - R * R), 0.6);
        if (R < 1.0) v += quartic_IA(R, sqrt(1.0 - R * R), 1.0);
    }
    v *= 2.0 * 15625.0 / 512.0 / M_PI;
    return v / (h * h);
}

static double w_gauss(double dx, double dy, double h, double t, double norm) {
    double dr = sqrt(dx * dx + dy * dy), u = dr / h / GSIG;
    double zmax = (t > u) ? sqrt(t * t - u * u) : 0.0;
    double c = h * sqrt(2.0) * GSIG;
    double v = 0.25 * erf(zmax / sqrt(2.0)) * (erf((dx + 0.5) / c) - erf((dx - 0.5) / c)) *
               (erf((dy + 0.5) / c) - erf((dy - 0.5) / c));
    if ((dr - sqrt(0.5)) / h / GSIG > t) v = 0.0;
    return v / norm;
}

double oracle_weight(int kind, double dx, double dy, double h, double t, double norm) {
    switch (kind) {
    case K_CUBIC: return w_cubic(dx, dy, h);
    case K_WC2: return w_wc2(dx, dy, h);
    case K_WC6: return w_wc6(dx, dy, h);
    case K_QUARTIC: return w_quartic(dx, dy, h);
    case K_GAUSS: return w_gauss(dx, dy, h, t, norm);
    case K_DIRAC: return (fabs(dx) < 0.5 && fabs(dy) < 0.5) ? 1.0 : 0.0;
    }
    return 0.0;
}

/* Spectrum row of one particle (Jy): edges in km/s, monotonic. */
static void spectrum_row(double v, double sigma, double amp, const double* edges, int nc,
                         int dirac, double* out) {
    int decreasing = edges[1] < edges[0];
    for (int c = 0; c < nc; ++c) {
        double lo = decreasing ? edges[c + 1] : edges[c];
        double hi = decreasing ? edges[c] : edges[c + 1];
        double phi;
        if (dirac) phi = ((v - lo) >= 0.0 ? 1.0 : 0.0) * ((hi - v) >= 0.0 ? 1.0 : 0.0);
        else phi = 0.5 * (erf((hi - v) / sqrt(2.0) / sigma) - erf((lo - v) / sqrt(2.0) / sigma));
        out[c] = amp * phi / fabs(edges[c + 1] - edges[c]) / 2.36e5;
    }
}

/*
 * n particles: x, y pixel coordinates; h_eff (h_px * rescale); r (mask range);
 * kind codes; gauss (t, norm) per particle (ignored unless kind == gauss);
 * v (km/s), sigma (km/s), amp = mHI / d^2.  edges: nc+1 velocity edges (km/s).
 * cube: nx*ny*nc, row-major [i][j][c], accumulated (+=).
 */
int oracle_cube(int64_t n, const double* x, const double* y, const double* h_eff,
                const double* r, const int32_t* kind, const double* gt, const double* gnorm,
                const double* v, const double* sigma, const double* amp, const double* edges,
                int nx, int ny, int nc, int dirac_spectrum, int nthreads, double* cube) {
    if (nthreads < 1) nthreads = 1;
    size_t cube_n = (size_t)nx * ny * nc;
    double* priv = (double*)calloc((size_t)nthreads * cube_n, sizeof(double));
    if (!priv) return -1;
#pragma omp parallel num_threads(nthreads)
    {
        int t = 0, nt = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        nt = omp_get_num_threads();
#endif
        double* mine = priv + (size_t)t * cube_n;
        double* S = (double*)malloc(sizeof(double) * nc);
        int64_t p0 = n * t / nt, p1 = n * (t + 1) / nt;
        for (int64_t p = p0; p < p1 && S; ++p) {
            spectrum_row(v[p], sigma[p], amp[p], edges, nc, dirac_spectrum, S);
            /* channels outside [c0, c1] hold exact zeros (erf saturates in FP64), and
             * adding w * 0 leaves a voxel unchanged: they are skipped */
            int c0 = 0, c1 = nc - 1;
            while (c0 <= c1 && S[c0] == 0.0) ++c0;
            while (c1 >= c0 && S[c1] == 0.0) --c1;
            if (c0 > c1) continue;
            int ilo = (int)ceil(x[p] - r[p]), ihi = (int)floor(x[p] + r[p]);
            int jlo = (int)ceil(y[p] - r[p]), jhi = (int)floor(y[p] + r[p]);
            if (ilo < 0) ilo = 0;
            if (jlo < 0) jlo = 0;
            if (ihi > nx - 1) ihi = nx - 1;
            if (jhi > ny - 1) jhi = ny - 1;
            for (int i = ilo; i <= ihi; ++i) {
                if (!(fabs(i - x[p]) <= r[p])) continue;
                for (int j = jlo; j <= jhi; ++j) {
                    if (!(fabs(j - y[p]) <= r[p])) continue;
                    double w = oracle_weight(kind[p], x[p] - i, y[p] - j, h_eff[p], gt[p], gnorm[p]);
                    if (w == 0.0) continue;
                    double* dst = mine + ((size_t)i * ny + j) * nc;
                    for (int c = c0; c <= c1; ++c) dst[c] += w * S[c];
                }
            }
        }
        free(S);
    }
    for (int t = 0; t < nthreads; ++t) {
        const double* src = priv + (size_t)t * cube_n;
        for (size_t k = 0; k < cube_n; ++k) cube[k] += src[k];
    }
    free(priv);
    return 0;
}

/* Vectorised weight evaluation for tests. */
void oracle_weights(int64_t n, int kind, const double* dx, const double* dy, const double* h,
                    double t, double norm, double* out) {
    for (int64_t k = 0; k < n; ++k) out[k] = oracle_weight(kind, dx[k], dy[k], h[k], t, norm);
}
