/*
 * csx_oracle.c -- FP64 CPU restatement of csim's E-field propagation path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP product and the
 * timed CPU baseline.  Product code never links or calls this file.
 *
 * The reference (stevepur/csim) cannot be built here (FFTW, Armadillo, nlopt
 * and BLAS headers are absent), so this is an independent restatement of its
 * algorithms.  Each function cites the reference file:line it follows.
 * Third-party arithmetic restated from published algorithms:
 *   - FFTW 3.3.7 transforms (FFTW_FORWARD = e^{-2 pi i jk/n}, unnormalised):
 *     replaced by an own iterative radix-2 FFT (power-of-two sizes) with an
 *     exact-twiddle DFT fallback for other sizes;
 *   - Armadillo 8.500.1 interp1(..., "*linear", extrap): nearest-node search
 *     plus linear weight (restated in orc_interp1_linear).
 * Column-major storage (Armadillo): element (r, c) at c*rows + r.
 */
#include "csx_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.14159265358979323846

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ FFT -- */
/* Twiddle cache: tw[k] = exp(sign * 2 pi i k / n), k < n/2, computed directly
 * with cos/sin so every twiddle is correctly rounded to within 1 ulp. */
typedef struct { int n, sign; ocx* tw; int* rev; } fft_plan_t;
#define MAX_PLANS 64
static fft_plan_t g_plans[MAX_PLANS];
static int g_nplans = 0;

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static const fft_plan_t* get_plan(int n, int sign) {
    const fft_plan_t* found = NULL;
#pragma omp critical(orc_fft_plan)
    {
        for (int i = 0; i < g_nplans; i++)
            if (g_plans[i].n == n && g_plans[i].sign == sign) { found = &g_plans[i]; break; }
        if (!found && g_nplans < MAX_PLANS) {
            fft_plan_t* p = &g_plans[g_nplans];
            p->n = n; p->sign = sign;
            p->tw = (ocx*)malloc(sizeof(ocx) * (size_t)(n / 2 + 1));
            p->rev = (int*)malloc(sizeof(int) * (size_t)n);
            for (int k = 0; k <= n / 2; k++) {
                double ang = sign * 2.0 * PI * (double)k / (double)n;
                p->tw[k] = cos(ang) + I * sin(ang);
            }
            int lg = 0; while ((1 << lg) < n) lg++;
            for (int i = 0; i < n; i++) {
                int r = 0;
                for (int b = 0; b < lg; b++) if (i & (1 << b)) r |= 1 << (lg - 1 - b);
                p->rev[i] = r;
            }
            g_nplans++;
            found = p;
        }
    }
    return found;
}

static void dft_naive(ocx* x, int n, int sign) {
    ocx* tmp = (ocx*)malloc(sizeof(ocx) * (size_t)n);
    for (int k = 0; k < n; k++) {
        ocx s = 0;
        for (int j = 0; j < n; j++) {
            long long jk = ((long long)j * k) % n;      /* exact argument reduction */
            double ang = sign * 2.0 * PI * (double)jk / (double)n;
            s += x[j] * (cos(ang) + I * sin(ang));
        }
        tmp[k] = s;
    }
    memcpy(x, tmp, sizeof(ocx) * (size_t)n);
    free(tmp);
}

void orc_fft1d(ocx* x, int n, int sign) {
    if (n <= 1) return;
    if (!is_pow2(n)) { dft_naive(x, n, sign); return; }
    const fft_plan_t* p = get_plan(n, sign);
    for (int i = 0; i < n; i++) {
        int r = p->rev[i];
        if (r > i) { ocx t = x[i]; x[i] = x[r]; x[r] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int j = 0; j < half; j++) {
                ocx w = p->tw[j * step];
                ocx u = x[i + j], v = x[i + j + half] * w;
                x[i + j] = u + v;
                x[i + j + half] = u - v;
            }
        }
    }
}

/* 2-D transform of a column-major rows x cols matrix (FFTW plan_dft_2d(cols, rows),
 * src/lib/csim_fft.cpp:111): 1-D transforms down every column, then along every row. */
void orc_fft2d(ocx* a, int rows, int cols, int sign) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < cols; c++) orc_fft1d(a + (size_t)c * rows, rows, sign);
#pragma omp parallel
    {
        ocx* buf = (ocx*)malloc(sizeof(ocx) * (size_t)cols);
#pragma omp for schedule(static)
        for (int r = 0; r < rows; r++) {
            for (int c = 0; c < cols; c++) buf[c] = a[(size_t)c * rows + r];
            orc_fft1d(buf, cols, sign);
            for (int c = 0; c < cols; c++) a[(size_t)c * rows + r] = buf[c];
        }
        free(buf);
    }
}

/* fft::execute (src/lib/csim_fft.cpp:98-160): forward, unnormalised. */
void orc_fft2_csim(ocx* a, int rows, int cols) { orc_fft2d(a, rows, cols, -1); }

/* ifft::execute (src/lib/csim_fft.cpp:165-228): backward, divided by n_elem (:193,226). */
void orc_ifft2_csim(ocx* a, int rows, int cols) {
    orc_fft2d(a, rows, cols, +1);
    double s = (double)rows * (double)cols;
    size_t n = (size_t)rows * cols;
    for (size_t i = 0; i < n; i++) a[i] /= s;
}

/* fft_shift (src/lib/csim_fft.cpp:296-320): circular shift by floor(n/2) down
 * the rows, then along the columns: out(r,c) = in(r - R/2 mod R, c - C/2 mod C). */
void orc_fft_shift(const ocx* in, ocx* out, int rows, int cols) {
    int sr = rows / 2, sc = cols / 2;
    for (int c = 0; c < cols; c++)
        for (int r = 0; r < rows; r++)
            out[(size_t)((c + sc) % cols) * rows + (r + sr) % rows] = in[(size_t)c * rows + r];
}
void orc_fft_shift_real(const double* in, double* out, int rows, int cols) {
    int sr = rows / 2, sc = cols / 2;
    for (int c = 0; c < cols; c++)
        for (int r = 0; r < rows; r++)
            out[(size_t)((c + sc) % cols) * rows + (r + sr) % rows] = in[(size_t)c * rows + r];
}

/* ------------------------------------------------------------ chirp-z -- */
/* nfft = 2^ceil(log2(m+k-1)) exactly as src/lib/csim_fft.cpp:464. */
int orc_czt_nfft(int m, int k) { return (int)pow(2, ceil(log2((double)(m + k - 1)))); }

/* czt::czt2d 1-D pass along columns (src/lib/csim_fft.cpp:442-665):
 * out(q, c) = sum_n in(n, c) exp(-i n (a - w q)), q < k, via Bluestein. */
void orc_czt1d_cols(const ocx* in, int m, int ncols, int k, double w, double a, ocx* out) {
    int nfft = orc_czt_nfft(m, k);
    int mm = m >= k ? m : k;
    int nww = m + mm - 1;
    double* ww = (double*)malloc(sizeof(double) * (size_t)nww);
    for (int i = 0; i < nww; i++) {                 /* :519-525 */
        double kk = (double)i - m + 1.0;
        ww[i] = w * (kk * kk / 2.0);
    }
    ocx* caa = (ocx*)malloc(sizeof(ocx) * (size_t)m);
    for (int i = 0; i < m; i++) {                   /* :530-541 */
        double aa = a * ((double)-i);
        caa[i] = cexp(I * (aa + ww[m - 1 + i]));
    }
    ocx* fv = (ocx*)calloc((size_t)nfft, sizeof(ocx));
    for (int j = 0; j <= m - 1 + k - 1; j++) fv[j] = cexp(-I * ww[j]);   /* :570 chirp filter */
    orc_fft1d(fv, nfft, -1);                        /* :573 */
    ocx* expww = (ocx*)malloc(sizeof(ocx) * (size_t)k);
    for (int i = 0; i < k; i++) expww[i] = cexp(I * ww[m - 1 + i]);     /* :637-644 */
#pragma omp parallel
    {
        ocx* y = (ocx*)malloc(sizeof(ocx) * (size_t)nfft);
#pragma omp for schedule(static)
        for (int c = 0; c < ncols; c++) {           /* :589-606 */
            const ocx* col = in + (size_t)c * m;
            for (int i = 0; i < m; i++) y[i] = col[i] * caa[i];
            for (int i = m; i < nfft; i++) y[i] = 0;
            orc_fft1d(y, nfft, -1);
            for (int i = 0; i < nfft; i++) y[i] *= fv[i];
            orc_fft1d(y, nfft, +1);
            for (int i = 0; i < k; i++)             /* g /= nfft; out = g(m-1..) % expww */
                out[(size_t)c * k + i] = (y[m - 1 + i] / (double)nfft) * expww[i];
        }
        free(y);
    }
    free(ww); free(caa); free(fv); free(expww);
}

static void transpose(const ocx* in, int rows, int cols, ocx* out) {
#pragma omp parallel for schedule(static)
    for (int c = 0; c < cols; c++)
        for (int r = 0; r < rows; r++)
            out[(size_t)r * cols + c] = in[(size_t)c * rows + r];
}

/* czt::czt2d (src/lib/csim_fft.cpp:422-440): y pass with (wy, ay), transpose,
 * x pass with (wx, ax), transpose.  Output k x k. */
void orc_czt2d(const ocx* in, int rows, int cols, int k,
               double wx, double ax, double wy, double ay, ocx* out) {
    ocx* o1 = (ocx*)malloc(sizeof(ocx) * (size_t)k * cols);
    ocx* o1t = (ocx*)malloc(sizeof(ocx) * (size_t)k * cols);
    ocx* o2 = (ocx*)malloc(sizeof(ocx) * (size_t)k * k);
    orc_czt1d_cols(in, rows, cols, k, wy, ay, o1);      /* k x cols */
    transpose(o1, k, cols, o1t);                         /* cols x k */
    orc_czt1d_cols(o1t, cols, k, k, wx, ax, o2);         /* k x k */
    transpose(o2, k, k, out);
    free(o1); free(o1t); free(o2);
}

/* zoomFft::init(arrayGeom, arrayGeom, lambdaF) + zoomFft::execute
 * (src/lib/csim_fft.cpp:670-708, 808-826). */
void orc_zoomfft(const ocx* in, int rows, int cols,
                 double inX0, double inY0, double inDx, double inDy,
                 int nOut, double outX0, double outY0, double outDx, double outDy,
                 double lambdaF, ocx* out) {
    double Ax = 2 * PI * inDx * outX0 / lambdaF;        /* :685 */
    double Wx = -2 * PI * inDx * outDx / lambdaF;       /* :686 */
    double Ay = Ax, Wy = Wx;
    if ((inDx != inDy) | (outDx != outDy) | (outY0 != outX0)) {   /* :687-693 */
        Ay = 2 * PI * inDy * outY0 / lambdaF;
        Wy = -2 * PI * inDy * outDy / lambdaF;
    }
    orc_czt2d(in, rows, cols, nOut, Wx, Ax, Wy, Ay, out);
    ocx f1 = -2 * PI * I / lambdaF;                     /* :816-817 */
    ocx f2 = inDx * inDy / (I * lambdaF);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nOut; c++) {
        double xo = outX0 + c * outDx;
        for (int r = 0; r < nOut; r++) {
            double yo = outY0 + r * outDy;
            double mxPmy = inX0 * xo + inY0 * yo;       /* :698 */
            out[(size_t)c * nOut + r] *= cexp(f1 * mxPmy) * f2;
        }
    }
}

/* Phase e^{-2 pi i t} with t = p*q/lambdaF in cycles, reduced to [-1/2, 1/2]
 * before the trig call so the FP64 argument keeps full accuracy. */
static inline ocx cis_cycles(double t) {
    double f = t - nearbyint(t);
    double ang = -2.0 * PI * f;
    return cos(ang) + I * sin(ang);
}

/* The same transform as zoomFft in closed form: out = s * A_y * in * B_x^T with
 * A_y[r][m] = e^{-2 pi i y_out[r] y_in[m]/lf}, B_x[c][n] = e^{-2 pi i x_out[c] x_in[n]/lf},
 * s = dx dy / (i lf)  (derivation from src/lib/csim_fft.cpp:519-541,685-686,794-801). */
void orc_mft_direct(const ocx* in, int rows, int cols,
                    double inX0, double inY0, double inDx, double inDy,
                    int outRows, int outCols, double outX0, double outY0,
                    double outDx, double outDy, double lambdaF, ocx* out) {
    ocx* Ay = (ocx*)malloc(sizeof(ocx) * (size_t)outRows * rows);   /* [m][r] col-major outRows x rows */
    ocx* Bx = (ocx*)malloc(sizeof(ocx) * (size_t)outCols * cols);
    for (int m = 0; m < rows; m++)
        for (int r = 0; r < outRows; r++)
            Ay[(size_t)m * outRows + r] = cis_cycles((outY0 + r * outDy) * (inY0 + m * inDy) / lambdaF);
    for (int n = 0; n < cols; n++)
        for (int c = 0; c < outCols; c++)
            Bx[(size_t)n * outCols + c] = cis_cycles((outX0 + c * outDx) * (inX0 + n * inDx) / lambdaF);
    /* T = A_y * in  (outRows x cols) */
    ocx* T = (ocx*)calloc((size_t)outRows * cols, sizeof(ocx));
#pragma omp parallel for schedule(static)
    for (int n = 0; n < cols; n++) {
        ocx* tc = T + (size_t)n * outRows;
        const ocx* ic = in + (size_t)n * rows;
        for (int m = 0; m < rows; m++) {
            ocx v = ic[m];
            if (v == 0) continue;
            const ocx* ac = Ay + (size_t)m * outRows;
            for (int r = 0; r < outRows; r++) tc[r] += ac[r] * v;
        }
    }
    ocx s = inDx * inDy / (I * lambdaF);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < outCols; c++) {
        ocx* oc = out + (size_t)c * outRows;
        for (int r = 0; r < outRows; r++) oc[r] = 0;
        for (int n = 0; n < cols; n++) {
            ocx b = Bx[(size_t)n * outCols + c];
            const ocx* tc = T + (size_t)n * outRows;
            for (int r = 0; r < outRows; r++) oc[r] += tc[r] * b;
        }
        for (int r = 0; r < outRows; r++) oc[r] *= s;
    }
    free(Ay); free(Bx); free(T);
}

/* ------------------------------------------------------------ Fresnel -- */
/* fresnelPropagateAS::execute (src/lib/csim_fft.cpp:371-392), Goodman 4-20:
 * E <- IFFT2( FFT2(E) . exp(i k z sqrt(1 - (lambda/apRad)^2/4 (p^2+q^2))) ),
 * p, q the fftshifted integer frequency indices.  Square n x n only. */
void orc_fresnel_as(ocx* E, int n, double lambda, double apRad, double z) {
    double k = 2. * PI / lambda;
    double c4 = 1. / 4. * pow(lambda / apRad, 2.);
    orc_fft2_csim(E, n, n);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; c++) {
        /* H(r,c) after fft_shift: vH(i) = i - n/2 shifted by floor(n/2) */
        int pc = ((c - n / 2) % n + n) % n - n / 2;
        for (int r = 0; r < n; r++) {
            int pr = ((r - n / 2) % n + n) % n - n / 2;
            double H = (double)pc * pc + (double)pr * pr;
            ocx cH = cexp(I * k * z * sqrt(1. - c4 * H));
            E[(size_t)c * n + r] *= cH;
        }
    }
    orc_ifft2_csim(E, n, n);
}

/* ----------------------------------------------------- geometry/resample -- */
/* arrayGeom::set_xy symmetric grid (src/lib/csim_geom.cpp:113-140):
 * x_i = -(n d - d)/2 + i d - origin. */
void orc_grid_symmetric(int n, double d, double origin, double* x) {
    double phys = d * n;
    for (int i = 0; i < n; i++) x[i] = -(phys - d) / 2.0 + i * d - origin;
}
/* Offset grid (i - n/2) d used by the doubleFraunhofer FPM plane
 * (src/elements/doubleFraunhoferFPM.cpp:59-64). */
void orc_grid_offset(int n, double d, double* x) {
    for (int i = 0; i < n; i++) x[i] = ((double)i - ((double)n) / 2.) * d;
}

/* Armadillo 8.500.1 interp1(XG, YG, XI, YI, "*linear", extrap) restated:
 * XG, XI ascending; outside [min XG, max XG] -> extrap; otherwise the nearest
 * node a (searched forward from the previous one) and its neighbour b on the
 * other side of XI, weight = |XG[a]-XI| / (|XG[a]-XI| + |XG[b]-XI|). */
void orc_interp1_linear(const double* xg, const double* yg, int ng,
                        const double* xi, double* yi, int ni, double extrap) {
    double xmin = xg[0], xmax = xg[0];
    for (int i = 1; i < ng; i++) { if (xg[i] < xmin) xmin = xg[i]; if (xg[i] > xmax) xmax = xg[i]; }
    int a_idx = 0, b_idx = 0;
    for (int i = 0; i < ni; i++) {
        double v = xi[i];
        if (v < xmin || v > xmax) { yi[i] = extrap; continue; }
        double a_err = INFINITY, b_err;
        for (int j = a_idx; j < ng; j++) {
            double t = xg[j] - v;
            double e = t >= 0 ? t : -t;
            if (e >= a_err) break;
            a_err = e; a_idx = j;
        }
        if (xg[a_idx] - v <= 0) b_idx = (a_idx + 1 < ng) ? a_idx + 1 : a_idx;
        else b_idx = (a_idx >= 1) ? a_idx - 1 : a_idx;
        b_err = fabs(xg[b_idx] - v);
        int lo = a_idx, hi = b_idx;
        double lo_err = a_err, hi_err = b_err;
        if (lo > hi) { int t = lo; lo = hi; hi = t; double te = lo_err; lo_err = hi_err; hi_err = te; }
        double wgt = (lo_err > 0) ? lo_err / (lo_err + hi_err) : 0.0;
        yi[i] = (1.0 - wgt) * yg[lo] + wgt * yg[hi];
    }
}

/* interp2 (src/lib/csim_geom.cpp:244-266): along every row with the x grids,
 * then along every column with the y grids.  out is orows x ocols. */
void orc_interp2(const double* in, int rows, int cols, double* out, int orows, int ocols,
                 const double* inX, const double* inY, const double* outX, const double* outY) {
    double* s1 = (double*)calloc((size_t)rows * ocols, sizeof(double));
#pragma omp parallel
    {
        double* vin = (double*)malloc(sizeof(double) * (size_t)cols);
        double* vout = (double*)malloc(sizeof(double) * (size_t)ocols);
#pragma omp for schedule(static)
        for (int r = 0; r < rows; r++) {
            for (int c = 0; c < cols; c++) vin[c] = in[(size_t)c * rows + r];
            orc_interp1_linear(inX, vin, cols, outX, vout, ocols, 0.0);
            for (int c = 0; c < ocols; c++) s1[(size_t)c * rows + r] = vout[c];
        }
        free(vin); free(vout);
    }
#pragma omp parallel for schedule(static)
    for (int c = 0; c < ocols; c++)
        orc_interp1_linear(inY, s1 + (size_t)c * rows, rows, outY, out + (size_t)c * orows, orows, 0.0);
    free(s1);
}

/* downsample::execute (src/elements/downsample.cpp:23-96): new symmetric grid of
 * pixel factor*d and size n/factor; Re and Im interpolated separately. */
int orc_downsample(const ocx* in, int rows, int cols, int factor,
                   double pixelSizeX, double pixelSizeY, double originX, double originY,
                   ocx* out) {
    int nr = rows / factor, nc = cols / factor;
    /* efield geometry: set_geometry(E->n_cols, E->n_rows, ...) -> nX = cols, nY = rows */
    double* inX = (double*)malloc(sizeof(double) * (size_t)cols);
    double* inY = (double*)malloc(sizeof(double) * (size_t)rows);
    double* outX = (double*)malloc(sizeof(double) * (size_t)nc);
    double* outY = (double*)malloc(sizeof(double) * (size_t)nr);
    orc_grid_symmetric(cols, pixelSizeX, originX, inX);
    orc_grid_symmetric(rows, pixelSizeY, originY, inY);
    /* newGeom.set_geometry(newNRows, newNCols, ...) -> nX = newNRows, nY = newNCols */
    orc_grid_symmetric(nr, factor * pixelSizeX, originX, outX);
    orc_grid_symmetric(nc, factor * pixelSizeY, originY, outY);
    size_t n = (size_t)rows * cols, no = (size_t)nr * nc;
    double* re = (double*)malloc(sizeof(double) * n);
    double* im = (double*)malloc(sizeof(double) * n);
    double* ro = (double*)malloc(sizeof(double) * no);
    double* io = (double*)malloc(sizeof(double) * no);
    for (size_t i = 0; i < n; i++) { re[i] = creal(in[i]); im[i] = cimag(in[i]); }
    orc_interp2(re, rows, cols, ro, nr, nc, inX, inY, outX, outY);
    orc_interp2(im, rows, cols, io, nr, nc, inX, inY, outX, outY);
    for (size_t i = 0; i < no; i++) out[i] = ro[i] + I * io[i];
    free(inX); free(inY); free(outX); free(outY); free(re); free(im); free(ro); free(io);
    return 0;
}

/* ------------------------------------------------------------ pointwise -- */
/* mask::execute / init (src/elements/mask.cpp:32-34,38-46,53-58): the real mask
 * is copied into both the real and imaginary part: E *= (1+i) m. */
void orc_mask_apply(ocx* E, const double* m, int n) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) E[i] *= (m[i] + I * m[i]);
}
/* mirror::set_mirrorPhase (src/elements/mirror.cpp:66-72):
 * E *= exp((-sign * 4 pi / lambda) i h). */
void orc_mirror_apply(ocx* E, const double* h, int n, double sign, double lambda) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) E[i] *= cexp((-sign * 2 * 2 * PI / lambda) * I * h[i]);
}
/* efield::init_point_sources (src/data/efield.cpp:179-206): sqrt(flux) is stored
 * by the pointSource key (:295-304, tip/tilt times magnification there).
 * E = flux * exp(i 2 pi (2 XX tipX / Lx + 2 YY tiltY / Ly)). */
void orc_point_source(ocx* E, int rows, int cols, double pixelScale,
                      double flux, double tipX, double tiltY) {
    double* x = (double*)malloc(sizeof(double) * (size_t)cols);
    double* y = (double*)malloc(sizeof(double) * (size_t)rows);
    orc_grid_symmetric(cols, pixelScale, 0, x);
    orc_grid_symmetric(rows, pixelScale, 0, y);
    double Lx = pixelScale * cols, Ly = pixelScale * rows;
    for (int c = 0; c < cols; c++)
        for (int r = 0; r < rows; r++)
            E[(size_t)c * rows + r] = flux * cexp(I * 2.0 * PI * (2 * x[c] * tipX / Lx + 2 * y[r] * tiltY / Ly));
    free(x); free(y);
}

/* ----------------------------------------------------------------- hex -- */
/* complexHexMaskFPM::compute_hex_centers (src/elements/complexHexMaskFPM.cpp:236-315):
 * lattice (3i, sqrt3 j) step plus the (1.5, sqrt3/2) offset row, kept when
 * (int)|c| < maskNRings, numbered ring by ring in loop order. */
int orc_hex_centers_csim(int maskNRings, double hexStep, double* hx, double* hy,
                         int* ring, int maxHex) {
    int ii1Max = (int)(maskNRings / 3 + 2);
    int jj1Max = (int)(maskNRings / sqrt(3) + 2);
    int cap = 8 * ii1Max * jj1Max + 8;   // two centres per (ii1, jj1) over [-max, max)^2
    double* X = (double*)malloc(sizeof(double) * cap);
    double* Y = (double*)malloc(sizeof(double) * cap);
    int* R = (int*)malloc(sizeof(int) * cap);
    long long* num = (long long*)malloc(sizeof(long long) * cap);
    const long long bad = 100000000LL;
    int n = 0;
    for (int ii1 = -ii1Max; ii1 < ii1Max; ii1++) {
        for (int jj1 = -jj1Max; jj1 < jj1Max; jj1++) {
            double x = hexStep * ii1 * 3.0, y = hexStep * jj1 * sqrt(3.0);
            int rg = (int)sqrt(x * x + y * y);
            X[n] = x; Y[n] = y; R[n] = rg; num[n] = rg < maskNRings ? 0 : bad + n; n++;
            x += hexStep * 1.5; y += hexStep * sqrt(3.0) / 2.0;
            rg = (int)sqrt(x * x + y * y);
            X[n] = x; Y[n] = y; R[n] = rg; num[n] = rg < maskNRings ? 0 : bad + n; n++;
        }
    }
    /* number by ring (arma::find order = index order) */
    long long hCount = 0;
    int* assigned = (int*)calloc(cap, sizeof(int));
    for (int r = 0; r <= maskNRings; r++)
        for (int i = 0; i < n; i++)
            if (R[i] == r && num[i] == 0 && !assigned[i]) { num[i] = hCount++; assigned[i] = 1; }
    /* sort by number, keep the legal ones: number h goes to slot h */
    int count = 0;
    for (int i = 0; i < n; i++) {
        if (num[i] < bad) {
            long long h = num[i];
            if (h < maxHex) { hx[h] = X[i]; hy[h] = Y[i]; ring[h] = R[i]; }
            count++;
        }
    }
    free(X); free(Y); free(R); free(num); free(assigned);
    return count;
}

/* Standard hex rings (1 + 3 n (n+1) cells) on csim's lattice: basis
 * a1 = (1.5, sqrt3/2) step, a2 = (0, sqrt3) step; ring = max(|p|,|q|,|p+q|);
 * numbered by ring, then by polar angle in [0, 2 pi).  Not in the reference:
 * its radial cut-off never produces these counts. */
int orc_hex_centers_standard(int nRings, double hexStep, double* hx, double* hy,
                             int* ring, int maxHex) {
    int count = 0;
    for (int rg = 0; rg <= nRings; rg++) {
        int start = count;
        for (int p = -rg; p <= rg; p++)
            for (int q = -rg; q <= rg; q++) {
                int d = abs(p); if (abs(q) > d) d = abs(q); if (abs(p + q) > d) d = abs(p + q);
                if (d != rg) continue;
                if (count < maxHex) {
                    hx[count] = hexStep * 1.5 * p;
                    hy[count] = hexStep * (sqrt(3.0) / 2.0 * p + sqrt(3.0) * q);
                    ring[count] = rg;
                }
                count++;
            }
        /* order within the ring by angle */
        int m = count - start;
        if (count <= maxHex)
            for (int i = 1; i < m; i++)
                for (int j = i; j > 0; j--) {
                    int a = start + j - 1, b = start + j;
                    double ta = atan2(hy[a], hx[a]), tb = atan2(hy[b], hx[b]);
                    if (ta < 0) ta += 2 * PI;
                    if (tb < 0) tb += 2 * PI;
                    if (rg == 0 || ta <= tb) break;
                    double t; t = hx[a]; hx[a] = hx[b]; hx[b] = t;
                    t = hy[a]; hy[a] = hy[b]; hy[b] = t;
                }
    }
    return count;
}

/* complexHexMaskFPM::make_hex_array (src/elements/complexHexMaskFPM.cpp:317-366).
 * design is fpmArraySize^2 column-major, -1 outside every hex.  The reference
 * rasterises hexes in parallel; where hexes overlap (negative hexGap) the
 * highest-numbered hex wins here. `layoutRadius` is maskNRings for csim's layout. */
void orc_hex_design_array(const double* hx, const double* hy, int nHex, double layoutRadius,
                          int A, double fpmScaleFactor, double hexGap, int32_t* design) {
    double vx = sqrt(3) / 2, vy = 0.5;
    double hexStepPix = (0.5 * A) / layoutRadius * fpmScaleFactor;
    double hexRadius = hexStepPix * (1.0 - hexGap) * (sqrt(3.0) / 2.0);
    double hexOuterRadius = 2 * hexRadius / sqrt(3.0);
    double closeHexSq = hexOuterRadius * hexOuterRadius;
    for (size_t i = 0; i < (size_t)A * A; i++) design[i] = -1;
    for (int h = 0; h < nHex; h++) {
        double xc = 0.5 * A + hx[h] * hexStepPix;
        double yc = 0.5 * A + hy[h] * hexStepPix;
        int c0 = (int)floor(xc - hexOuterRadius) - 1, c1 = (int)ceil(xc + hexOuterRadius) + 1;
        int r0 = (int)floor(yc - hexOuterRadius) - 1, r1 = (int)ceil(yc + hexOuterRadius) + 1;
        if (c0 < 0) c0 = 0;
        if (r0 < 0) r0 = 0;
        if (c1 > A - 1) c1 = A - 1;
        if (r1 > A - 1) r1 = A - 1;
        for (int ii = r0; ii <= r1; ii++)
            for (int jj = c0; jj <= c1; jj++) {
                double dx = (double)jj - xc, dy = (double)ii - yc;
                if (dx * dx + dy * dy > closeHexSq) continue;
                double dpx = dx * vx, dpy = dy * vy;
                if ((fabs(dy) < hexRadius) & (fabs(dpx + dpy) < hexRadius) & (fabs(dpx - dpy) < hexRadius))
                    design[(size_t)jj * A + ii] = h;
            }
    }
}

/* complexHexMaskFPM::make_interpolation_data (src/elements/complexHexMaskFPM.cpp:372-469).
 * nsub / hexnum are S x S x 3 (column-major, slice-major).  Slot 0 counts the
 * off-design / outside samples and also the first hex met, as in the reference. */
void orc_hex_interp_data(const int32_t* design, int A, int S, int nSubPix, double pixelScale,
                         double fpmRadius, double fpmScaleFactor, int32_t* nsub, int32_t* hexnum) {
    size_t plane = (size_t)S * S;
    double ac1 = -S / 2.0 - 0.5;
    double xc1 = (0.5 / fpmRadius) * fpmScaleFactor;
    for (size_t i = 0; i < 3 * plane; i++) { nsub[i] = 0; hexnum[i] = -1; }
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < S; i++) {
        int iii = (int)(nSubPix / 2.0);
        double x = (i + ac1 + (0.5 + iii) / ((double)nSubPix)) * pixelScale;
        int ii1 = (int)((0.5 + x * xc1) * A + 0.5);
        for (int j = 0; j < S; j++) {
            int jjj = (int)(nSubPix / 2.0);
            double y = (j + ac1 + (0.5 + jjj) / ((double)nSubPix)) * pixelScale;
            int jj1 = (int)((0.5 + y * xc1) * A + 0.5);
            if ((ii1 < 0) | (ii1 >= A) | (jj1 < 0) | (jj1 >= A)) continue;
            int list[3], nl = 0;
            size_t p = (size_t)j * S + i;
            for (int a = 0; a < nSubPix; a++)
                for (int b = 0; b < nSubPix; b++) {
                    double xs = (i + ac1 + (0.5 + a) / ((double)nSubPix)) * pixelScale;
                    double ys = (j + ac1 + (0.5 + b) / ((double)nSubPix)) * pixelScale;
                    int i1 = (int)((0.5 + xs * xc1) * A + 0.5);
                    int j1 = (int)((0.5 + ys * xc1) * A + 0.5);
                    if ((i1 < 0) | (i1 >= A) | (j1 < 0) | (j1 >= A)) { nsub[p]++; continue; }
                    int h = design[(size_t)j1 * A + i1];
                    if (h < 0) { nsub[p]++; continue; }
                    int slot = -1;
                    for (int s = 0; s < nl; s++) if (list[s] == h) slot = s;
                    if (slot < 0) { if (nl >= 3) continue; list[nl] = h; slot = nl++; }
                    nsub[p + slot * plane]++;
                    hexnum[p + slot * plane] = h;
                }
        }
    }
    for (size_t p = 0; p < plane; p++)           /* "fix zeros" :462-468 */
        if (nsub[p] + nsub[p + plane] + nsub[p + 2 * plane] == 0) nsub[p] = nSubPix * nSubPix;
}

/* complexHexMaskFPM::make_complex_intpolated_mask (src/elements/complexHexMaskFPM.cpp:44-99):
 * mask = sum_s n_s exp(phaseSign 2 pi i 2 sag(h_s) / lambda) / nSubPix^2, sag 0 off-hex.
 * useOnlyThisHex >= 0 keeps only that hex's samples (== nHex keeps the off-hex ones). */
void orc_hex_mask(const int32_t* nsub, const int32_t* hexnum, int S, int nSubPix,
                  const double* sags, int nHex, double lambda, double phaseSign,
                  int useOnlyThisHex, ocx* mask) {
    size_t plane = (size_t)S * S;
    double norm = (double)nSubPix * nSubPix;
#pragma omp parallel for schedule(static)
    for (long long p = 0; p < (long long)plane; p++) {
        ocx acc = 0;
        for (int s = 0; s < 3; s++) {
            int h = hexnum[p + s * plane];
            double ns = nsub[p + s * plane];
            if (useOnlyThisHex > -1) {
                if (h == useOnlyThisHex) ;
                else if (useOnlyThisHex == nHex && h == -1) ;
                else ns = 0;
            }
            double sag = h > -1 ? sags[h] : 0.0;
            acc += ns * cexp(phaseSign * 2 * PI * I * 2.0 * sag / lambda);
        }
        mask[p] = acc / norm;
    }
}

/* doubleFraunhoferFPM centred mask multiply (src/elements/doubleFraunhoferFPM.cpp:84-97). */
void orc_center_mask_multiply(ocx* F, int rows, int cols, const ocx* m, int mrows, int mcols) {
    int row0 = rows / 2 - mrows / 2, col0 = cols / 2 - mcols / 2;
    for (int c = 0; c < mcols; c++)
        for (int r = 0; r < mrows; r++)
            F[(size_t)(col0 + c) * rows + row0 + r] *= m[(size_t)c * mrows + r];
}

/* fpmPupToLyot::execute, one slice (src/elements/fpmPupToLyot.cpp:24-167):
 * maskHat = zoomFftSign * zoomFft(M - 1 -> 2n pupil grid); F = FFT2(fftshift(maskHat));
 * E <- crop( IFFT2(F . FFT2(pad(E))) * dx^2 / (-i fRatioSign lambdaF) [+ pad(E)] ).
 * maskMinus1 is the S x S mask already minus 1 when babinet (and fliplr'd by the
 * caller for the interpolated hex variant, :444-445); lambdaF is signed. */
void orc_fpm_pup_to_lyot(ocx* E, int n, double dx, const ocx* maskMinus1, int S,
                         double maskPixelSize, double lambdaF, double zoomFftSign,
                         double fRatioSign, int babinet) {
    int n2 = 2 * n;
    size_t np = (size_t)n2 * n2;
    double mx0 = -(S * maskPixelSize - maskPixelSize) / 2.0;     /* maskGeom.set_xy */
    double px0 = -(n2 * dx - dx) / 2.0;                          /* paddedGeom.set_xy */
    ocx* hat = (ocx*)malloc(sizeof(ocx) * np);
    ocx* F = (ocx*)malloc(sizeof(ocx) * np);
    orc_zoomfft(maskMinus1, S, S, mx0, mx0, maskPixelSize, maskPixelSize,
                n2, px0, px0, dx, dx, lambdaF, hat);
    for (size_t i = 0; i < np; i++) hat[i] *= zoomFftSign;
    orc_fft_shift(hat, F, n2, n2);
    orc_fft2_csim(F, n2, n2);
    ocx* P = (ocx*)calloc(np, sizeof(ocx));
    for (int c = 0; c < n; c++)
        for (int r = 0; r < n; r++)
            P[(size_t)(c + n / 2) * n2 + r + n / 2] = E[(size_t)c * n + r];
    memcpy(hat, P, sizeof(ocx) * np);
    orc_fft2_csim(hat, n2, n2);
    for (size_t i = 0; i < np; i++) hat[i] *= F[i];
    orc_ifft2_csim(hat, n2, n2);
    ocx scale = pow(dx, 2) / (-I * fRatioSign * lambdaF);
    for (size_t i = 0; i < np; i++) {
        hat[i] *= scale;
        if (babinet) hat[i] += P[i];
    }
    for (int c = 0; c < n; c++)
        for (int r = 0; r < n; r++)
            E[(size_t)c * n + r] = hat[(size_t)(c + n / 2) * n2 + r + n / 2];
    free(hat); free(F); free(P);
}

/* ------------------------------------------------------------ contrast -- */
/* regionContrast::compute_intensity (src/tools/regionContrast.cpp:124-143): I += |E|^2. */
void orc_intensity_accumulate(const ocx* E, int n, double* Iacc) {
    for (int i = 0; i < n; i++) Iacc[i] += creal(E[i] * conj(E[i]));
}

/* regionContrast::get_region_pixels (src/tools/regionContrast.cpp:89-101) on a
 * symmetric n x n grid: radii inclusive, angles strict.  Returns the count and
 * writes column-major linear indices (arma::find order). */
int orc_region_pixels(int n, double d, double loD, double xlim1, double xlim2,
                      double r1, double r2, double a1deg, double a2deg, int32_t* idx) {
    double* x = (double*)malloc(sizeof(double) * (size_t)n);
    orc_grid_symmetric(n, d, 0, x);
    int cnt = 0;
    for (int c = 0; c < n; c++)
        for (int r = 0; r < n; r++) {
            double XX = x[c], YY = x[r];
            double RR = sqrt(pow(XX, 2) + pow(YY, 2)), TT = atan2(YY, XX);
            if ((XX / loD >= xlim1) & (XX / loD <= xlim2) & (RR / loD >= r1) & (RR / loD <= r2) &
                (TT > a1deg * PI / 180) & (TT < a2deg * PI / 180)) {
                if (idx) idx[cnt] = c * n + r;
                cnt++;
            }
        }
    free(x);
    return cnt;
}

/* optHexRespObjective::execute (src/optimization/optObjective.cpp:24-69):
 * per slice, E = M * [exp(i phaseSign 4 pi X / lambda); exp(0)], pixEnergy += |E|^2;
 * objective = mean (or max) over pixels / calib.  M is npix x ncol x nslices
 * column-major with ncol = nopt + 1; X is nopt x ncand column-major. */
void orc_hex_resp_objective(const ocx* M, int npix, int ncol, int nslices,
                            const double* lambdas, double phaseSign, double calib,
                            const double* X, int nopt, int ncand, int objMax, double* obj) {
#pragma omp parallel
    {
        ocx* R = (ocx*)malloc(sizeof(ocx) * (size_t)ncol);
        double* pe = (double*)malloc(sizeof(double) * (size_t)npix);
#pragma omp for schedule(dynamic, 1)
        for (int j = 0; j < ncand; j++) {
            for (int p = 0; p < npix; p++) pe[p] = 0;
            for (int sl = 0; sl < nslices; sl++) {
                double a = phaseSign * 4 * PI / lambdas[sl];
                for (int i = 0; i < ncol; i++) {
                    double xv = i < nopt ? X[(size_t)j * nopt + i] : 0.0;
                    R[i] = cexp(I * a * xv);
                }
                const ocx* Ms = M + (size_t)sl * npix * ncol;
                for (int p = 0; p < npix; p++) {
                    ocx e = 0;
                    for (int i = 0; i < ncol; i++) e += Ms[(size_t)i * npix + p] * R[i];
                    pe[p] += creal(e * conj(e));
                }
            }
            double v = 0;
            if (objMax) { v = pe[0]; for (int p = 1; p < npix; p++) if (pe[p] > v) v = pe[p]; }
            else { for (int p = 0; p < npix; p++) v += pe[p]; v /= npix; }
            obj[j] = v / calib;
        }
        free(R); free(pe);
    }
}

/* ----------------------------------------------------------------- MLP -- */
/* 3-layer perceptron in FP64: h1 = x W1 + b1, a1 = gelu(h1), h2 = a1 W2 + b2,
 * a2 = gelu(h2), y = a2 W3 + b3.  Row-major: x nb x nin, W1 nin x nh1 etc.
 * gelu(u) = u/2 (1 + erf(u / sqrt2)).  (No reference counterpart.) */
static double gelu(double u) { return 0.5 * u * (1.0 + erf(u / sqrt(2.0))); }
static double gelu_d(double u) {
    return 0.5 * (1.0 + erf(u / sqrt(2.0))) + u * exp(-0.5 * u * u) / sqrt(2.0 * PI);
}
static void dense(const double* x, int nb, int nin, const double* W, const double* b, int nout, double* y) {
    for (int i = 0; i < nb; i++)
        for (int o = 0; o < nout; o++) {
            double s = b[o];
            for (int k = 0; k < nin; k++) s += x[(size_t)i * nin + k] * W[(size_t)k * nout + o];
            y[(size_t)i * nout + o] = s;
        }
}
void orc_mlp_forward(const double* x, int nb, int nin, const double* W1, const double* b1,
                     int nh1, const double* W2, const double* b2, int nh2,
                     const double* W3, const double* b3, int nout,
                     double* h1, double* a1, double* h2, double* a2, double* y) {
    dense(x, nb, nin, W1, b1, nh1, h1);
    for (size_t i = 0; i < (size_t)nb * nh1; i++) a1[i] = gelu(h1[i]);
    dense(a1, nb, nh1, W2, b2, nh2, h2);
    for (size_t i = 0; i < (size_t)nb * nh2; i++) a2[i] = gelu(h2[i]);
    dense(a2, nb, nh2, W3, b3, nout, y);
}
/* loss = sum (y - t)^2 / (nb * nout); writes dloss/dparam. */
double orc_mlp_loss_grad(const double* x, const double* t, int nb, int nin,
                         const double* W1, const double* b1, int nh1,
                         const double* W2, const double* b2, int nh2,
                         const double* W3, const double* b3, int nout,
                         double* gW1, double* gb1, double* gW2, double* gb2,
                         double* gW3, double* gb3) {
    double* h1 = malloc(sizeof(double) * (size_t)nb * nh1);
    double* a1 = malloc(sizeof(double) * (size_t)nb * nh1);
    double* h2 = malloc(sizeof(double) * (size_t)nb * nh2);
    double* a2 = malloc(sizeof(double) * (size_t)nb * nh2);
    double* y = malloc(sizeof(double) * (size_t)nb * nout);
    orc_mlp_forward(x, nb, nin, W1, b1, nh1, W2, b2, nh2, W3, b3, nout, h1, a1, h2, a2, y);
    double loss = 0, sc = 1.0 / ((double)nb * nout);
    double* dy = malloc(sizeof(double) * (size_t)nb * nout);
    for (size_t i = 0; i < (size_t)nb * nout; i++) {
        double d = y[i] - t[i];
        loss += d * d * sc;
        dy[i] = 2.0 * d * sc;
    }
    double* d2 = calloc((size_t)nb * nh2, sizeof(double));
    double* d1 = calloc((size_t)nb * nh1, sizeof(double));
    memset(gW3, 0, sizeof(double) * (size_t)nh2 * nout); memset(gb3, 0, sizeof(double) * nout);
    memset(gW2, 0, sizeof(double) * (size_t)nh1 * nh2); memset(gb2, 0, sizeof(double) * nh2);
    memset(gW1, 0, sizeof(double) * (size_t)nin * nh1); memset(gb1, 0, sizeof(double) * nh1);
    for (int i = 0; i < nb; i++)
        for (int o = 0; o < nout; o++) {
            double g = dy[(size_t)i * nout + o];
            gb3[o] += g;
            for (int k = 0; k < nh2; k++) {
                gW3[(size_t)k * nout + o] += a2[(size_t)i * nh2 + k] * g;
                d2[(size_t)i * nh2 + k] += W3[(size_t)k * nout + o] * g;
            }
        }
    for (size_t i = 0; i < (size_t)nb * nh2; i++) d2[i] *= gelu_d(h2[i]);
    for (int i = 0; i < nb; i++)
        for (int o = 0; o < nh2; o++) {
            double g = d2[(size_t)i * nh2 + o];
            gb2[o] += g;
            for (int k = 0; k < nh1; k++) {
                gW2[(size_t)k * nh2 + o] += a1[(size_t)i * nh1 + k] * g;
                d1[(size_t)i * nh1 + k] += W2[(size_t)k * nh2 + o] * g;
            }
        }
    for (size_t i = 0; i < (size_t)nb * nh1; i++) d1[i] *= gelu_d(h1[i]);
    for (int i = 0; i < nb; i++)
        for (int o = 0; o < nh1; o++) {
            double g = d1[(size_t)i * nh1 + o];
            gb1[o] += g;
            for (int k = 0; k < nin; k++) gW1[(size_t)k * nh1 + o] += x[(size_t)i * nin + k] * g;
        }
    free(h1); free(a1); free(h2); free(a2); free(y); free(dy); free(d2); free(d1);
    return loss;
}
