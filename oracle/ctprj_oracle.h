/*
 * FP64 CPU oracle for the CT strip-integral projector -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library, and only as the checker.  The product path (libctprj.so) never calls it.
 *
 * It restates the reference projector's system matrix in double precision and
 * evaluates every (angle, bin) ray directly, without the reference's 8-fold angular
 * symmetry, so any angle count works:
 *   - geometry / conventions: prj/cpuPrj.c:319-616 (rayDrivePar, rayDriveFan),
 *     prj/cpuPrj.c:46-317 (pixelDrivePar, pixelDriveFan), prj/prj.h:8-57;
 *   - pixel-strip footprint weight: prj/utils.h:12-92 (getWeight);
 *   - output quirks: /effectiveRate (prj/cpuPrj.c:405,579), sinogram bin 0 never
 *     computed for an even detector (prj/cpuPrj.c:703-705), image row 0 / column 0
 *     never read or written for an even image size (prj/cpuPrj.c:713-716).
 *   - FBP: fan cosine pre-weight prj/cpuPrj.c:733-751, Hann-windowed ramp filter
 *     prj/utils.h:94-151, backprojection scale prj/cpuPrj.c:161-163,258,304-306.
 *
 * Parity pins (tests/test_oracle_*.py): the reference's own cpuPrj compiled from its
 * sources into oracle/_ref (golden fixtures), exact polygon-clipping areas of
 * the pixel/strip overlap, analytic ellipse line integrals, brute-force dense A on
 * tiny images, and <Ax,y> = <x,A^T y>.
 */
#ifndef CTPRJ_ORACLE_H
#define CTPRJ_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_conf {
    int n;              /* image is n x n, row-major, pixel (x,y) at [(y+n/2)*n + x+n/2] */
    int prjWidth;       /* detector bins per view */
    int np;             /* views computed: angle index a in [0,np) */
    int prjFull;        /* views per 360 degrees: theta_a = a*2*pi/prjFull */
    double dSize;       /* detector bin pitch (pixels) */
    double effectiveRate; /* beam width = dSize*effectiveRate; outputs are divided by it */
    double d;           /* source-to-rotation-centre distance; <= 0 means parallel beam */
    int angle_mode;     /* 0: theta in double; 1: the reference's float theta
                           (thetaIdx*2*PI/prjFull with PI=3.14159265359f, prj/cpuPrj.c:342);
                           2, 3, 4: symmetric angles -- view a takes the float theta of
                           its canonical view a0 under the 8-fold (2), 4-fold (3) or
                           2-fold (4) symmetry of the pixel grid, rotated / reflected
                           exactly (prj/cpuPrj.c:383-460 reuses the canonical view's
                           weights for all eight images in the same way) */
    int support_mode;   /* 0: reference support -- a ray touches a pixel iff the ray's
                              detector strip meets the pixel (prj/cpuPrj.c:365-369,530-536,
                              82-105,219-244);
                           1: every pixel whose footprint weight is non-zero */
} orc_conf;

/* prj/utils.h:12-92 in double precision: overlap area of a unit pixel and a strip of
 * width bw whose centre line is `dist` away from the pixel centre, ray direction
 * (cosR, sinR). */
double orc_get_weight(double dist, double bw, double cosR, double sinR);

/* Exact area of {u in [-1/2,1/2]^2 : |u.(c,s) + dist| <= bw/2} by polygon clipping
 * (independent check of orc_get_weight; c^2+s^2 = 1). */
double orc_strip_area(double dist, double bw, double c, double s);

/* sino[np][prjWidth] = A img ; returns nnz(A) visited (weights != 0). */
long long orc_forward(const orc_conf* c, const double* img, double* sino, int nthreads);
/* img[n][n] = A^T sino ; returns nnz(A). */
long long orc_adjoint(const orc_conf* c, const double* sino, double* img, int nthreads);
/* nnz(A) only. */
long long orc_nnz(const orc_conf* c, int nthreads);
/* Dense A (np*prjWidth x n*n), row-major, for tiny problems. */
int orc_dense(const orc_conf* c, double* A);

/* FBP (the 'FBP' command): fan cosine pre-weight, ramp filter per view, then the
 * FBP-weighted backprojection.  img[n][n] from sino[np][prjWidth]. */
int orc_fbp(const orc_conf* c, const double* sino, double* img, int nthreads);
/* the Hann-windowed ramp filter of one view of length `size` (prj/utils.h:94-151),
 * by direct O(size^2) circular convolution in double */
void orc_ramp_filter(double* signal, int size);

/* A stored in compressed rows (row = sinogram index a * prjWidth + b, column = image
 * index), built from the same visit as orc_forward: orc_csr_counts gives the entries per
 * row, the caller forms rowptr (exclusive prefix sum, np * prjWidth + 1 entries) and
 * orc_csr_fill writes each row in visiting order.  orc_csr_apply multiplies K stacked
 * vectors: trans = 0 gives out[k] = A in[k] (in: K images, out: K sinograms; each row
 * adds its terms in orc_forward's order), trans = 1 gives out[k] = A^T in[k]. */
int orc_csr_counts(const orc_conf* c, long long* rowcnt, int nthreads);
int orc_csr_fill(const orc_conf* c, const long long* rowptr, int* col, double* val, int nthreads);
int orc_csr_apply(long long rows, long long ncols, const long long* rowptr, const int* col,
                  const double* val, const double* in, double* out, int K, int trans,
                  int nthreads);

#ifdef __cplusplus
}
#endif
#endif
