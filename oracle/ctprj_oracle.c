/*
 * FP64 CPU oracle for the CT projector -- TEST INFRASTRUCTURE ONLY (see ctprj_oracle.h).
 *
 * Every ray (a, b) is evaluated directly in double precision.  For each ray the
 * candidate pixels are taken along the ray's major axis with a 2-pixel margin and each
 * candidate is then tested with the support rule of `support_mode`, so the enumeration
 * never depends on rounding of the bounds.
 */
#include "ctprj_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define REF_PI_F 3.14159265359f /* prj/prj.h:9 */
#define ORC_PI 3.14159265358979323846

/* ---------------------------------------------------------------- weight --- */

/* prj/utils.h:12-92 restated in double precision (same case split). */
double orc_get_weight(double dist, double beamWidth, double cosR, double sinR)
{
    double c = fabs(cosR), s = fabs(sinR), tmp;
    if (c < s) { tmp = c; c = s; s = tmp; }
    const double height = 1.0 / c;
    const double bias2 = (c + s) / 2.0;
    const double bias1 = fabs((c - s) / 2.0);
    const double areaT = 0.5 * height * (bias2 - bias1);
    const double areaR = height * bias1 * 2.0;
    const double area = height * (bias1 + bias2);
    double w = 0.0;
    dist = fabs(dist);
    const double lfoot = dist - bias2, rfoot = dist + bias2;
    const double lshoulder = dist - bias1, rshoulder = dist + bias1;
    const double hb = beamWidth / 2.0;
    const double slope = bias2 - bias1;
    if (-hb <= lfoot) {
        if (hb <= lfoot) w = 0.0;
        else if (hb <= lshoulder) { tmp = hb - lfoot; w = 0.5 * tmp * tmp * height / slope; }
        else if (hb <= rshoulder) w = areaT + height * (hb - lshoulder);
        else if (hb <= rfoot) { tmp = rfoot - hb; w = areaT + areaR + areaT - 0.5 * tmp * tmp * height / slope; }
        else w = area;
    } else if (-hb <= lshoulder) {
        if (hb <= lshoulder) w = height * (-lfoot) * beamWidth / slope;
        else if (hb <= rshoulder) {
            tmp = -hb - lfoot;
            w = areaT + height * (hb - lshoulder) - 0.5 * tmp * height * tmp / slope;
        } else if (hb <= rfoot) {
            tmp = -hb - lfoot;
            w = areaT + areaR - 0.5 * tmp * height * tmp / slope;
            tmp = rfoot - hb;
            w = w + areaT - 0.5 * tmp * height * tmp / slope;
        } else {
            tmp = -hb - lfoot;
            w = areaT - 0.5 * tmp * height * tmp / slope + areaR + areaT;
        }
    } else if (-hb <= rshoulder) {
        if (hb <= rshoulder) w = height * beamWidth;
        else if (hb <= rfoot) {
            tmp = rfoot - hb;
            w = height * (rshoulder + hb) + areaT - 0.5 * tmp * height * tmp / slope;
        } else w = height * (rshoulder + hb) + areaT;
    } else {
        if (-hb <= rfoot) {
            if (hb <= rfoot) w = height * rfoot * beamWidth / slope;
            else { tmp = rfoot + hb; w = 0.5 * height * tmp * tmp / slope; }
        }
    }
    return w;
}

/* Sutherland-Hodgman clip of a convex polygon by {p : nx*px + ny*py <= r}. */
static int clip_half(const double* in, int nin, double* out, double nx, double ny, double r)
{
    int nout = 0;
    for (int i = 0; i < nin; i++) {
        const double* p = in + 2 * i;
        const double* q = in + 2 * ((i + 1) % nin);
        double fp = nx * p[0] + ny * p[1] - r, fq = nx * q[0] + ny * q[1] - r;
        if (fp <= 0) { out[2 * nout] = p[0]; out[2 * nout + 1] = p[1]; nout++; }
        if ((fp < 0 && fq > 0) || (fp > 0 && fq < 0)) {
            double t = fp / (fp - fq);
            out[2 * nout] = p[0] + t * (q[0] - p[0]);
            out[2 * nout + 1] = p[1] + t * (q[1] - p[1]);
            nout++;
        }
    }
    return nout;
}

double orc_strip_area(double dist, double bw, double c, double s)
{
    double a[16] = {-0.5, -0.5, 0.5, -0.5, 0.5, 0.5, -0.5, 0.5}, b[16];
    int n = clip_half(a, 4, b, c, s, bw / 2 - dist);     /* u.n + dist <= bw/2 */
    n = clip_half(b, n, a, -c, -s, bw / 2 + dist);       /* -(u.n + dist) <= bw/2 */
    double area = 0;
    for (int i = 0; i < n; i++) {
        int j = (i + 1) % n;
        area += a[2 * i] * a[2 * j + 1] - a[2 * j] * a[2 * i + 1];
    }
    return fabs(area) / 2;
}

/* -------------------------------------------------------------- geometry --- */

typedef struct {
    double cosT, sinT;
} view_t;

static double ref_theta(const orc_conf* c, int a)
{
    /* prj/cpuPrj.c:342 -- float theta, then cos()/sin() of that float in double */
    float thf = (float)(a * 2) * REF_PI_F / (float)c->prjFull;
    return (double)thf;
}

/* Canonical view a0 and group element R^k M^m with view a = g a0 (views mod P):
 *   R^k   maps a0 -> a0 + k P/4,   M maps a0 -> P/2 - a0,
 * rotations preferred where both reach a (mirror views).  Group by angle_mode:
 * 2 = all eight (P % 4 == 0), 3 = {1, R^2, M, R^2 M} (P even), 4 = {1, R^2 M}. */
static void canonical_of(const orc_conf* c, int a, int* a0, int* k, int* m)
{
    const int P = c->prjFull;
    if (c->angle_mode == 2) {
        const int Q = P / 4, j = a / Q, r = a % Q;
        if (8 * r <= P) { *a0 = r; *k = j; *m = 0; }
        else { *a0 = Q - r; *k = (j + 3) & 3; *m = 1; }
    } else if (c->angle_mode == 3) {
        const int H = P / 2, j = a / H, r = a % H;
        if (4 * r <= P) { *a0 = r; *k = 2 * j; *m = 0; }
        else { *a0 = H - r; *k = 2 * j; *m = 1; }
    } else {
        if (2 * a <= P) { *a0 = a; *k = 0; *m = 0; }
        else { *a0 = P - a; *k = 2; *m = 1; }
    }
}

static view_t view_of(const orc_conf* c, int a)
{
    view_t v;
    if (c->angle_mode >= 2) {
        /* the ray normal (cos, sin) of the canonical view, mapped by g: M takes theta to
         * pi - theta, R to theta + pi/2 (exact in double) */
        int a0, k, m;
        canonical_of(c, a, &a0, &k, &m);
        const double th = ref_theta(c, a0);
        double cs = cos(th), sn = sin(th);
        if (m) cs = -cs;
        for (int i = 0; i < (k & 3); i++) {
            const double t = cs;
            cs = -sn;
            sn = t;
        }
        v.cosT = cs;
        v.sinT = sn;
        return v;
    }
    const double th = (c->angle_mode == 1) ? ref_theta(c, a)
                                           : (double)a * 2.0 * ORC_PI / (double)c->prjFull;
    v.cosT = cos(th);
    v.sinT = sin(th);
    return v;
}

typedef struct {
    double t;            /* detector position of the bin centre */
    double cosR, sinR;   /* ray direction normal (parallel: the view's) */
    double dc;           /* signed distance of the ray from the rotation centre */
    double bwc;          /* beam width at the isocentre, perpendicular to the ray */
} ray_t;

static ray_t ray_of(const orc_conf* c, view_t v, int b)
{
    ray_t r;
    const double pC = c->prjWidth / 2;
    const double bw = c->dSize * c->effectiveRate;
    r.t = (b - pC) * c->dSize;
    if (c->d > 0) {
        /* prj/cpuPrj.c:503-507 */
        double d = c->d, bq = sqrt(d * d + r.t * r.t);
        double cosB = d / bq, sinB = r.t / bq;
        r.cosR = cosB * v.cosT - sinB * v.sinT;
        r.sinR = sinB * v.cosT + cosB * v.sinT;
        r.dc = d * r.t / bq;
        r.bwc = bw * cosB;
    } else {
        r.cosR = v.cosT;
        r.sinR = v.sinT;
        r.dc = r.t;
        r.bwc = bw;
    }
    return r;
}

/* The weight of pixel (x, y) for ray r, 0 when outside the support rule. */
static double ray_pixel_weight(const orc_conf* c, view_t v, ray_t r, int x, int y)
{
    const double hb = c->dSize * c->effectiveRate / 2;
    double dist = x * r.cosR + y * r.sinR - r.dc;
    double bw = r.bwc;
    if (c->d > 0) bw = r.bwc + (x * v.sinT - y * v.cosT) * r.bwc / c->d; /* prj/cpuPrj.c:549 */
    if (c->support_mode == 0) {
        if (c->d > 0) {
            /* pixel shadow on the virtual detector, prj/cpuPrj.c:219-239 */
            double tl = 1e300, tr = -1e300;
            for (int k = 0; k < 4; k++) {
                double X = x + ((k == 1 || k == 2) ? 0.5 : -0.5);
                double Y = y + ((k >= 2) ? 0.5 : -0.5);
                double qe = X * v.sinT - Y * v.cosT + c->d;
                double oc = X * v.cosT + Y * v.sinT;
                double t = c->d * oc / qe;
                if (t < tl) tl = t;
                if (t > tr) tr = t;
            }
            if (!(tl < r.t + hb && tr > r.t - hb)) return 0.0;
        } else {
            double b2 = (fabs(r.cosR) + fabs(r.sinR)) / 2;
            if (!(fabs(dist) < b2 + hb)) return 0.0;
        }
    }
    return orc_get_weight(dist, bw, r.cosR, r.sinR) / c->effectiveRate;
}

/* Visit every (pixel, weight) of ray (a, b).  op: 0 fwd, 1 adj, 2 count, 3 FBP adj,
 * 4 record the row of A (pixel indices to col, weights to val, in visiting order). */
static long long visit_ray_rec(const orc_conf* c, int a, int b, const double* in, double* out,
                               int op, int* col, double* val)
{
    const int N = c->n, H = (N - 1) / 2;
    view_t v = view_of(c, a);
    ray_t r = ray_of(c, v, b);
    const double hbmax = (c->d > 0)
        ? r.bwc * (1.0 + 0.75 * N / c->d) / 2 + 1.0
        : r.bwc / 2;
    const int rowmajor = fabs(r.cosR) >= fabs(r.sinR);
    const double cu = rowmajor ? r.cosR : r.sinR; /* coefficient of the inner coordinate */
    const double cv = rowmajor ? r.sinR : r.cosR;
    const double half = (fabs(r.cosR) + fabs(r.sinR)) / 2 + hbmax;
    long long nnz = 0;
    double acc = 0;
    const int sidx = a * c->prjWidth + b;
    for (int v2 = -H; v2 <= H; v2++) {
        double uc = (r.dc - v2 * cv) / cu;
        int lo = (int)floor(uc - half / fabs(cu)) - 2, hi = (int)ceil(uc + half / fabs(cu)) + 2;
        if (lo < -H) lo = -H;
        if (hi > H) hi = H;
        for (int u = lo; u <= hi; u++) {
            int x = rowmajor ? u : v2, y = rowmajor ? v2 : u;
            double w = ray_pixel_weight(c, v, r, x, y);
            if (w == 0.0) continue;
            nnz++;
            long pidx = (long)(y + N / 2) * N + x + N / 2;
            if (op == 4) {
                col[nnz - 1] = (int)pidx;
                val[nnz - 1] = w;
            }
            if (op == 0) acc += w * in[pidx];
            else if (op == 1) out[pidx] += w * in[sidx];
            else if (op == 3) {
                double f = 1.0;
                if (c->d > 0) { /* FBP fan factor d^2/qe^2, prj/cpuPrj.c:258 */
                    double qe = c->d + x * v.sinT - y * v.cosT;
                    f = c->d * c->d / (qe * qe);
                }
                out[pidx] += w * f * in[sidx];
            }
        }
    }
    if (op == 0) out[sidx] = acc;
    return nnz;
}

static long long visit_ray(const orc_conf* c, int a, int b, const double* in, double* out, int op)
{
    return visit_ray_rec(c, a, b, in, out, op, NULL, NULL);
}

/* ------------------------------------------------------------ threading --- */

typedef struct {
    const orc_conf* c;
    const double* in;
    double* out;
    int op, tid, nth;
    long long nnz;
} job_t;

static int bin_lo(const orc_conf* c) { return (c->prjWidth % 2 == 0) ? 1 : 0; }

static void* worker(void* p)
{
    job_t* j = (job_t*)p;
    const orc_conf* c = j->c;
    for (int a = j->tid; a < c->np; a += j->nth)
        for (int b = bin_lo(c); b < c->prjWidth; b++)
            j->nnz += visit_ray(c, a, b, j->in, j->out, j->op);
    return NULL;
}

static long long run(const orc_conf* c, const double* in, double* out, int op, int nth, long outsz)
{
    if (nth < 1) nth = 1;
    if (nth > 64) nth = 64;
    pthread_t th[64];
    job_t jobs[64];
    double* priv[64];
    for (int i = 0; i < nth; i++) {
        priv[i] = NULL;
        if (op == 1 || op == 3) priv[i] = (double*)calloc(outsz, sizeof(double));
        jobs[i] = (job_t){c, in, (op == 1 || op == 3) ? priv[i] : out, op, i, nth, 0};
        pthread_create(&th[i], NULL, worker, &jobs[i]);
    }
    long long nnz = 0;
    for (int i = 0; i < nth; i++) {
        pthread_join(th[i], NULL);
        nnz += jobs[i].nnz;
    }
    if (op == 1 || op == 3) {
        memset(out, 0, outsz * sizeof(double));
        for (int i = 0; i < nth; i++) {
            for (long k = 0; k < outsz; k++) out[k] += priv[i][k];
            free(priv[i]);
        }
    }
    return nnz;
}

long long orc_forward(const orc_conf* c, const double* img, double* sino, int nthreads)
{
    memset(sino, 0, sizeof(double) * (size_t)c->np * c->prjWidth);
    return run(c, img, sino, 0, nthreads, (long)c->np * c->prjWidth);
}

long long orc_adjoint(const orc_conf* c, const double* sino, double* img, int nthreads)
{
    return run(c, sino, img, 1, nthreads, (long)c->n * c->n);
}

long long orc_nnz(const orc_conf* c, int nthreads)
{
    return run(c, NULL, NULL, 2, nthreads, 0);
}

int orc_dense(const orc_conf* c, double* A)
{
    const int N = c->n;
    const long cols = (long)N * N;
    memset(A, 0, sizeof(double) * (size_t)c->np * c->prjWidth * cols);
    double* e = (double*)calloc(cols, sizeof(double));
    double* s = (double*)calloc((size_t)c->np * c->prjWidth, sizeof(double));
    for (long p = 0; p < cols; p++) {
        e[p] = 1.0;
        orc_forward(c, e, s, 1);
        for (long r = 0; r < (long)c->np * c->prjWidth; r++) A[r * cols + p] = s[r];
        e[p] = 0.0;
    }
    free(e);
    free(s);
    return 0;
}

/* ------------------------------------------------------------------ FBP --- */

/* prj/utils.h:94-151.  With L = 2*size, the filter spectrum is the DFT of the periodic
 * Ram-Lak kernel h[0]=1/4, h[odd k]=-1/(pi^2 k^2), h[even k]=0 (indices mod L), times a
 * Hann window 0.5+0.5cos(2 pi k/L).  The window's inverse DFT is 0.5 delta[0] +
 * 0.25 (delta[1] + delta[-1]), so the filtered view is the circular convolution of the
 * zero-padded view with g = 0.5 h + 0.25 h(.-1) + 0.25 h(.+1); only the first `size`
 * outputs are kept. */
static double ramlak(long k, long L)
{
    k = ((k % L) + L) % L;
    if (k > L / 2) k = L - k;
    if (k == 0) return 0.25;
    if (k % 2 == 0) return 0.0;
    return -1.0 / (ORC_PI * ORC_PI * (double)k * (double)k);
}

void orc_ramp_filter(double* signal, int size)
{
    const long L = 2L * size;
    double* g = (double*)malloc(sizeof(double) * L);
    double* out = (double*)malloc(sizeof(double) * size);
    for (long k = 0; k < L; k++)
        g[k] = 0.5 * ramlak(k, L) + 0.25 * ramlak(k - 1, L) + 0.25 * ramlak(k + 1, L);
    for (long i = 0; i < size; i++) {
        double acc = 0;
        for (long j = 0; j < size; j++) acc += signal[j] * g[((i - j) % L + L) % L];
        out[i] = acc;
    }
    memcpy(signal, out, sizeof(double) * size);
    free(g);
    free(out);
}

int orc_fbp(const orc_conf* c, const double* sino, double* img, int nthreads)
{
    const int W = c->prjWidth, pC = W / 2;
    const long ssz = (long)c->np * W, isz = (long)c->n * c->n;
    double* ws = (double*)malloc(sizeof(double) * ssz);
    for (int a = 0; a < c->np; a++)
        for (int b = 0; b < W; b++) {
            double v = sino[(long)a * W + b];
            if (c->d > 0) { /* prj/cpuPrj.c:737-751 */
                double t = (b - pC) * c->dSize;
                v *= c->d / sqrt(c->d * c->d + t * t);
            }
            ws[(long)a * W + b] = v;
        }
    for (int a = 0; a < c->np; a++) orc_ramp_filter(ws + (long)a * W, W);
    run(c, ws, img, 3, nthreads, isz);
    const double scale = ORC_PI / c->np / c->dSize / c->dSize; /* prj/cpuPrj.c:161,304 */
    for (long k = 0; k < isz; k++) img[k] *= scale;
    free(ws);
    return 0;
}

/* ------------------------------------------------------------ stored A --- */

/* A in compressed rows: row r = a * prjWidth + b (the sinogram index), columns the image
 * index, each row in visit_ray's order, so a stored forward adds the same terms in the
 * same order as orc_forward.  Rows of bin 0 of an even detector are empty. */
typedef struct {
    const orc_conf* c;
    long long* rowcnt;
    const long long* rowptr;
    int* col;
    double* val;
    int tid, nth;
} csr_job_t;

static void* csr_worker(void* p)
{
    csr_job_t* j = (csr_job_t*)p;
    const orc_conf* c = j->c;
    const int W = c->prjWidth;
    for (int a = j->tid; a < c->np; a += j->nth)
        for (int b = bin_lo(c); b < W; b++) {
            const long r = (long)a * W + b;
            if (j->rowptr)
                visit_ray_rec(c, a, b, NULL, NULL, 4, j->col + j->rowptr[r], j->val + j->rowptr[r]);
            else
                j->rowcnt[r] = visit_ray(c, a, b, NULL, NULL, 2);
        }
    return NULL;
}

static void csr_run(const orc_conf* c, long long* rowcnt, const long long* rowptr, int* col,
                    double* val, int nth)
{
    if (nth < 1) nth = 1;
    if (nth > 64) nth = 64;
    pthread_t th[64];
    csr_job_t jobs[64];
    for (int i = 0; i < nth; i++) {
        jobs[i] = (csr_job_t){c, rowcnt, rowptr, col, val, i, nth};
        pthread_create(&th[i], NULL, csr_worker, &jobs[i]);
    }
    for (int i = 0; i < nth; i++) pthread_join(th[i], NULL);
}

int orc_csr_counts(const orc_conf* c, long long* rowcnt, int nthreads)
{
    memset(rowcnt, 0, sizeof(long long) * (size_t)c->np * c->prjWidth);
    csr_run(c, rowcnt, NULL, NULL, NULL, nthreads);
    return 0;
}

int orc_csr_fill(const orc_conf* c, const long long* rowptr, int* col, double* val, int nthreads)
{
    csr_run(c, NULL, rowptr, col, val, nthreads);
    return 0;
}

typedef struct {
    long long rows, ncols;
    const long long* rowptr;
    const int* col;
    const double* val;
    const double* in;
    double* out;
    int K, tid, nth, trans;
} spmv_job_t;

static void* spmv_worker(void* p)
{
    spmv_job_t* j = (spmv_job_t*)p;
    /* contiguous row blocks: forward writes disjoint outputs; the transpose adds into
     * the thread's own image buffers */
    const long long r0 = j->rows * j->tid / j->nth, r1 = j->rows * (j->tid + 1) / j->nth;
    for (long long r = r0; r < r1; r++) {
        const long long e0 = j->rowptr[r], e1 = j->rowptr[r + 1];
        for (int k = 0; k < j->K; k++) {
            if (!j->trans) {
                const double* x = j->in + (size_t)k * j->ncols;
                double acc = 0;
                for (long long e = e0; e < e1; e++) acc += j->val[e] * x[j->col[e]];
                j->out[(size_t)k * j->rows + r] = acc;
            } else {
                const double y = j->in[(size_t)k * j->rows + r];
                double* x = j->out + (size_t)k * j->ncols;
                if (y != 0.0)
                    for (long long e = e0; e < e1; e++) x[j->col[e]] += j->val[e] * y;
            }
        }
    }
    return NULL;
}

int orc_csr_apply(long long rows, long long ncols, const long long* rowptr, const int* col,
                  const double* val, const double* in, double* out, int K, int trans,
                  int nthreads)
{
    int nth = nthreads < 1 ? 1 : (nthreads > 64 ? 64 : nthreads);
    pthread_t th[64];
    spmv_job_t jobs[64];
    double* priv[64];
    const size_t osz = (size_t)K * (trans ? ncols : rows);
    for (int i = 0; i < nth; i++) {
        priv[i] = trans ? (double*)calloc(osz, sizeof(double)) : NULL;
        if (trans && !priv[i]) {
            for (int k = 0; k < i; k++) free(priv[k]);
            return -1;
        }
        jobs[i] = (spmv_job_t){rows, ncols, rowptr, col, val, in, trans ? priv[i] : out,
                               K, i, nth, trans};
        pthread_create(&th[i], NULL, spmv_worker, &jobs[i]);
    }
    for (int i = 0; i < nth; i++) pthread_join(th[i], NULL);
    if (trans) {
        memset(out, 0, osz * sizeof(double));
        for (int i = 0; i < nth; i++) {
            for (size_t k = 0; k < osz; k++) out[k] += priv[i][k];
            free(priv[i]);
        }
    }
    return 0;
}
