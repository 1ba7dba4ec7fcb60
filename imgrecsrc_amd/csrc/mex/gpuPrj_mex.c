/*
 * gpuPrj MEX gateway over libctprj: the MATLAB operator API of prj/mPrj.c:37-173 with
 * the same call forms and the same marshalling:
 *
 *   gpuPrj(0, conf, 'config')        conf.{n,prjWidth,np,prjFull,dSize,effectiveRate,d}
 *   s = gpuPrj(img, 0, 'forward')    img n x n (double or single)  -> prjWidth*np x 1 double
 *   x = gpuPrj(sino, 0, 'backward')  sino, prjWidth*np elements     -> n*n x 1 double
 *   x = gpuPrj(sino, 0, 'FBP')       as 'backward'
 *   gpuPrj(0, 0, <other string>)     prints the configuration
 *
 * MATLAB column-major n x n <-> row-major image with the same orientation: MATLAB
 * element (i, j) is image row i, column j (prj/mPrj.c:82-84, 134-136).  Sinograms are
 * view-major in both (prj/mPrj.c:110-111, 118-123).
 *
 * Unlike the reference gateway, every failure (argument count, type, size, device)
 * raises a MATLAB error through mexErrMsgIdAndTxt instead of printing and returning
 * (prj/mPrj.c:43-46,101-104) or exit()ing (prj/gpuPrj.cu:60-68).
 *
 * Build in MATLAB:  mex -I<repo>/include gpuPrj_mex.c -L<repo>/imgrecsrc_amd/lib -lctprj
 *                       -output gpuPrj
 */
#include <stdlib.h>
#include <string.h>

#include "ctprj.h"
#include "mex.h"

static ctprj_geom g_geom;
static int g_configured = 0;
static float* g_img = NULL;
static float* g_sino = NULL;

static void release(void)
{
    free(g_img);
    free(g_sino);
    g_img = g_sino = NULL;
}

static void lib_error(int rc, const char* what)
{
    mexErrMsgIdAndTxt("gpuPrj:device", "%s failed (%d): %s", what, rc, ctprj_last_error());
}

static double field(const mxArray* s, const char* name)
{
    const mxArray* f;
    if (!mxIsStruct(s)) mexErrMsgIdAndTxt("gpuPrj:config", "'config' needs a struct argument");
    f = mxGetField(s, 0, name);
    if (!f || mxGetNumberOfElements(f) < 1)
        mexErrMsgIdAndTxt("gpuPrj:config", "config struct has no field '%s'", name);
    return mxGetScalar(f);
}

static void need_config(void)
{
    if (!g_configured) mexErrMsgIdAndTxt("gpuPrj:state", "call gpuPrj(0, conf, 'config') first");
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[])
{
    char cmd[32];
    int rc;
    (void)nlhs;
    if (nrhs != 3) mexErrMsgIdAndTxt("gpuPrj:nrhs", "number of input arguments are: %d (need 3)", nrhs);
    if (!mxIsChar(prhs[2]) || mxGetString(prhs[2], cmd, sizeof(cmd)) != 0)
        mexErrMsgIdAndTxt("gpuPrj:cmd", "third argument must be a command string");

    if (strcmp(cmd, "config") == 0) {
        ctprj_geom g;
        g.n = (int)field(prhs[1], "n");
        g.prjWidth = (int)field(prhs[1], "prjWidth");
        g.np = (int)field(prhs[1], "np");
        g.prjFull = (int)field(prhs[1], "prjFull");
        g.dSize = (float)field(prhs[1], "dSize");
        g.effectiveRate = (float)field(prhs[1], "effectiveRate");
        g.d = (float)field(prhs[1], "d");
        rc = prj_setup(g.n, g.prjWidth, g.np, g.prjFull, g.dSize, g.effectiveRate, g.d);
        if (rc) lib_error(rc, "prj_setup");
        release();
        g_img = (float*)malloc(sizeof(float) * (size_t)g.n * g.n);
        g_sino = (float*)malloc(sizeof(float) * (size_t)g.np * g.prjWidth);
        if (!g_img || !g_sino) {
            release();
            g_configured = 0;
            mexErrMsgIdAndTxt("gpuPrj:nomem", "host staging allocation failed");
        }
        g_geom = g;
        g_configured = 1;
        return;
    }

    if (strcmp(cmd, "forward") == 0) {
        const int n = g_geom.n;
        const size_t ni = (size_t)n * n, ns = (size_t)g_geom.np * g_geom.prjWidth;
        size_t i, j;
        double* out;
        need_config();
        if (mxGetNumberOfElements(prhs[0]) != ni)
            mexErrMsgIdAndTxt("gpuPrj:size", "forward: image has %d elements, config needs %d",
                              (int)mxGetNumberOfElements(prhs[0]), (int)ni);
        if (mxIsDouble(prhs[0])) {
            const double* in = mxGetPr(prhs[0]);
            for (j = 0; j < (size_t)n; j++)
                for (i = 0; i < (size_t)n; i++) g_img[i * n + j] = (float)in[i + j * n];
        } else if (mxIsSingle(prhs[0])) {
            const float* in = (const float*)mxGetData(prhs[0]);
            for (j = 0; j < (size_t)n; j++)
                for (i = 0; i < (size_t)n; i++) g_img[i * n + j] = in[i + j * n];
        } else {
            mexErrMsgIdAndTxt("gpuPrj:type", "unknown input parameter type in gpuPrj forward");
        }
        rc = prj_run(g_img, g_sino, CTPRJ_FWD_BIT);
        if (rc) lib_error(rc, "forward");
        plhs[0] = mxCreateNumericMatrix(ns, 1, mxDOUBLE_CLASS, mxREAL);
        out = mxGetPr(plhs[0]);
        for (i = 0; i < ns; i++) out[i] = g_sino[i];
        return;
    }

    if (strcmp(cmd, "backward") == 0 || strcmp(cmd, "FBP") == 0) {
        const int n = g_geom.n;
        const size_t ni = (size_t)n * n, ns = (size_t)g_geom.np * g_geom.prjWidth;
        size_t i, j;
        double* out;
        need_config();
        if (mxGetNumberOfElements(prhs[0]) != ns)
            mexErrMsgIdAndTxt("gpuPrj:size", "%s: sinogram has %d elements, config needs %d", cmd,
                              (int)mxGetNumberOfElements(prhs[0]), (int)ns);
        if (mxIsDouble(prhs[0])) {
            const double* in = mxGetPr(prhs[0]);
            for (i = 0; i < ns; i++) g_sino[i] = (float)in[i];
        } else if (mxIsSingle(prhs[0])) {
            memcpy(g_sino, mxGetData(prhs[0]), sizeof(float) * ns);
        } else {
            mexErrMsgIdAndTxt("gpuPrj:type", "unknown input parameter type in gpuPrj %s", cmd);
        }
        rc = prj_run(g_img, g_sino, cmd[0] == 'b' ? CTPRJ_BWD_BIT : CTPRJ_FBP_BIT);
        if (rc) lib_error(rc, cmd);
        plhs[0] = mxCreateNumericMatrix(ni, 1, mxDOUBLE_CLASS, mxREAL);
        out = mxGetPr(plhs[0]);
        for (j = 0; j < (size_t)n; j++)
            for (i = 0; i < (size_t)n; i++) out[i * n + j] = g_img[i + j * n];
        return;
    }

    need_config();
    mexPrintf("\nPrinting the current configuration ...\n");
    mexPrintf("config.n=%d\nconfig.prjWidth=%d\nconfig.np=%d\nconfig.prjFull=%d\n", g_geom.n,
              g_geom.prjWidth, g_geom.np, g_geom.prjFull);
    mexPrintf("config.dSize=%g\nconfig.effectiveRate=%g\nconfig.d=%g\n", (double)g_geom.dSize,
              (double)g_geom.effectiveRate, (double)g_geom.d);
}
