/* Drives the gpuPrj gateway (imgrecsrc_amd/csrc/mex/gpuPrj_mex.c) through the stand-in
 * MEX runtime (tests/mexmock) and the stand-in library (ctprj_stub.c), built with
 * -fsanitize=address,undefined: every command form, the layout conversions, and the
 * error paths including a failed host staging allocation ('gpuPrj:nomem').  Exit
 * status 0 when every check holds. */
#include <stdio.h>
#include <string.h>

#include "mex.h"

mxArray* mock_string(const char* s);
mxArray* mock_struct(void);
void mock_set_field(mxArray* s, const char* name, double v);
void mock_free(mxArray* a);
int mock_call(int nlhs, mxArray** plhs, int nrhs, const mxArray** prhs);
const char* mock_error_id(void);
void mock_fail_malloc(int k);

static int failures = 0;
#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            failures++;                                                  \
        }                                                                \
    } while (0)

static mxArray* scalar(double v)
{
    mxArray* a = mxCreateNumericMatrix(1, 1, mxDOUBLE_CLASS, mxREAL);
    mxGetPr(a)[0] = v;
    return a;
}

static mxArray* conf(int n, int W, int np)
{
    mxArray* s = mock_struct();
    mock_set_field(s, "n", n);
    mock_set_field(s, "prjWidth", W);
    mock_set_field(s, "np", np);
    mock_set_field(s, "prjFull", 2 * np);
    mock_set_field(s, "dSize", 1);
    mock_set_field(s, "effectiveRate", 1);
    mock_set_field(s, "d", 0);
    return s;
}

/* calls the gateway; returns the error id ("" on success) and the output in *out */
static const char* call3(mxArray* a, mxArray* b, const char* cmd, mxArray** out)
{
    mxArray* c = mock_string(cmd);
    const mxArray* prhs[3] = {a, b, c};
    mxArray* plhs[1] = {NULL};
    const int rc = mock_call(1, plhs, 3, prhs);
    mock_free(c);
    if (out) *out = plhs[0];
    else mock_free(plhs[0]);
    return rc ? mock_error_id() : "";
}

int main(void)
{
    const int n = 5, W = 9, np = 4;  /* 25 pixels, 36 bins */
    mxArray *zero = scalar(0), *cf = conf(n, W, np), *out = NULL;
    int i, j;

    /* state: nothing configured yet */
    mxArray* img = mxCreateNumericMatrix(n, n, mxDOUBLE_CLASS, mxREAL);
    for (j = 0; j < n; j++)
        for (i = 0; i < n; i++) mxGetPr(img)[i + j * n] = 10 * i + j;
    CHECK(!strcmp(call3(img, zero, "forward", NULL), "gpuPrj:state"));

    /* the first and the second staging allocation fail in turn: a MATLAB error, and the
     * gateway stays unconfigured */
    for (i = 1; i <= 2; i++) {
        mock_fail_malloc(i);
        CHECK(!strcmp(call3(zero, cf, "config", NULL), "gpuPrj:nomem"));
        mock_fail_malloc(0);
        CHECK(!strcmp(call3(img, zero, "forward", NULL), "gpuPrj:state"));
    }

    /* a good configuration, then forward: MATLAB element (i, j) is image row i, col j */
    CHECK(!strcmp(call3(zero, cf, "config", NULL), ""));
    CHECK(!strcmp(call3(img, zero, "forward", &out), ""));
    CHECK(out && mxGetNumberOfElements(out) == (size_t)W * np);
    if (out) {
        for (i = 0; i < n; i++)
            for (j = 0; j < n; j++) CHECK(mxGetPr(out)[i * n + j] == 10 * i + j + 1);
        for (i = n * n; i < W * np; i++) CHECK(mxGetPr(out)[i] == 0);
    }

    /* backward: the sinogram's first n*n entries come back as image row i, col j */
    mxArray* back = NULL;
    CHECK(!strcmp(call3(out, zero, "backward", &back), ""));
    CHECK(back && mxGetNumberOfElements(back) == (size_t)n * n);
    if (back)
        for (j = 0; j < n; j++)
            for (i = 0; i < n; i++) CHECK(mxGetPr(back)[i + j * n] == 10 * i + j);
    mock_free(back);
    mxArray* fbp = NULL;
    CHECK(!strcmp(call3(out, zero, "FBP", &fbp), ""));
    mock_free(fbp);

    /* size and type errors */
    mxArray* small = mxCreateNumericMatrix(n, n + 1, mxDOUBLE_CLASS, mxREAL);
    CHECK(!strcmp(call3(small, zero, "forward", NULL), "gpuPrj:size"));
    CHECK(!strcmp(call3(small, zero, "backward", NULL), "gpuPrj:size"));
    mxArray* chars = mock_string("abcdefghijklmnopqrstuvwxy");
    CHECK(!strcmp(call3(chars, zero, "forward", NULL), "gpuPrj:type"));
    CHECK(!strcmp(call3(zero, zero, "show", NULL), ""));

    /* a failed reconfiguration leaves the gateway unconfigured, a good one restores it */
    mock_fail_malloc(1);
    CHECK(!strcmp(call3(zero, cf, "config", NULL), "gpuPrj:nomem"));
    mock_fail_malloc(0);
    CHECK(!strcmp(call3(img, zero, "forward", NULL), "gpuPrj:state"));
    CHECK(!strcmp(call3(zero, cf, "config", NULL), ""));

    mock_free(small);
    mock_free(chars);
    mock_free(out);
    mock_free(img);
    mock_free(cf);
    mock_free(zero);
    if (failures) fprintf(stderr, "%d checks failed\n", failures);
    else printf("shim driver: all checks passed\n");
    return failures ? 1 : 0;
}
