/* Stand-in MEX runtime for tests/test_mex_shim.py (see mex.h). */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

#define MAXF 16
struct mxArray_tag {
    mxClassID cls;
    size_t m, n;
    void* data;
    int nf;
    char fname[MAXF][32];
    mxArray* fval[MAXF];
};

static jmp_buf g_jmp;
static char g_errid[128], g_errmsg[512];

mxArray* mxCreateNumericMatrix(size_t m, size_t n, mxClassID c, mxComplexity x)
{
    (void)x;
    mxArray* a = (mxArray*)calloc(1, sizeof(mxArray));
    size_t es = c == mxSINGLE_CLASS ? 4 : (c == mxCHAR_CLASS ? 1 : 8);
    a->cls = c;
    a->m = m;
    a->n = n;
    a->data = calloc(m * n + 1, es);
    return a;
}
mxArray* mock_string(const char* s)
{
    mxArray* a = mxCreateNumericMatrix(1, strlen(s), mxCHAR_CLASS, mxREAL);
    memcpy(a->data, s, strlen(s));
    return a;
}
mxArray* mock_struct(void)
{
    mxArray* a = (mxArray*)calloc(1, sizeof(mxArray));
    a->cls = mxSTRUCT_CLASS;
    a->m = a->n = 1;
    return a;
}
void mock_set_field(mxArray* s, const char* name, double v)
{
    mxArray* f = mxCreateNumericMatrix(1, 1, mxDOUBLE_CLASS, mxREAL);
    ((double*)f->data)[0] = v;
    snprintf(s->fname[s->nf], 32, "%s", name);
    s->fval[s->nf++] = f;
}
void mock_free(mxArray* a)
{
    if (!a) return;
    for (int i = 0; i < a->nf; i++) mock_free(a->fval[i]);
    free(a->data);
    free(a);
}
double* mxGetPr(const mxArray* a) { return (double*)a->data; }
void* mxGetData(const mxArray* a) { return a->data; }
size_t mxGetNumberOfElements(const mxArray* a) { return a->m * a->n; }
int mxIsDouble(const mxArray* a) { return a->cls == mxDOUBLE_CLASS; }
int mxIsSingle(const mxArray* a) { return a->cls == mxSINGLE_CLASS; }
int mxIsChar(const mxArray* a) { return a->cls == mxCHAR_CLASS; }
int mxIsStruct(const mxArray* a) { return a->cls == mxSTRUCT_CLASS; }
int mxGetString(const mxArray* a, char* buf, size_t len)
{
    size_t n = a->m * a->n;
    if (a->cls != mxCHAR_CLASS || n + 1 > len) return 1;
    memcpy(buf, a->data, n);
    buf[n] = 0;
    return 0;
}
mxArray* mxGetField(const mxArray* s, size_t idx, const char* name)
{
    (void)idx;
    for (int i = 0; i < s->nf; i++)
        if (!strcmp(s->fname[i], name)) return s->fval[i];
    return NULL;
}
double mxGetScalar(const mxArray* a)
{
    if (a->cls == mxSINGLE_CLASS) return ((float*)a->data)[0];
    return ((double*)a->data)[0];
}
int mexPrintf(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    int r = vprintf(fmt, ap);
    va_end(ap);
    fflush(stdout);
    return r;
}
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...)
{
    va_list ap;
    snprintf(g_errid, sizeof(g_errid), "%s", id);
    va_start(ap, fmt);
    vsnprintf(g_errmsg, sizeof(g_errmsg), fmt, ap);
    va_end(ap);
    longjmp(g_jmp, 1);
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);

/* 0 on success, 1 when the gateway raised a MATLAB error (see mock_error_id). */
int mock_call(int nlhs, mxArray** plhs, int nrhs, const mxArray** prhs)
{
    g_errid[0] = g_errmsg[0] = 0;
    if (setjmp(g_jmp)) return 1;
    mexFunction(nlhs, plhs, nrhs, prhs);
    return 0;
}
const char* mock_error_id(void) { return g_errid; }
const char* mock_error_msg(void) { return g_errmsg; }
