/* Minimal stand-in for MATLAB's MEX API, enough to build and drive the gpuPrj gateway
 * in tests: numeric (double/single/char) matrices and 1x1 structs.
 * mexErrMsgIdAndTxt longjmps back to mock_call() with the identifier recorded. */
#ifndef MEXMOCK_MEX_H
#define MEXMOCK_MEX_H
#include <stddef.h>

typedef enum { mxDOUBLE_CLASS = 6, mxSINGLE_CLASS = 7, mxCHAR_CLASS = 4, mxSTRUCT_CLASS = 2 } mxClassID;
typedef enum { mxREAL = 0 } mxComplexity;
typedef struct mxArray_tag mxArray;

mxArray* mxCreateNumericMatrix(size_t m, size_t n, mxClassID c, mxComplexity x);
double* mxGetPr(const mxArray* a);
void* mxGetData(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
int mxIsDouble(const mxArray* a);
int mxIsSingle(const mxArray* a);
int mxIsChar(const mxArray* a);
int mxIsStruct(const mxArray* a);
int mxGetString(const mxArray* a, char* buf, size_t len);
mxArray* mxGetField(const mxArray* s, size_t idx, const char* name);
double mxGetScalar(const mxArray* a);
int mexPrintf(const char* fmt, ...);
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
#endif
