/* Test stand-in for the MATLAB MEX API subset the gateway knn-toolbox_amd/mex/mexknns.cpp uses.
 * It lets the gateway run without MATLAB or Octave: tests build the gateway against this header
 * and drive mexFunction through shim_* calls. Error raising (mexErrMsgIdAndTxt) throws, and
 * shim_call turns that into a status and a message. With MATLAB or Octave the gateway is built
 * with their own mex.h instead. */
#ifndef KNNS_TEST_MEX_H
#define KNNS_TEST_MEX_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef enum { mxUNKNOWN_CLASS = 0, mxDOUBLE_CLASS = 6, mxSINGLE_CLASS = 7, mxINT32_CLASS = 12 } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef struct mxArray_tag mxArray;

mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity cplx);
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity cplx);
void* mxGetData(const mxArray* a);
double* mxGetPr(const mxArray* a);
mwSize mxGetM(const mxArray* a);
mwSize mxGetN(const mxArray* a);
mwSize mxGetNumberOfElements(const mxArray* a);
mxClassID mxGetClassID(const mxArray* a);
int mxIsComplex(const mxArray* a);
double mxGetScalar(const mxArray* a);
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);

/* test driver side */
mxArray* shim_matrix(int cls, mwSize m, mwSize n, const void* data, int complex_flag);
void shim_free_all(void);
int shim_call(int nlhs, mxArray** plhs, int nrhs, mxArray** prhs);
const char* shim_error_id(void);
const char* shim_error_msg(void);

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);

#ifdef __cplusplus
}
#endif
#endif
