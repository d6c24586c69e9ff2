// MEX gateway over the drop-in C API: the counterpart of src/mexknnsGen.cu:55-268 and
// src/mexknnsHeap.cu:144-421, built four times (as the reference's Makefile builds its MEX files):
//
//   mexknnsBitonic   float   gpuknnsBitonic      mexknnsHeap    float   gpuknnsHeap
//   mexknnsDBitonic  double  gpuknnsBitonic      mexknnsDHeap   double  gpuknnsHeap
//
// (-D__DOUBLE__ selects double / libcuknn_double, -DKNNS_MEX_HEAP selects the heap entry point.)
//
// [dist idx time dTime sTime] = mexknns...(query, data, k, numStreams)
//   query: d x q (the columns are the query vectors), data: d x N, both single (or double)
//   dist, idx: k x q of the input class; idx holds 0-based corpus columns (the toolbox .m files
//   add 1); time, dTime, sTime: seconds of the whole call, the distance stage and the search
//   stage (mexknnsGen.cu:90-98).
// Column-major d x n is the row-major n x d layout of the C API, so the MATLAB arrays are passed
// to gpuknns* without copies. Unlike the reference gateways, the arguments are checked and a
// failed search raises a MATLAB error (id "knns:search") instead of returning garbage, and the
// device is not reset after the call.
#include <cmath>
#include <cstdint>

#include "mex.h"
#include "cuknns.h"

#ifdef __DOUBLE__
#define KNNS_MX_CLASS mxDOUBLE_CLASS
#define KNNS_MX_NAME "double"
#else
#define KNNS_MX_CLASS mxSINGLE_CLASS
#define KNNS_MX_NAME "single"
#endif

#ifdef KNNS_MEX_HEAP
#define KNNS_MEX_SEARCH gpuknnsHeap
#else
#define KNNS_MEX_SEARCH gpuknnsBitonic
#endif

namespace {

bool scalar_int(const mxArray* a, long lo, long hi, long* out) {
  if (mxGetNumberOfElements(a) != 1 || mxIsComplex(a)) return false;
  const double v = mxGetScalar(a);
  if (!(v >= (double)lo && v <= (double)hi) || v != std::floor(v)) return false;
  *out = (long)v;
  return true;
}

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs != 4)
    mexErrMsgIdAndTxt("knns:args", "expected (query, data, k, numStreams), got %d arguments", nrhs);
  if (nlhs > 5) mexErrMsgIdAndTxt("knns:args", "at most 5 outputs: [dist idx time dTime sTime]");
  const mxArray* Q = prhs[0];
  const mxArray* X = prhs[1];
  if (mxGetClassID(Q) != KNNS_MX_CLASS || mxGetClassID(X) != KNNS_MX_CLASS || mxIsComplex(Q) || mxIsComplex(X))
    mexErrMsgIdAndTxt("knns:class", "query and data must be real %s matrices", KNNS_MX_NAME);
  const long d = (long)mxGetM(Q);
  const long q = (long)mxGetN(Q);
  const long N = (long)mxGetN(X);
  if ((long)mxGetM(X) != d)
    mexErrMsgIdAndTxt("knns:dims", "query has %ld rows (dimensions) but data has %ld", d, (long)mxGetM(X));
  if (d < 1 || N < 1 || N > 0x7fffff80L || q > 0x7fffffffL)
    mexErrMsgIdAndTxt("knns:dims", "empty or oversized inputs (d=%ld, N=%ld, q=%ld)", d, N, q);
  long k = 0, streams = 0;
  if (!scalar_int(prhs[2], 1, 1024, &k)) mexErrMsgIdAndTxt("knns:k", "k must be an integer in [1, 1024]");
  if (!scalar_int(prhs[3], 1, 1024, &streams))
    mexErrMsgIdAndTxt("knns:streams", "numStreams must be a positive integer");

  plhs[0] = mxCreateNumericMatrix((size_t)k, (size_t)q, KNNS_MX_CLASS, mxREAL);
  mxArray* idx = mxCreateNumericMatrix((size_t)k, (size_t)q, KNNS_MX_CLASS, mxREAL);
  double t = 0.0;
  if (q > 0) {
    // dTime / sTime (outputs 4 and 5) are per-stage device times: recorded only when asked for
    knnsSetStageTimes(nlhs > 3 ? 1 : 0);
    t = KNNS_MEX_SEARCH((knntype*)mxGetData(Q), (knntype*)mxGetData(X), (knntype*)mxGetData(plhs[0]),
                        (knntype*)mxGetData(idx), (int)N, (int)q, (int)d, (int)k, (int)streams);
    const char* err = knnsGetLastError();
    if (err && err[0]) mexErrMsgIdAndTxt("knns:search", "search failed: %s", err);
  }
  const knntimes tm = knnsGetLastTimes();
  const double outs[3] = {t, tm.dst_time * 1e-3, tm.srch_time * 1e-3};
  if (nlhs > 1) plhs[1] = idx;
  for (int o = 2; o < nlhs; ++o) {
    plhs[o] = mxCreateDoubleMatrix(1, 1, mxREAL);
    mxGetPr(plhs[o])[0] = outs[o - 2];
  }
}
