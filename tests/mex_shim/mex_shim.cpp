// Implementation of the test stand-in MEX API (tests/mex_shim/mex.h).
#include "mex.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

struct mxArray_tag {
  mxClassID cls;
  mwSize m, n;
  int cplx;
  std::vector<char> data;
};

namespace {
std::vector<std::unique_ptr<mxArray_tag>> g_arrays;
std::string g_id, g_msg;
struct MexError {};

size_t elem_size(mxClassID c) { return c == mxDOUBLE_CLASS ? 8 : 4; }
}  // namespace

extern "C" {

mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity cplx) {
  auto a = std::make_unique<mxArray_tag>();
  a->cls = cls;
  a->m = m;
  a->n = n;
  a->cplx = cplx == mxCOMPLEX;
  a->data.assign(m * n * elem_size(cls), 0);
  g_arrays.push_back(std::move(a));
  return g_arrays.back().get();
}
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity cplx) {
  return mxCreateNumericMatrix(m, n, mxDOUBLE_CLASS, cplx);
}
void* mxGetData(const mxArray* a) { return (void*)a->data.data(); }
double* mxGetPr(const mxArray* a) { return (double*)a->data.data(); }
mwSize mxGetM(const mxArray* a) { return a->m; }
mwSize mxGetN(const mxArray* a) { return a->n; }
mwSize mxGetNumberOfElements(const mxArray* a) { return a->m * a->n; }
mxClassID mxGetClassID(const mxArray* a) { return a->cls; }
int mxIsComplex(const mxArray* a) { return a->cplx; }
double mxGetScalar(const mxArray* a) {
  if (a->m * a->n == 0) return 0.0;
  switch (a->cls) {
    case mxDOUBLE_CLASS: return ((const double*)a->data.data())[0];
    case mxSINGLE_CLASS: return ((const float*)a->data.data())[0];
    case mxINT32_CLASS: return ((const int*)a->data.data())[0];
    default: return 0.0;
  }
}
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_id = id;
  g_msg = buf;
  throw MexError{};
}

mxArray* shim_matrix(int cls, mwSize m, mwSize n, const void* data, int complex_flag) {
  mxArray* a = mxCreateNumericMatrix(m, n, (mxClassID)cls, complex_flag ? mxCOMPLEX : mxREAL);
  if (data) std::memcpy(a->data.data(), data, a->data.size());
  return a;
}
void shim_free_all(void) { g_arrays.clear(); }
int shim_call(int nlhs, mxArray** plhs, int nrhs, mxArray** prhs) {
  g_id.clear();
  g_msg.clear();
  try {
    mexFunction(nlhs, plhs, nrhs, (const mxArray**)prhs);
  } catch (const MexError&) {
    return 1;
  }
  return 0;
}
const char* shim_error_id(void) { return g_id.c_str(); }
const char* shim_error_msg(void) { return g_msg.c_str(); }

}  // extern "C"
