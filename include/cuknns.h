/* Drop-in C API of the brute-force kNN library, compatible with the reference's include/cuknns.h.
 *
 * Build libcuknn.so (knntype = float) and libcuknn_double.so (knntype = double, -D__DOUBLE__):
 * both export the same symbol names, like the reference's libcuknn.a / libcuknn_double.a
 * (Makefile:57,64). Unlike the reference header (cuknns.h:5-6) this one includes no CUDA or
 * cuBLAS header: declarations that carry cublasHandle_t / CUstream (cuknns.h:32, 62-73) are per-
 * chunk internals of the CUDA build and are replaced by the native API in knns_native.h.
 *
 * Layouts (readme.txt:61-69): `data` is N x D and `queries` Q x D with D leading (row-major);
 * KNNdist / KNNidx are Q x k with k leading. Indices are 0-based values stored in knntype
 * (src/cuknnBitonic.cu:68), exact below 2^24 rows for float. Results are sorted by ascending
 * distance with ties broken by the lower index, a refinement of the reference's "not always
 * ordered" output (readme.txt:66-69).
 */
#ifndef CUKNN
#define CUKNN

#ifndef __DOUBLE__
typedef float knntype;   /* cuknns.h:8-12 */
#else
typedef double knntype;
#endif

#define MEMTEST_ONLY 10        /* cuknns.h:14 */
#define EXACT_KNN 5            /* cuknns.h:15 */
#define FORCE_SINGLE_STREAM -5 /* cuknns.h:16 */

/* cuknns.h:18 */
typedef double (*knnptr2function)(knntype*, knntype*, knntype*, knntype*, int, int, int, int, int);

/* cuknns.h:34-41 (field names kept, including "dimentions") */
typedef struct {
  knnptr2function pt2Function;
  int objects;
  int numQueries;
  int dimentions;
  int k;
  int numStreams;
} knnplan;

/* cuknns.h:43-47: stage timers of the last call, in milliseconds like the reference's
 * cudaEventElapsedTime accumulators (src/cuknnBitonic.cu:352-368). */
typedef struct {
  double dst_time;
  double srch_time;
  double knn_time;
} knntimes;

#ifdef __cplusplus
extern "C" {
#endif

/* cuknns.h:75. Returns elapsed seconds, FLT_MAX on failure (src/cuknnBitonic.cu:581). */
double gpuknnsBitonic(knntype* query, knntype* data, knntype* values, knntype* indices,
                      int objects, int numQueries, int attributes, int k, int numStreams);

/* cuknns.h:83 */
double gpuknnsHeap(knntype* query, knntype* data, knntype* values, knntype* indices,
                   int objects, int numQueries, int attributes, int k, int numStreams);

/* cuknns.h:79 (src/cuknnBitonic.cu:621-847, unfinished there: its search call is commented
 * out). Bucketed search: the corpus is c consecutive buckets of bucketSize[j] rows; the queries
 * query_offsets[j] .. +query_sizes[j] get their exact top-k within bucket j, at their own rows of
 * values / indices (k leading), indices = corpus rows; slots past a bucket smaller than k hold
 * +inf and -1. dp and numStreams are ignored. Seconds, FLT_MAX on failure. */
float gpuknnLSH(knntype* query, knntype* data, knntype* values, knntype* indices, knntype* dp,
                int objects, int numQueries, int attributes, int k, int numStreams,
                int* bucketSize, int* query_offsets, int c, int* query_sizes);

/* cuknns.h:85. A fixed rule, no trial runs on random data (src/knnsplan.cu:43-57 ran 80):
 * gpuknnsHeap with one stream (the two entry points return identical results and the heap form
 * measured faster at every size tried). */
void knnsplan(knnplan* plan, long int N, long int Q, long int D, long int k);

/* cuknns.h:87. Prints "Time Elapsed: %f\n" like src/knnsexecute.cu:16 (KNNS_QUIET=1 silences). */
void knnsexecute(knnplan plan, knntype* data, knntype* queries, knntype* KNNdist,
                 knntype* KNNidx);

/* Additions with no counterpart in the reference: the reason of the last failure (empty string
 * after a success) and the timers of the last call. knn_time (the whole call) is always set;
 * dst_time and srch_time (distance and search stages) only while stage timing is on, because it
 * records a device event around every kernel launch of the call. Stage timing is off unless
 * knnsSetStageTimes(1) was called or the environment holds KNNS_STAGE_TIMES=1. */
const char* knnsGetLastError(void);
knntimes knnsGetLastTimes(void);
void knnsSetStageTimes(int enable);

#ifdef __cplusplus
}
#endif

#endif /* CUKNN */
