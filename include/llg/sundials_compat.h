/*
 * sundials_compat.h -- the SUNDIALS v7 vocabulary the LLG plug-ins are
 * written against: scalar types, the generic N_Vector object with its ops
 * table, the CVODE/CVLS callback signatures, and thin dispatchers
 * (N_VLinearSum, N_VDotProd, ...).
 *
 * SUNDIALS itself is a third-party dependency of the reference that is not
 * vendored there (src/irregular/2d_fft/2d_i4/Makefile:2-13, version >= 7.0
 * inferred from SUN_COMM_NULL at 2d_i4/2d_fft.cu:478 and the SUNErrCode
 * op signature at 2d_i4/deferred_nvector.cu:146).  These declarations are
 * written from its public API so that the plug-in signatures below are the
 * ones the reference registers with CVodeInit / CVodeSetJacTimes /
 * CVodeSetPreconditioner (2d_i4/2d_fft.cu:625,642,643).  The ops-table
 * field order follows the public SUNDIALS 7 `struct _generic_N_Vector_Ops`
 * (every slot, including the XBraid buffer ops), checked below.
 */
#ifndef LLG_SUNDIALS_COMPAT_H
#define LLG_SUNDIALS_COMPAT_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double  sunrealtype;
typedef int64_t sunindextype;
typedef int     sunbooleantype;
typedef int     SUNErrCode;
typedef int     SUNComm;

#define SUNTRUE        1
#define SUNFALSE       0
#define SUN_SUCCESS    0
#define SUN_COMM_NULL  0
#define SUN_RCONST(x)  (x)

/* preconditioning sides (SUNLinSol pretype) */
#define SUN_PREC_NONE  0
#define SUN_PREC_LEFT  1
#define SUN_PREC_RIGHT 2
#define SUN_PREC_BOTH  3

/* Gram-Schmidt flavours (SUNLinSol_SPGMRSetGSType) */
#define SUN_MODIFIED_GS  1
#define SUN_CLASSICAL_GS 2

typedef struct _SUNContext* SUNContext;

typedef enum {
  SUNDIALS_NVEC_SERIAL,
  SUNDIALS_NVEC_PARALLEL,
  SUNDIALS_NVEC_OPENMP,
  SUNDIALS_NVEC_PTHREADS,
  SUNDIALS_NVEC_PARHYP,
  SUNDIALS_NVEC_PETSC,
  SUNDIALS_NVEC_CUDA,
  SUNDIALS_NVEC_HIP,
  SUNDIALS_NVEC_SYCL,
  SUNDIALS_NVEC_RAJA,
  SUNDIALS_NVEC_KOKKOS,
  SUNDIALS_NVEC_OPENMPDEV,
  SUNDIALS_NVEC_TRILINOS,
  SUNDIALS_NVEC_MANYVECTOR,
  SUNDIALS_NVEC_MPIMANYVECTOR,
  SUNDIALS_NVEC_MPIPLUSX,
  SUNDIALS_NVEC_CUSTOM
} N_Vector_ID;

typedef struct _generic_N_Vector_Ops* N_Vector_Ops;
typedef struct _generic_N_Vector*     N_Vector;
typedef N_Vector*                     N_Vector_S;

struct _generic_N_Vector_Ops {
  /* constructors, destructors, utility */
  N_Vector_ID (*nvgetvectorid)(N_Vector);
  N_Vector (*nvclone)(N_Vector);
  N_Vector (*nvcloneempty)(N_Vector);
  void (*nvdestroy)(N_Vector);
  void (*nvspace)(N_Vector, sunindextype*, sunindextype*);
  sunrealtype* (*nvgetarraypointer)(N_Vector);
  sunrealtype* (*nvgetdevicearraypointer)(N_Vector);
  void (*nvsetarraypointer)(sunrealtype*, N_Vector);
  SUNComm (*nvgetcommunicator)(N_Vector);
  sunindextype (*nvgetlength)(N_Vector);
  sunindextype (*nvgetlocallength)(N_Vector);
  /* standard vector operations */
  void (*nvlinearsum)(sunrealtype, N_Vector, sunrealtype, N_Vector, N_Vector);
  void (*nvconst)(sunrealtype, N_Vector);
  void (*nvprod)(N_Vector, N_Vector, N_Vector);
  void (*nvdiv)(N_Vector, N_Vector, N_Vector);
  void (*nvscale)(sunrealtype, N_Vector, N_Vector);
  void (*nvabs)(N_Vector, N_Vector);
  void (*nvinv)(N_Vector, N_Vector);
  void (*nvaddconst)(N_Vector, sunrealtype, N_Vector);
  sunrealtype (*nvdotprod)(N_Vector, N_Vector);
  sunrealtype (*nvmaxnorm)(N_Vector);
  sunrealtype (*nvwrmsnorm)(N_Vector, N_Vector);
  sunrealtype (*nvwrmsnormmask)(N_Vector, N_Vector, N_Vector);
  sunrealtype (*nvmin)(N_Vector);
  sunrealtype (*nvwl2norm)(N_Vector, N_Vector);
  sunrealtype (*nvl1norm)(N_Vector);
  void (*nvcompare)(sunrealtype, N_Vector, N_Vector);
  sunbooleantype (*nvinvtest)(N_Vector, N_Vector);
  sunbooleantype (*nvconstrmask)(N_Vector, N_Vector, N_Vector);
  sunrealtype (*nvminquotient)(N_Vector, N_Vector);
  /* fused vector operations */
  SUNErrCode (*nvlinearcombination)(int, sunrealtype*, N_Vector*, N_Vector);
  SUNErrCode (*nvscaleaddmulti)(int, sunrealtype*, N_Vector, N_Vector*, N_Vector*);
  SUNErrCode (*nvdotprodmulti)(int, N_Vector, N_Vector*, sunrealtype*);
  /* vector array operations */
  SUNErrCode (*nvlinearsumvectorarray)(int, sunrealtype, N_Vector*, sunrealtype,
                                       N_Vector*, N_Vector*);
  SUNErrCode (*nvscalevectorarray)(int, sunrealtype*, N_Vector*, N_Vector*);
  SUNErrCode (*nvconstvectorarray)(int, sunrealtype, N_Vector*);
  SUNErrCode (*nvwrmsnormvectorarray)(int, N_Vector*, N_Vector*, sunrealtype*);
  SUNErrCode (*nvwrmsnormmaskvectorarray)(int, N_Vector*, N_Vector*, N_Vector,
                                          sunrealtype*);
  SUNErrCode (*nvscaleaddmultivectorarray)(int, int, sunrealtype*, N_Vector*,
                                           N_Vector**, N_Vector**);
  SUNErrCode (*nvlinearcombinationvectorarray)(int, int, sunrealtype*,
                                               N_Vector**, N_Vector*);
  /* local reductions (single-rank: identical to the global ones) */
  sunrealtype (*nvdotprodlocal)(N_Vector, N_Vector);
  sunrealtype (*nvmaxnormlocal)(N_Vector);
  sunrealtype (*nvminlocal)(N_Vector);
  sunrealtype (*nvl1normlocal)(N_Vector);
  sunbooleantype (*nvinvtestlocal)(N_Vector, N_Vector);
  sunbooleantype (*nvconstrmasklocal)(N_Vector, N_Vector, N_Vector);
  sunrealtype (*nvminquotientlocal)(N_Vector, N_Vector);
  sunrealtype (*nvwsqrsumlocal)(N_Vector, N_Vector);
  sunrealtype (*nvwsqrsummasklocal)(N_Vector, N_Vector, N_Vector);
  /* single-buffer reductions */
  SUNErrCode (*nvdotprodmultilocal)(int, N_Vector, N_Vector*, sunrealtype*);
  SUNErrCode (*nvdotprodmultiallreduce)(int, N_Vector, sunrealtype*);
  /* XBraid buffer interface (left NULL by the HIP vector) */
  SUNErrCode (*nvbufsize)(N_Vector, sunindextype*);
  SUNErrCode (*nvbufpack)(N_Vector, void*);
  SUNErrCode (*nvbufunpack)(N_Vector, void*);
  /* debugging */
  void (*nvprint)(N_Vector);
  void (*nvprintfile)(N_Vector, FILE*);
};

struct _generic_N_Vector {
  void*        content;
  N_Vector_Ops ops;
  SUNContext   sunctx;
};

/* SUNDIALS 7 layout of the ops table: 56 function pointers, the XBraid
 * buffer ops at slots 51-53 and the two print ops last.  A real SUNDIALS
 * build copies and indexes the table with these offsets (N_VCopyOps,
 * N_VPrint, N_VBufSize), so they are checked at compile time. */
#define LLG_NV_OPS_SLOTS 56
#ifdef __cplusplus
#define LLG_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define LLG_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif
LLG_STATIC_ASSERT(sizeof(struct _generic_N_Vector_Ops) == LLG_NV_OPS_SLOTS * sizeof(void*),
                  "N_Vector ops table must hold the 56 SUNDIALS 7 slots");
LLG_STATIC_ASSERT(__builtin_offsetof(struct _generic_N_Vector_Ops, nvlinearcombination) ==
                      30 * sizeof(void*), "fused ops start at slot 30");
LLG_STATIC_ASSERT(__builtin_offsetof(struct _generic_N_Vector_Ops, nvdotprodlocal) ==
                      40 * sizeof(void*), "local reductions start at slot 40");
LLG_STATIC_ASSERT(__builtin_offsetof(struct _generic_N_Vector_Ops, nvbufsize) ==
                      51 * sizeof(void*), "XBraid buffer ops start at slot 51");
LLG_STATIC_ASSERT(__builtin_offsetof(struct _generic_N_Vector_Ops, nvprint) ==
                      54 * sizeof(void*), "nvprint is slot 54");
LLG_STATIC_ASSERT(sizeof(struct _generic_N_Vector) == 3 * sizeof(void*),
                  "N_Vector object is {content, ops, sunctx}");

/* CVODE / CVLS user-supplied callback types (public SUNDIALS signatures) */
typedef int (*CVRhsFn)(sunrealtype t, N_Vector y, N_Vector ydot, void* user_data);
typedef int (*CVLsJacTimesSetupFn)(sunrealtype t, N_Vector y, N_Vector fy,
                                   void* user_data);
typedef int (*CVLsJacTimesVecFn)(N_Vector v, N_Vector Jv, sunrealtype t,
                                 N_Vector y, N_Vector fy, void* user_data,
                                 N_Vector tmp);
typedef int (*CVLsPrecSetupFn)(sunrealtype t, N_Vector y, N_Vector fy,
                               sunbooleantype jok, sunbooleantype* jcurPtr,
                               sunrealtype gamma, void* user_data);
typedef int (*CVLsPrecSolveFn)(sunrealtype t, N_Vector y, N_Vector fy,
                               N_Vector r, N_Vector z, sunrealtype gamma,
                               sunrealtype delta, int lr, void* user_data);

/* ---- context ---------------------------------------------------------- */
struct _SUNContext { int in_use; };

/* ---- generic dispatchers ---------------------------------------------- */
static inline N_Vector N_VClone(N_Vector w) { return w->ops->nvclone(w); }
static inline void N_VDestroy(N_Vector v) { if (v && v->ops && v->ops->nvdestroy) v->ops->nvdestroy(v); }
static inline sunindextype N_VGetLength(N_Vector v) { return v->ops->nvgetlength(v); }
static inline sunrealtype* N_VGetArrayPointer(N_Vector v) { return v->ops->nvgetarraypointer(v); }
static inline sunrealtype* N_VGetDeviceArrayPointer(N_Vector v) {
  return v->ops->nvgetdevicearraypointer ? v->ops->nvgetdevicearraypointer(v)
                                         : v->ops->nvgetarraypointer(v);
}
static inline void N_VLinearSum(sunrealtype a, N_Vector x, sunrealtype b, N_Vector y, N_Vector z) {
  z->ops->nvlinearsum(a, x, b, y, z);
}
static inline void N_VConst(sunrealtype c, N_Vector z) { z->ops->nvconst(c, z); }
static inline void N_VProd(N_Vector x, N_Vector y, N_Vector z) { z->ops->nvprod(x, y, z); }
static inline void N_VDiv(N_Vector x, N_Vector y, N_Vector z) { z->ops->nvdiv(x, y, z); }
static inline void N_VScale(sunrealtype c, N_Vector x, N_Vector z) { z->ops->nvscale(c, x, z); }
static inline void N_VAbs(N_Vector x, N_Vector z) { z->ops->nvabs(x, z); }
static inline void N_VInv(N_Vector x, N_Vector z) { z->ops->nvinv(x, z); }
static inline void N_VAddConst(N_Vector x, sunrealtype b, N_Vector z) { z->ops->nvaddconst(x, b, z); }
static inline sunrealtype N_VDotProd(N_Vector x, N_Vector y) { return y->ops->nvdotprod(x, y); }
static inline sunrealtype N_VMaxNorm(N_Vector x) { return x->ops->nvmaxnorm(x); }
static inline sunrealtype N_VWrmsNorm(N_Vector x, N_Vector w) { return x->ops->nvwrmsnorm(x, w); }
static inline sunrealtype N_VWrmsNormMask(N_Vector x, N_Vector w, N_Vector id) {
  return x->ops->nvwrmsnormmask(x, w, id);
}
static inline sunrealtype N_VMin(N_Vector x) { return x->ops->nvmin(x); }
static inline sunrealtype N_VWL2Norm(N_Vector x, N_Vector w) { return x->ops->nvwl2norm(x, w); }
static inline sunrealtype N_VL1Norm(N_Vector x) { return x->ops->nvl1norm(x); }
static inline void N_VCompare(sunrealtype c, N_Vector x, N_Vector z) { z->ops->nvcompare(c, x, z); }
static inline sunbooleantype N_VInvTest(N_Vector x, N_Vector z) { return z->ops->nvinvtest(x, z); }
static inline sunbooleantype N_VConstrMask(N_Vector c, N_Vector x, N_Vector m) {
  return x->ops->nvconstrmask(c, x, m);
}
static inline sunrealtype N_VMinQuotient(N_Vector num, N_Vector denom) {
  return num->ops->nvminquotient(num, denom);
}

/* Fused ops fall back to sequences of standard ops when the vector does not
 * provide them (the behaviour CVODE relies on when a fused pointer is NULL,
 * e.g. the reference's v4 override at 2d_i4/deferred_nvector.cu:193-196). */
static inline SUNErrCode N_VLinearCombination(int nv, sunrealtype* c, N_Vector* X, N_Vector z) {
  if (z->ops->nvlinearcombination) return z->ops->nvlinearcombination(nv, c, X, z);
  if (nv < 1) return SUN_SUCCESS;
  N_VScale(c[0], X[0], z);
  for (int i = 1; i < nv; i++) N_VLinearSum(1.0, z, c[i], X[i], z);
  return SUN_SUCCESS;
}
static inline SUNErrCode N_VScaleAddMulti(int nv, sunrealtype* a, N_Vector x, N_Vector* Y, N_Vector* Z) {
  if (x->ops->nvscaleaddmulti) return x->ops->nvscaleaddmulti(nv, a, x, Y, Z);
  for (int i = 0; i < nv; i++) N_VLinearSum(a[i], x, 1.0, Y[i], Z[i]);
  return SUN_SUCCESS;
}
static inline SUNErrCode N_VDotProdMulti(int nv, N_Vector x, N_Vector* Y, sunrealtype* d) {
  if (x->ops->nvdotprodmulti) return x->ops->nvdotprodmulti(nv, x, Y, d);
  for (int i = 0; i < nv; i++) d[i] = N_VDotProd(x, Y[i]);
  return SUN_SUCCESS;
}
static inline SUNErrCode N_VScaleVectorArray(int nv, sunrealtype* c, N_Vector* X, N_Vector* Z) {
  if (nv < 1) return SUN_SUCCESS;
  if (X[0]->ops->nvscalevectorarray) return X[0]->ops->nvscalevectorarray(nv, c, X, Z);
  for (int i = 0; i < nv; i++) N_VScale(c[i], X[i], Z[i]);
  return SUN_SUCCESS;
}
static inline SUNErrCode N_VWrmsNormVectorArray(int nv, N_Vector* X, N_Vector* W, sunrealtype* nrm) {
  if (nv < 1) return SUN_SUCCESS;
  if (X[0]->ops->nvwrmsnormvectorarray) return X[0]->ops->nvwrmsnormvectorarray(nv, X, W, nrm);
  for (int i = 0; i < nv; i++) nrm[i] = N_VWrmsNorm(X[i], W[i]);
  return SUN_SUCCESS;
}
static inline SUNErrCode N_VLinearSumVectorArray(int nv, sunrealtype a, N_Vector* X, sunrealtype b,
                                                 N_Vector* Y, N_Vector* Z) {
  if (nv < 1) return SUN_SUCCESS;
  if (X[0]->ops->nvlinearsumvectorarray)
    return X[0]->ops->nvlinearsumvectorarray(nv, a, X, b, Y, Z);
  for (int i = 0; i < nv; i++) N_VLinearSum(a, X[i], b, Y[i], Z[i]);
  return SUN_SUCCESS;
}

#ifdef __cplusplus
}
#endif
#endif /* LLG_SUNDIALS_COMPAT_H */
