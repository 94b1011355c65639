/*
 * Test oracle — NOT product code (see oracle.h).  Builds the FP32 and FP64
 * twins of the reference restatement from oracle_body.h.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define R float
#define SUF f32
#define ORC_IS_F64 0
#define FMA_ fmaf
#define SQRT_ sqrtf
#define FMIN_ fminf
#define FMAX_ fmaxf
#define FABS_ fabsf
#include "oracle_body.h"
#undef R
#undef SUF
#undef ORC_IS_F64
#undef FMA_
#undef SQRT_
#undef FMIN_
#undef FMAX_
#undef FABS_

#define R double
#define SUF f64
#define ORC_IS_F64 1
#define FMA_ fma
#define SQRT_ sqrt
#define FMIN_ fmin
#define FMAX_ fmax
#define FABS_ fabs
#include "oracle_body.h"
