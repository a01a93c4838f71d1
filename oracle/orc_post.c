/*
 * orc_post.c -- posterior and colouring in FP32 (reference arithmetic) and FP64.
 * Test infrastructure only; see oracle.h and orc_post_impl.h for the restated code.
 */
#include <math.h>
#include <stdint.h>
#include "oracle.h"

#define ORC_MAXT 16
#define ORC_TOL 1e-20 /* lib_matr.c:19 */

#define REAL float
#define SUF f32
#include "orc_post_impl.h"
#undef REAL
#undef SUF

#define REAL double
#define SUF f64
#include "orc_post_impl.h"
#undef REAL
#undef SUF
