/* CPU alpha-blending oracle -- TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
 * Built by oracle/build.py into oracle/liboracle_blend.so.  Exports
 * blend_forward_f32 / blend_backward_f32 (float arithmetic in the kernels'
 * operation order, libm expf) and the _f64 pair (double arithmetic), and
 * _mt variants of each that share tile rows among OpenMP threads (the CPU
 * baseline). */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REAL float
#define IN float
#define SFX _f32
#define EXPF expf
#include "blend_body.h"
#undef REAL
#undef IN
#undef SFX
#undef EXPF

#define REAL double
#define IN double
#define SFX _f64
#define EXPF exp
#include "blend_body.h"
#undef REAL
#undef IN
#undef SFX
#undef EXPF
