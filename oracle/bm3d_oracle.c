/*
 * CPU oracle for the BM3D / V-BM3D hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithm
 * (WolframRhodium/VapourSynth-BM3DCUDA: cpu_source/source.cpp:91-911 and
 * source/kernel.cu:55-757), used by tests/, by smoke() and by the CPU
 * baseline leg of bench.py as the checker. The product never links it.
 *
 * Built with -ffp-contract=off so that every fused multiply-add is an
 * explicit FMA()/fma() call and every other product and sum rounds on its
 * own, exactly as written.
 *
 * Pinning (see tests/test_oracle.py): the DCT against the cosine matrix,
 * the distance against a float64 brute force, the top-8 selection against
 * an order-free lexicographic sort on (error, scan index), the temporal
 * predictive search against an independent numpy restatement on small
 * inputs, shrinkage and aggregation against closed forms (constant
 * image, sigma -> 0 identity, DC-only group mean).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define DCT_R2 1.41421356237309504880168872420969808
#define DCT_K1 1.84775906502257351225636637879357657 /* 2 cos(pi/8)   */
#define DCT_K3 0.76536686473017954345691996806916186 /* 2 cos(3pi/8)  */
#define DCT_C1 1.96157056080646086901156744484086917 /* 2 cos(pi/16)  */
#define DCT_C3 1.66293922460509004760575233432468773 /* 2 cos(3pi/16) */
#define DCT_C5 1.11114046603920391150735136234616736 /* 2 cos(5pi/16) */
#define DCT_C7 0.39018064403225654488047858302562963 /* 2 cos(7pi/16) */

#define T float
#define SFX(name) name##_f32
#define FMA fmaf
#define FABS fabsf
#include "bm3d_oracle_body.h"
#undef T
#undef SFX
#undef FMA
#undef FABS

#define T double
#define SFX(name) name##_f64
#define FMA fma
#define FABS fabs
#include "bm3d_oracle_body.h"
#undef T
#undef SFX
#undef FMA
#undef FABS

/* Pre-rounding helper exposed for tests: (v + E) - E in float. */
float oracle_preround_f32(float v, float E) { return (v + E) - E; }
