/*
 * CPU oracle for the BM3D / V-BM3D hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithm
 * (WolframRhodium/VapourSynth-BM3DCUDA: cpu_source/source.cpp:91-911 and
 * source/kernel.cu:55-757), used by tests/, by smoke() and by the CPU
 * baseline leg of bench.py, to check results against. The product never
 * links it.
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

/* 8-point transforms other than the DCT, as 8 x 8 matrices M (out[r] =
 * sum_k M[r][k] v[k]), from the transform definitions of the NVRTC variant
 * (rtc_source/kernel.hpp:200-361), each 4 x an orthonormal (or, for
 * bior1.5, biorthogonal) basis so that inverse(forward(x)) = 16 x:
 *   haar     rows: scaling function, then the Haar wavelets of widths 8, 4,
 *            2. The reference's inverse (kernel.hpp:239-259) reads
 *            coefficient 2 where coefficient 3 belongs and coefficient 4
 *            where 5, 6 and 7 belong, so it is not the inverse of the
 *            forward transform; ORACLE_HAAR_INV restates that map as it is.
 *   wht      sqrt(2) x the Walsh (sequency-ordered) Hadamard rows; the same
 *            matrix both ways (built in xform_matrix).
 *   bior1.5  the forward / inverse pair of kernel.hpp:296-361.
 * Index: [kind - 1][inverse]. The WHT slot is filled at first use. */
#define XS2 1.41421356237309504880168872420969808
#define XS8 2.82842712474619009760337744841939616
#define BA 0.877670597010003062405456290501163901200537360
#define BB 1.797135031972863413496886690073811797696338403
#define BC 2.277437593371384746027611934034934695963515886
#define BD 1.609389232649111887192845766718020518480884560
#define BE 0.334024180361136429417383083658457088741315663
#define BG 1.495435764250674860795011090214036706658653686
#define BH 2.058234225009388964222454285384072231477027482
#define BI 0.486135912065751423025580498947083714508324707
static double oracle_xform[3][2][64] = {
    {{XS2, XS2, XS2, XS2, XS2, XS2, XS2, XS2,
      XS2, XS2, XS2, XS2, -XS2, -XS2, -XS2, -XS2,
      2, 2, -2, -2, 0, 0, 0, 0,
      0, 0, 0, 0, 2, 2, -2, -2,
      XS8, -XS8, 0, 0, 0, 0, 0, 0,
      0, 0, XS8, -XS8, 0, 0, 0, 0,
      0, 0, 0, 0, XS8, -XS8, 0, 0,
      0, 0, 0, 0, 0, 0, XS8, -XS8},
     /* ORACLE_HAAR_INV */
     {XS2, XS2, 2, 0, XS8, 0, 0, 0,
      XS2, XS2, 2, 0, -XS8, 0, 0, 0,
      XS2, XS2, -2, 0, XS8, 0, 0, 0,
      XS2, XS2, -2, 0, -XS8, 0, 0, 0,
      XS2, -XS2, 2, 0, XS8, 0, 0, 0,
      XS2, -XS2, 2, 0, -XS8, 0, 0, 0,
      XS2, -XS2, -2, 0, XS8, 0, 0, 0,
      XS2, -XS2, -2, 0, -XS8, 0, 0, 0}},
    {{0}, {0}},
    {{XS2, XS2, XS2, XS2, XS2, XS2, XS2, XS2,
      BA, BB, BB, BA, -BA, -BB, -BB, -BA,
      BC, BD, -BD, -BC, -BE, BE, -BE, BE,
      -BE, BE, -BE, BE, BC, BD, -BD, -BC,
      XS8, -XS8, 0, 0, 0, 0, 0, 0,
      0, 0, XS8, -XS8, 0, 0, 0, 0,
      0, 0, 0, 0, XS8, -XS8, 0, 0,
      0, 0, 0, 0, 0, 0, XS8, -XS8},
     {XS2, BG, BH, 0, XS8, -BI, 0, BI,
      XS2, BG, BH, 0, -XS8, -BI, 0, BI,
      XS2, BG, -BH, 0, BI, XS8, -BI, 0,
      XS2, BG, -BH, 0, BI, -XS8, -BI, 0,
      XS2, -BG, 0, BH, 0, BI, XS8, -BI,
      XS2, -BG, 0, BH, 0, BI, -XS8, -BI,
      XS2, -BG, 0, -BH, -BI, 0, BI, XS8,
      XS2, -BG, 0, -BH, -BI, 0, BI, -XS8}},
};

/* Matrix of a non-DCT transform kind (1 haar, 2 wht, 3 bior1.5). Walsh row
 * k is Hadamard row bitrev3(gray(k)): the sign of entry n is
 * (-1)^popcount(row & n). */
static const double *xform_matrix(int kind, int inverse) {
    if (kind == 2 && oracle_xform[1][0][0] == 0) {
        for (int k = 0; k < 8; ++k) {
            int g = k ^ (k >> 1);
            int row = ((g & 1) << 2) | (g & 2) | ((g >> 2) & 1);
            for (int n = 0; n < 8; ++n) {
                int bits = row & n, odd = 0;
                while (bits) { odd ^= bits & 1; bits >>= 1; }
                oracle_xform[1][0][k * 8 + n] = odd ? -XS2 : XS2;
                oracle_xform[1][1][k * 8 + n] = odd ? -XS2 : XS2;
            }
        }
    }
    return oracle_xform[kind - 1][inverse ? 1 : 0];
}

/* The matrix of transform `kind` (0 dct: built from the cosine formula) into
 * out[64], for tests. */
void oracle_xform_matrix(int kind, int inverse, double *out) {
    if (kind == 0) {
        const double pi = 3.14159265358979323846264338327950288;
        for (int r = 0; r < 8; ++r)
            for (int k = 0; k < 8; ++k) {
                double f = r == 0 ? XS2 : 2.0 * cos((2 * k + 1) * r * pi / 16.0);
                if (inverse) out[k * 8 + r] = f;
                else out[r * 8 + k] = f;
            }
        return;
    }
    memcpy(out, xform_matrix(kind, inverse), sizeof(double) * 64);
}

#define T float
#define SFX(name) name##_f32
#define FMA fmaf
#define FABS fabsf
#define SQRT sqrtf
#include "bm3d_oracle_body.h"
#undef T
#undef SFX
#undef FMA
#undef FABS
#undef SQRT

#define T double
#define SFX(name) name##_f64
#define FMA fma
#define FABS fabs
#define SQRT sqrt
#include "bm3d_oracle_body.h"
#undef T
#undef SFX
#undef FMA
#undef FABS
#undef SQRT

/* Pre-rounding helper exposed for tests: (v + E) - E in float. */
float oracle_preround_f32(float v, float E) { return (v + E) - E; }
