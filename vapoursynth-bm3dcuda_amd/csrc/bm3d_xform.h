// 8-point transforms other than the DCT (the transform_2d_s / transform_1d_s
// options of the NVRTC variant, rtc_source/kernel.hpp:200-361), as 8 x 8
// matrices M with out[r] = sum_k M[r][k] v[k]. Each is 4 x an orthonormal
// (bior1.5: biorthogonal) basis, like the DCT pair:
//
//   haar     rows: the scaling function, then the Haar wavelets of widths 8,
//            4, 2, in the reference's output order (kernel.hpp:225-232). Its
//            inverse (kernel.hpp:239-259) reads coefficient 2 where 3 belongs
//            and coefficient 4 where 5, 6 and 7 belong; the matrix below is
//            that map as the reference computes it, so results match the
//            reference, not a true inverse.
//   wht      sqrt(2) x the Walsh-ordered (by sign changes) Hadamard rows:
//            row k is Hadamard row bitrev3(gray(k)); its own inverse.
//   bior1.5  the biorthogonal spline 1.5 pair of kernel.hpp:296-361.
//
// Internal. Host code reads the double table; the FP32 filter kernel a
// float copy in constant memory.
#pragma once

namespace bm3d {

struct XformMatrices {
    double m[3][2][64];  // [kind - 1][inverse][r * 8 + k]
};

constexpr XformMatrices make_xform_matrices() {
    constexpr double s2 = 1.41421356237309504880168872420969808;
    constexpr double s8 = 2.82842712474619009760337744841939616;
    constexpr double ba = 0.877670597010003062405456290501163901200537360;
    constexpr double bb = 1.797135031972863413496886690073811797696338403;
    constexpr double bc = 2.277437593371384746027611934034934695963515886;
    constexpr double bd = 1.609389232649111887192845766718020518480884560;
    constexpr double be = 0.334024180361136429417383083658457088741315663;
    constexpr double bg = 1.495435764250674860795011090214036706658653686;
    constexpr double bh = 2.058234225009388964222454285384072231477027482;
    constexpr double bi = 0.486135912065751423025580498947083714508324707;
    XformMatrices t{};
    double* hf = t.m[0][0];
    double* hi = t.m[0][1];
    double* bf = t.m[2][0];
    double* bv = t.m[2][1];
    for (int k = 0; k < 8; ++k) {
        // scaling function, then the width-8 Haar wavelet
        hf[k] = s2;
        hf[8 + k] = k < 4 ? s2 : -s2;
        bf[k] = s2;
    }
    // width-4 Haar wavelets
    for (int k = 0; k < 4; ++k) {
        hf[16 + k] = k < 2 ? 2.0 : -2.0;
        hf[24 + 4 + k] = k < 2 ? 2.0 : -2.0;
    }
    // width-2 Haar wavelets (also the finest bior1.5 analysis rows)
    for (int r = 0; r < 4; ++r) {
        hf[(4 + r) * 8 + 2 * r] = s8;
        hf[(4 + r) * 8 + 2 * r + 1] = -s8;
        bf[(4 + r) * 8 + 2 * r] = s8;
        bf[(4 + r) * 8 + 2 * r + 1] = -s8;
    }
    // the reference's Haar inverse: out[n] = s2 c0 +- s2 c1 +- 2 c2 +- s8 c4
    for (int n = 0; n < 8; ++n) {
        hi[n * 8 + 0] = s2;
        hi[n * 8 + 1] = n < 4 ? s2 : -s2;
        hi[n * 8 + 2] = (n & 2) ? -2.0 : 2.0;
        hi[n * 8 + 4] = (n & 1) ? -s8 : s8;
    }
    // bior1.5 analysis rows 1-3
    const double r1[8] = {ba, bb, bb, ba, -ba, -bb, -bb, -ba};
    const double r2[8] = {bc, bd, -bd, -bc, -be, be, -be, be};
    const double r3[8] = {-be, be, -be, be, bc, bd, -bd, -bc};
    for (int k = 0; k < 8; ++k) {
        bf[8 + k] = r1[k];
        bf[16 + k] = r2[k];
        bf[24 + k] = r3[k];
    }
    // bior1.5 synthesis
    const double inv[8][8] = {
        {s2, bg, bh, 0, s8, -bi, 0, bi},   {s2, bg, bh, 0, -s8, -bi, 0, bi},
        {s2, bg, -bh, 0, bi, s8, -bi, 0},  {s2, bg, -bh, 0, bi, -s8, -bi, 0},
        {s2, -bg, 0, bh, 0, bi, s8, -bi},  {s2, -bg, 0, bh, 0, bi, -s8, -bi},
        {s2, -bg, 0, -bh, -bi, 0, bi, s8}, {s2, -bg, 0, -bh, -bi, 0, bi, -s8},
    };
    for (int r = 0; r < 8; ++r)
        for (int k = 0; k < 8; ++k) bv[r * 8 + k] = inv[r][k];
    // Walsh-ordered Hadamard, both directions
    for (int k = 0; k < 8; ++k) {
        const int g = k ^ (k >> 1);
        const int row = ((g & 1) << 2) | (g & 2) | ((g >> 2) & 1);
        for (int n = 0; n < 8; ++n) {
            const int b = row & n;
            const bool odd = ((b ^ (b >> 1) ^ (b >> 2)) & 1) != 0;
            t.m[1][0][k * 8 + n] = odd ? -s2 : s2;
            t.m[1][1][k * 8 + n] = odd ? -s2 : s2;
        }
    }
    return t;
}

constexpr XformMatrices kXformMatrices = make_xform_matrices();

}  // namespace bm3d
