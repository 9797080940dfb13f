// tgk_plain_model.h -- the reference-semantics (plain FP32) local models and
// orientation updates, device side.
//
// Results are bit-identical to a literal FP32 reading of the reference OpenCL
// (every product and sum rounded in source order, no contraction) given the
// elementary functions of tgk_math.h. Include only from translation units
// compiled with -ffp-contract=off.
#pragma once

#include "tgk_device.h"

namespace tgk {
namespace plain {

struct ModelValue {
    float value, dtheta, dphi;
};

// core.cl:142-153
__device__ __forceinline__ void cart2sph(V3 c, float &azimuth, float &colatitude)
{
    const float radius = length3(c);
    if (radius <= 0.0f) {
        azimuth = 0.0f;
        colatitude = 0.0f;
        return;
    }
    float az = tgk_atan2f(c.y, c.x);
    if (az < 0.0f)
        az += 2.0f * 3.14159265359f;
    azimuth = az;
    colatitude = tgk_acosf(fminf(fmaxf(c.z / radius, -1.0f), 1.0f));
}

// core.cl:101-109
__device__ __forceinline__ V3 exps2(V3 p, V3 x)
{
    const float n = length3(x);
    if (n <= 0.0f)
        return p;
    float s, c;
    tgk_sincosf(n, &s, &c);
    const float q = s / n;
    return v3(c * p.x + x.x * q, c * p.y + x.y * q, c * p.z + x.z * q);
}

// core.cl:111-120, k = 100
__device__ __forceinline__ float softplus(float x) { return x * 100.0f < 30.0f ? tgk_logf(1.0f + tgk_expf(100.0f * x)) / 100.0f : x; }
__device__ __forceinline__ float dsoftplus(float x) { return 1.0f / (1.0f + tgk_expf(-100.0f * x)); }

// core.cl:240-245
__device__ __forceinline__ float randn(uint2 &s)
{
    float n = randu(s);
    while (n <= 0.0f)
        n = randu(s);
    return tgk_sqrtf(-2.0f * tgk_logf(n)) * tgk_cosf((2.0f * 3.14159265359f) * randu(s));
}

// The l <= 8 real SH basis (45 terms) and its two angular derivatives
// (spharm.cl:216-413; the values alone are spharm.cl:10-207, the same
// expressions), handed to visit(i, Y_i, dY_i/dphi / sin(theta), dY_i/dtheta)
// in index order. The basis is generated band by band: for each (l, m > 0)
// the reference's prefactor of the value (A), of the azimuth derivative over
// sin(theta) (D) and of the colatitude derivative (T) is formed once and
// multiplied by sin(m phi) for index l(l+1)/2 - m and by cos(m phi) for index
// l(l+1)/2 + m; the D term of the cos partner is -(D sin(m phi)), equal bit
// for bit to the reference's negated-constant form. A visitor that ignores
// the derivatives lets the compiler drop them.
template <class Visit>
__device__ __forceinline__ void sh45_terms(float phi, float theta, Visit &&visit)
{
    float st, ct, sp, cp, s2p, c2p;
    tgk_sincosf(theta, &st, &ct);
    tgk_sincosf(phi, &sp, &cp);
    tgk_sincosf(2.0f * phi, &s2p, &c2p);
    const float s3p = 2.0f * cp * s2p - sp, c3p = 2.0f * cp * c2p - cp;
    const float s4p = 2.0f * cp * s3p - s2p, c4p = 2.0f * cp * c3p - c2p;
    const float s5p = 2.0f * cp * s4p - s3p, c5p = 2.0f * cp * c4p - c3p;
    const float s6p = 2.0f * cp * s5p - s4p, c6p = 2.0f * cp * c5p - c4p;
    const float s7p = 2.0f * cp * s6p - s5p, c7p = 2.0f * cp * c6p - c5p;
    const float s8p = 2.0f * cp * s7p - s6p, c8p = 2.0f * cp * c7p - c6p;

    const float st2 = st * st, st3 = st2 * st, st4 = st3 * st, st5 = st4 * st;
    const float st6 = st5 * st, st7 = st6 * st, st8 = st7 * st;
    const float ct2 = ct * ct, ct3 = ct2 * ct, ct4 = ct3 * ct, ct5 = ct4 * ct;
    const float ct6 = ct5 * ct, ct7 = ct6 * ct, ct8 = ct7 * ct;
    const float stct = st * ct;

#define TGK_ACC(i, Y, DP, DT) visit(i, (Y), (DP), (DT));
#define TGK_PAIR(ineg, ipos, A, D, T, S, C) \
    TGK_ACC(ineg, (A) * (S), (D) * (C), (T) * (S))
#define TGK_PAIR_POS(ipos, A, D, T, S, C) TGK_ACC(ipos, (A) * (C), -((D) * (S)), (T) * (C))

    // l = 0
    TGK_ACC(0, 0.28209479177387814f, 0.0f, 0.0f);

    // l = 2 (indices 1..5)
    {
        const float A2 = 0.54627421529f * st2, D2 = 0.54627421529f * st * 2.0f, T2 = 0.54627421529f * 2.0f * stct;
        const float A1 = -1.0925484305920792f * stct, D1 = -1.0925484305920792f * ct;
        const float T1 = -1.0925484305920792f * (ct2 - st2);
        TGK_PAIR(1, 5, A2, D2, T2, s2p, c2p);
        TGK_PAIR(2, 4, A1, D1, T1, sp, cp);
        TGK_ACC(3, 0.31539156525252005f * (3.0f * ct2 - 1.0f), 0.0f, -0.31539156525252005f * 6.0f * ct * st);
        TGK_PAIR_POS(4, A1, D1, T1, sp, cp);
        TGK_PAIR_POS(5, A2, D2, T2, s2p, c2p);
    }

    // l = 4 (indices 6..14)
    {
        const float A4 = 0.62583573544f * st4, D4 = 0.62583573544f * st3 * 4.0f;
        const float T4 = 0.62583573544f * 4.0f * st3 * ct;
        const float A3 = -1.77013076978f * st3 * ct, D3 = -1.77013076978f * st2 * ct * 3.0f;
        const float T3 = -1.77013076978f * (3.0f * st2 * ct2 - st4);
        const float q2 = 7.0f * ct2 - 1.0f;
        const float A2 = 0.47308734787f * st2 * q2, D2 = 0.47308734787f * st * q2 * 2.0f;
        const float T2 = 0.47308734787f * (2.0f * stct * q2 - 14.0f * st3 * ct);
        const float q1 = 7.0f * ct2 - 3.0f;
        const float A1 = -0.66904654355f * st * q1 * ct, D1 = -0.66904654355f * ct * q1;
        const float T1 = -0.66904654355f * ((ct2 - st2) * q1 - (14.0f * st2 * ct2));
        TGK_PAIR(6, 14, A4, D4, T4, s4p, c4p);
        TGK_PAIR(7, 13, A3, D3, T3, s3p, c3p);
        TGK_PAIR(8, 12, A2, D2, T2, s2p, c2p);
        TGK_PAIR(9, 11, A1, D1, T1, sp, cp);
        TGK_ACC(10, 0.10578554691520431f * (35.0f * ct2 * ct2 - 30.0f * ct2 + 3.0f), 0.0f,
                0.10578554691520431f * (-35.0f * 4.0f * ct2 * stct + 30.0f * 2.0f * stct));
        TGK_PAIR_POS(11, A1, D1, T1, sp, cp);
        TGK_PAIR_POS(12, A2, D2, T2, s2p, c2p);
        TGK_PAIR_POS(13, A3, D3, T3, s3p, c3p);
        TGK_PAIR_POS(14, A4, D4, T4, s4p, c4p);
    }

    // l = 6 (indices 15..27)
    {
        const float A6 = 0.68318410519f * st6, D6 = 0.68318410519f * st5 * 6.0f;
        const float T6 = 0.68318410519f * 6.0f * st5 * ct;
        const float A5 = -2.36661916223f * st5 * ct, D5 = -2.36661916223f * st4 * ct * 5.0f;
        const float T5 = -2.36661916223f * (5.0f * st4 * ct2 - st6);
        const float q4 = 11.0f * ct2 - 1.0f;
        const float A4 = 0.50456490072f * st4 * q4, D4 = 0.50456490072f * st3 * q4 * 4.0f;
        const float T4 = 0.50456490072f * (4.0f * st3 * ct * q4 - 22.0f * st5 * ct);
        const float q3 = 11.0f * ct3 - 3.0f * ct;
        const float A3 = -0.92120525951f * st3 * q3, D3 = -0.92120525951f * st2 * q3 * 3.0f;
        const float T3 = -0.92120525951f * (3.0f * st2 * ct * q3 - st4 * (33.0f * ct2 - 3.0f));
        const float q2 = 33.0f * ct4 - 18.0f * ct2 + 1.0f;
        const float A2 = 0.46060262975f * st2 * q2, D2 = 0.46060262975f * st * q2 * 2.0f;
        const float T2 = 0.46060262975f * (2.0f * st * ct * q2 - st3 * (33.0f * 4.0f * ct3 - 36.0f * ct));
        const float q1 = 33.0f * ct5 - 30.0f * ct3 + 5.0f * ct;
        const float A1 = -0.58262136251f * st * q1, D1 = -0.58262136251f * q1;
        const float T1 = -0.58262136251f * (ct * q1 - st2 * (33.0f * 5.0f * ct4 - 90.0f * ct2 + 5.0f));
        TGK_PAIR(15, 27, A6, D6, T6, s6p, c6p);
        TGK_PAIR(16, 26, A5, D5, T5, s5p, c5p);
        TGK_PAIR(17, 25, A4, D4, T4, s4p, c4p);
        TGK_PAIR(18, 24, A3, D3, T3, s3p, c3p);
        TGK_PAIR(19, 23, A2, D2, T2, s2p, c2p);
        TGK_PAIR(20, 22, A1, D1, T1, sp, cp);
        TGK_ACC(21, 0.06356920226f * (231.0f * ct6 - 315.0f * ct4 + 105.0f * ct2 - 5.0f), 0.0f,
                -0.06356920226f * (1386.0f * ct5 - 1260.0f * ct3 + 210.0f * ct) * st);
        TGK_PAIR_POS(22, A1, D1, T1, sp, cp);
        TGK_PAIR_POS(23, A2, D2, T2, s2p, c2p);
        TGK_PAIR_POS(24, A3, D3, T3, s3p, c3p);
        TGK_PAIR_POS(25, A4, D4, T4, s4p, c4p);
        TGK_PAIR_POS(26, A5, D5, T5, s5p, c5p);
        TGK_PAIR_POS(27, A6, D6, T6, s6p, c6p);
    }

    // l = 8 (indices 28..44)
    {
        const float A8 = 0.72892666017f * st8, D8 = 0.72892666017f * st7 * 8.0f;
        const float T8 = 0.72892666017f * 8.0f * st7 * ct;
        const float A7 = -2.9157066407f * st7 * ct, D7 = -2.9157066407f * st6 * ct * 7.0f;
        const float T7 = -2.9157066407f * (7.0f * st6 * ct2 - st8);
        const float q6 = 15.0f * ct2 - 1.0f;
        const float A6 = 0.53233276606f * st6 * q6, D6 = 0.53233276606f * st5 * q6 * 6.0f;
        const float T6 = 0.53233276606f * (6.0f * st5 * ct * q6 - st7 * 30.0f * ct);
        const float q5 = 5.0f * ct3 - ct;
        const float A5 = -3.4499106221f * st5 * q5, D5 = -3.4499106221f * st4 * q5 * 5.0f;
        const float T5 = -3.4499106221f * (5.0f * st4 * ct * q5 - st6 * q6);
        const float q4 = 65.0f * ct4 - 26.0f * ct2 + 1.0f;
        const float A4 = 0.47841652475f * st4 * q4, D4 = 0.47841652475f * st3 * q4 * 4.0f;
        const float T4 = 0.47841652475f * (4.0f * st3 * ct * q4 - st5 * (260.0f * ct3 - 52.0f * ct));
        const float q3 = 39.0f * ct5 - 26.0f * ct3 + 3.0f * ct;
        const float A3 = -1.2352661553f * st3 * q3, D3 = -1.2352661553f * st2 * q3 * 3.0f;
        const float T3 = -1.2352661553f * (3.0f * st2 * ct * q3 - st4 * (195.0f * ct4 - 78.0f * ct2 + 3.0f));
        const float q2 = 143.0f * ct6 - 143.0f * ct4 + 33.0f * ct2 - 1.0f;
        const float A2 = 0.45615225843f * st2 * q2, D2 = 0.45615225843f * st * q2 * 2.0f;
        const float T2 = 0.45615225843f * (2.0f * st * ct * q2 - st3 * (858.0f * ct5 - 572.0f * ct3 + 66.0f * ct));
        const float q1 = 715.0f * ct7 - 1001.0f * ct5 + 385.0f * ct3 - 35.0f * ct;
        const float A1 = -0.10904124589f * st * q1, D1 = -0.10904124589f * q1;
        const float T1 =
            -0.10904124589f * (ct * q1 - st2 * (5005.0f * ct6 - 5005.0f * ct4 + 1155.0f * ct2 - 35.0f));
        TGK_PAIR(28, 44, A8, D8, T8, s8p, c8p);
        TGK_PAIR(29, 43, A7, D7, T7, s7p, c7p);
        TGK_PAIR(30, 42, A6, D6, T6, s6p, c6p);
        TGK_PAIR(31, 41, A5, D5, T5, s5p, c5p);
        TGK_PAIR(32, 40, A4, D4, T4, s4p, c4p);
        TGK_PAIR(33, 39, A3, D3, T3, s3p, c3p);
        TGK_PAIR(34, 38, A2, D2, T2, s2p, c2p);
        TGK_PAIR(35, 37, A1, D1, T1, sp, cp);
        TGK_ACC(36, 0.00908677049f * (6435.0f * ct8 - 12012.0f * ct6 + 6930.0f * ct4 - 1260.0f * ct2 + 35.0f), 0.0f,
                -0.00908677049f *
                    (8.0f * 6435.0f * ct7 - 6.0f * 12012.0f * ct5 + 4.0f * 6930.0f * ct3 - 2.0f * 1260.0f * ct) * st);
        TGK_PAIR_POS(37, A1, D1, T1, sp, cp);
        TGK_PAIR_POS(38, A2, D2, T2, s2p, c2p);
        TGK_PAIR_POS(39, A3, D3, T3, s3p, c3p);
        TGK_PAIR_POS(40, A4, D4, T4, s4p, c4p);
        TGK_PAIR_POS(41, A5, D5, T5, s5p, c5p);
        TGK_PAIR_POS(42, A6, D6, T6, s6p, c6p);
        TGK_PAIR_POS(43, A7, D7, T7, s7p, c7p);
        TGK_PAIR_POS(44, A8, D8, T8, s8p, c8p);
    }
#undef TGK_PAIR_POS
#undef TGK_PAIR
#undef TGK_ACC
}

// No use for the basis values beyond the contraction.
struct NoTerms {
    __device__ __forceinline__ void operator()(int, float) const {}
};

// Model value and angular derivatives: terms accumulated in index order, so
// sums round exactly as the reference's `for (i < K)` loop. Coefficients past
// K arrive as +0 and leave the sums unchanged. `terms(i, Y_i)` sees every
// basis value on the way (all 45, whatever K).
template <class Terms = NoTerms>
__device__ __forceinline__ void sh45_contract(const float c[45], float phi, float theta, float &v, float &th,
                                              float &ph, Terms &&terms = Terms{})
{
    float sv = 0.0f, sth = 0.0f, sph = 0.0f;
    sh45_terms(phi, theta, [&](int i, float y, float dp, float dt) {
        const float ci = c[i];
        sv += ci * y;
        sth += ci * dt;
        sph += ci * dp;
        terms(i, y);
    });
    v = sv;
    th = sth;
    ph = sph;
}

// Loads the first K of a voxel's coefficients (K <= 45, rows 16-byte aligned)
// as float4 words; entries from K on are +0.
template <int KT>
__device__ __forceinline__ void load_sh_coefficients(const float *__restrict__ row, int K, float c[45])
{
    const float4 *row4 = reinterpret_cast<const float4 *>(row);
    if constexpr (KT == 45) {
        float4 w[12];
#pragma unroll
        for (int j = 0; j < 12; j++)
            w[j] = row4[j];
#pragma unroll
        for (int j = 0; j < 11; j++) {
            c[4 * j] = w[j].x;
            c[4 * j + 1] = w[j].y;
            c[4 * j + 2] = w[j].z;
            c[4 * j + 3] = w[j].w;
        }
        c[44] = w[11].x;
    } else {
        const int words = (K + 3) >> 2;
#pragma unroll
        for (int j = 0; j < 12; j++) {
            const float4 w = j < words ? row4[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const float e[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (4 * j + q < 45)
                    c[4 * j + q] = (4 * j + q) < K ? e[q] : 0.0f;
        }
    }
}

// local.py:91-120 (SH with softplus, K <= 45)
template <int KT, class Terms = NoTerms>
__device__ __forceinline__ ModelValue sh_model(const float *__restrict__ row, int K, V3 o, Terms &&terms = Terms{})
{
    float c[45];
    load_sh_coefficients<KT>(row, K, c);
    float az, col;
    cart2sph(o, az, col);
    float v, th, ph;
    sh45_contract(c, az, col, v, th, ph, terms);
    const float d = dsoftplus(v);
    ModelValue r;
    r.value = softplus(v);
    r.dtheta = th * d;
    r.dphi = ph * d;
    return r;
}

// local.py:212-230 (DTI: Dxx Dyy Dzz Dxy Dxz Dyz)
__device__ __forceinline__ ModelValue dti_model(const float *__restrict__ row, V3 o)
{
    const float4 w0 = reinterpret_cast<const float4 *>(row)[0];
    const float4 w1 = reinterpret_cast<const float4 *>(row)[1];
    const float m0 = w0.x, m1 = w0.y, m2 = w0.z, m3 = w0.w, m4 = w1.x, m5 = w1.y;
    const float t0 = (m0 * o.x + m3 * o.y) + m4 * o.z;
    const float t1 = (m3 * o.x + m1 * o.y) + m5 * o.z;
    const float t2 = (m4 * o.x + m5 * o.y) + m2 * o.z;
    float az, col, sp, cp, st, ct;
    cart2sph(o, az, col);
    tgk_sincosf(az, &sp, &cp);
    tgk_sincosf(col, &st, &ct);
    const float etx = ct * cp, ety = ct * sp, etz = -st;
    const float epx = -sp, epy = cp, epz = 0.0f;
    ModelValue r;
    r.value = (o.x * t0 + o.y * t1) + o.z * t2;
    r.dtheta = 2.0f * ((t0 * etx + t1 * ety) + t2 * etz);
    r.dphi = 2.0f * ((t0 * epx + t1 * epy) + t2 * epz);
    return r;
}

// `terms` sees the SH basis values at o (SH model only; the DTI model has
// none).
template <int MODEL, int KT, class Terms = NoTerms>
__device__ __forceinline__ ModelValue model_evaluate(const float *__restrict__ row, int K, V3 o,
                                                     Terms &&terms = Terms{})
{
    if constexpr (MODEL == MODEL_DTI)
        return dti_model(row, o);
    else
        return sh_model<KT>(row, K, o, terms);
}

// ------------------------------------------ Probabilistic, Deterministic ---

__device__ __forceinline__ V3 cross3(V3 a, V3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// The OpenCL normalize builtin, taken as v / length(v).
__device__ __forceinline__ V3 normalize3(V3 v)
{
    const float n = length3(v);
    return v3(v.x / n, v.y / n, v.z / n);
}

// One voxel's model, loaded once for the many evaluations of the
// Probabilistic and Deterministic updates.
template <int MODEL, int KT>
struct VoxelModel {
    float c[MODEL == MODEL_SH ? 45 : 6];

    __device__ __forceinline__ void load(const float *__restrict__ row, int K)
    {
        if constexpr (MODEL == MODEL_SH) {
            load_sh_coefficients<KT>(row, K, c);
        } else {
            const float4 w0 = reinterpret_cast<const float4 *>(row)[0];
            const float4 w1 = reinterpret_cast<const float4 *>(row)[1];
            c[0] = w0.x, c[1] = w0.y, c[2] = w0.z, c[3] = w0.w, c[4] = w1.x, c[5] = w1.y;
        }
    }

    // local_model_evaluate: local.py:91-120 (SH, softplus) or 212-230 (DTI).
    __device__ __forceinline__ ModelValue eval(V3 o) const
    {
        ModelValue r;
        if constexpr (MODEL == MODEL_SH) {
            float az, col, v, th, ph;
            cart2sph(o, az, col);
            sh45_contract(c, az, col, v, th, ph);
            const float d = dsoftplus(v);
            r.value = softplus(v);
            r.dtheta = th * d;
            r.dphi = ph * d;
        } else {
            const float t0 = (c[0] * o.x + c[3] * o.y) + c[4] * o.z;
            const float t1 = (c[3] * o.x + c[1] * o.y) + c[5] * o.z;
            const float t2 = (c[4] * o.x + c[5] * o.y) + c[2] * o.z;
            float az, col, sp, cp, st, ct;
            cart2sph(o, az, col);
            tgk_sincosf(az, &sp, &cp);
            tgk_sincosf(col, &st, &ct);
            r.value = (o.x * t0 + o.y * t1) + o.z * t2;
            r.dtheta = 2.0f * ((t0 * (ct * cp) + t1 * (ct * sp)) + t2 * (-st));
            r.dphi = 2.0f * ((t0 * (-sp) + t1 * cp) + t2 * 0.0f);
        }
        return r;
    }

    // The value alone; the unused derivative sums are dead code here.
    __device__ __forceinline__ float value(V3 o) const { return eval(o).value; }
};

// algorithm.py:305-376 (Probabilistic): n_samples candidates uniform on the
// spherical cap of radius max_angle around the normalised orientation, one
// chosen with probability proportional to its non-negative model value
// (choose_candidate: the reference's two passes, the second replaying only
// the block of candidates where the running weight crosses the target). The
// state after the update is the one left after the chosen candidate's two
// draws, or after all candidates' draws when none is chosen.
template <int MODEL, int KT>
__device__ __forceinline__ V3 probabilistic_update(const AlgoParams &a, const float *__restrict__ row, V3 o_in,
                                                   uint2 &rng)
{
    VoxelModel<MODEL, KT> vox;
    vox.load(row, a.n_parameters);
    const float norm = length3(o_in);
    const V3 o = norm <= 1e-8f ? v3(0.0f, 0.0f, 1.0f) : v3(o_in.x / norm, o_in.y / norm, o_in.z / norm);
    const V3 helper = fabsf(o.z) < 0.9f ? v3(0.0f, 0.0f, 1.0f) : v3(0.0f, 1.0f, 0.0f);
    const V3 t1 = normalize3(cross3(helper, o));
    const V3 t2 = cross3(o, t1);
    const float cos_max = tgk_cosf(a.max_angle);
    auto candidate = [&](uint2 &state) {
        const float cos_angle = 1.0f - randu(state) * (1.0f - cos_max);
        const float sin_angle = tgk_sqrtf(fmaxf(0.0f, 1.0f - cos_angle * cos_angle));
        const float azimuth = (2.0f * 3.14159265359f) * randu(state);
        float sa, ca;
        tgk_sincosf(azimuth, &sa, &ca);
        return v3(cos_angle * o.x + sin_angle * (ca * t1.x + sa * t2.x),
                  cos_angle * o.y + sin_angle * (ca * t1.y + sa * t2.y),
                  cos_angle * o.z + sin_angle * (ca * t1.z + sa * t2.z));
    };
    auto weight = [&](V3 c) {
        const float v = vox.value(c);
        return __builtin_isfinite(v) ? fmaxf(v, 0.0f) : 0.0f;
    };
    V3 chosen;
    return choose_candidate(a.n_samples, rng, candidate, weight, chosen) ? chosen : o;
}

// Gradient of the model on the sphere at o (algorithm.py:500-513): dphi is
// already divided by sin(theta), so it multiplies the unit azimuth tangent.
template <int MODEL, int KT>
__device__ __forceinline__ V3 sphere_gradient(const VoxelModel<MODEL, KT> &vox, V3 o)
{
    const ModelValue m = vox.eval(o);
    float az, col, sp, cp, st, ct;
    cart2sph(o, az, col);
    tgk_sincosf(az, &sp, &cp);
    tgk_sincosf(col, &st, &ct);
    return v3(m.dtheta * (ct * cp) + m.dphi * (-sp), m.dtheta * (ct * sp) + m.dphi * cp,
              m.dtheta * (-st) + m.dphi * 0.0f);
}

// algorithm.py:515-598 (Deterministic): Newton ascent in the tangent frame
// (e_theta, e_phi) with a Hessian from central differences of the gradient
// at geodesic offsets h = 0.01, a fixed 0.1 gradient step where that Hessian
// is not negative definite, the step capped at max_step_size, and a
// backtracking line search (8 halvings) that accepts the first strict
// increase over the current value + 1e-7. Stops on a gradient norm at or
// below the tolerance, a step at or below 1e-7, a rejected line search, or
// after max_iterations.
template <int MODEL, int KT>
__device__ __forceinline__ V3 deterministic_update(const AlgoParams &a, const float *__restrict__ row, V3 o_in)
{
    VoxelModel<MODEL, KT> vox;
    vox.load(row, a.n_parameters);
    const float h = 1e-2f;
    V3 o = normalize3(o_in);
    for (int iteration = 0; iteration < a.max_iterations; iteration++) {
        float az, col, sp, cp, st, ct;
        cart2sph(o, az, col);
        tgk_sincosf(az, &sp, &cp);
        tgk_sincosf(col, &st, &ct);
        const V3 e1 = v3(ct * cp, ct * sp, -st);
        const V3 e2 = v3(-sp, cp, 0.0f);
        const ModelValue current = vox.eval(o);
        const float f0 = current.value, g1 = current.dtheta, g2 = current.dphi;
        if (tgk_sqrtf(g1 * g1 + g2 * g2) <= a.tolerance)
            break;
        const V3 step1 = v3(h * e1.x, h * e1.y, h * e1.z);
        const V3 step2 = v3(h * e2.x, h * e2.y, h * e2.z);
        const V3 gp1 = sphere_gradient(vox, exps2(o, step1));
        const V3 gm1 = sphere_gradient(vox, exps2(o, v3(-step1.x, -step1.y, -step1.z)));
        const V3 gp2 = sphere_gradient(vox, exps2(o, step2));
        const V3 gm2 = sphere_gradient(vox, exps2(o, v3(-step2.x, -step2.y, -step2.z)));
        const V3 dg1 = v3(gp1.x - gm1.x, gp1.y - gm1.y, gp1.z - gm1.z);
        const V3 dg2 = v3(gp2.x - gm2.x, gp2.y - gm2.y, gp2.z - gm2.z);
        const float h11 = dot3(dg1, e1) / (2.0f * h);
        const float h22 = dot3(dg2, e2) / (2.0f * h);
        const float h12a = dot3(dg2, e1) / (2.0f * h);
        const float h12b = dot3(dg1, e2) / (2.0f * h);
        const float h12 = 0.5f * (h12a + h12b);
        const float determinant = h11 * h22 - h12 * h12;
        float d1, d2;
        if (h11 < -1e-7f && h22 < -1e-7f && determinant > 1e-12f) {
            d1 = (-h22 * g1 + h12 * g2) / determinant;
            d2 = (h12 * g1 - h11 * g2) / determinant;
        } else {
            const float gradient_norm = tgk_sqrtf(g1 * g1 + g2 * g2);
            d1 = 0.1f * g1 / gradient_norm;
            d2 = 0.1f * g2 / gradient_norm;
        }
        const float step_norm = tgk_sqrtf(d1 * d1 + d2 * d2);
        if (step_norm > a.max_step_size) {
            const float scale = a.max_step_size / step_norm;
            d1 *= scale;
            d2 *= scale;
        }
        if (tgk_sqrtf(d1 * d1 + d2 * d2) <= 1e-7f)
            break;
        const V3 tangent = v3(d1 * e1.x + d2 * e2.x, d1 * e1.y + d2 * e2.y, d1 * e1.z + d2 * e2.z);
        bool accepted = false;
        float scale = 1.0f;
        for (int search = 0; search < 8; search++) {
            const V3 c = exps2(o, v3(scale * tangent.x, scale * tangent.y, scale * tangent.z));
            if (vox.value(c) > f0 + 1e-7f) {
                accepted = true;
                o = c;
                break;
            }
            scale *= 0.5f;
        }
        if (!accepted)
            break;
    }
    return o;
}

// algorithm.py:83-112 (Transport), 198-228 (Diffusion), 305-376
// (Probabilistic) and 484-617 (Deterministic); for Diffusion the left randn
// of the noise expression is drawn first. For the SH model under Transport
// and Diffusion, `terms(i, Y_i)` receives the basis at cart2sph(o) that the
// model evaluation computes (the other two algorithms never evaluate at o
// alone and ignore it).
template <int MODEL, int ALGO, int KT, class Terms = NoTerms>
__device__ __forceinline__ V3 update(const AlgoParams &a, const float *__restrict__ row, V3 o, uint2 &rng,
                                     Terms &&terms = Terms{})
{
    if constexpr (ALGO == ALGO_PROBABILISTIC)
        return probabilistic_update<MODEL, KT>(a, row, o, rng);
    if constexpr (ALGO == ALGO_DETERMINISTIC)
        return deterministic_update<MODEL, KT>(a, row, o);
    const ModelValue mv = model_evaluate<MODEL, KT>(row, a.n_parameters, o, terms);
    float az, col, sp, cp, st, ct;
    cart2sph(o, az, col);
    tgk_sincosf(az, &sp, &cp);
    tgk_sincosf(col, &st, &ct);
    const float etx = ct * cp, ety = ct * sp, etz = -st;
    const float epx = -sp, epy = cp, epz = 0.0f;
    const float m = fmaxf(mv.value, 1e-6f);
    const float dx = (mv.dtheta * etx + mv.dphi * epx) / m;
    const float dy = (mv.dtheta * ety + mv.dphi * epy) / m;
    const float dz = (mv.dtheta * etz + mv.dphi * epz) / m;
    const float scale_drift = a.gamma * a.step_size;
    if constexpr (ALGO == ALGO_TRANSPORT) {
        return exps2(o, v3(scale_drift * dx, scale_drift * dy, scale_drift * dz));
    } else {
        float nx = 0.0f, ny = 0.0f, nz = 0.0f;
        if (a.noise_variance > 0.0f) {
            const float r1 = randn(rng);
            const float r2 = randn(rng);
            nx = r1 * etx + r2 * epx;
            ny = r1 * ety + r2 * epy;
            nz = r1 * etz + r2 * epz;
        }
        const float scale_noise = tgk_sqrtf(a.noise_variance * a.gamma * a.step_size);
        return exps2(o, v3(scale_drift * dx + scale_noise * nx, scale_drift * dy + scale_noise * ny,
                           scale_drift * dz + scale_noise * nz));
    }
}

}  // namespace plain
}  // namespace tgk
