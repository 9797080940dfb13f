/*
 * tgo_impl.h -- body of the host restatement, instantiated twice by tgo.c.
 *
 * TEST INFRASTRUCTURE ONLY (see tgo.h).
 *
 * Expected macros:
 *   R          real type (float or double)
 *   FN(x)      name mangling, tgo32_##x or tgo64_##x
 *   K(c)       literal of type R written as in the reference (c##f or c)
 *   PI_R       the reference's PI (core.cl:4) in FP32; exact pi in FP64
 *   SQRT SINCOS ATAN2 ACOS EXP LOG FLOOR RINT FMAX FMIN FABS
 *
 * All vector expressions are restated per component, left to right, as the
 * OpenCL C source reads, with no fused multiply-add (compile with
 * -ffp-contract=off). The fourth (w) lane of the reference's float4 values is
 * dropped: it is 1 for points and 0 for directions throughout the path.
 */

typedef struct {
    R x, y, z;
} FN(vec);
#define VEC FN(vec)

static inline VEC FN(mkvec)(R x, R y, R z)
{
    VEC v;
    v.x = x;
    v.y = y;
    v.z = z;
    return v;
}

/* OpenCL length() over xyz. */
static inline R FN(length)(VEC v) { return SQRT((v.x * v.x + v.y * v.y) + v.z * v.z); }

/* core.cl:142-153 */
static inline void FN(cart2sph_one)(VEC c, R *azimuth, R *colatitude)
{
    const R radius = FN(length)(c);
    if (radius <= K(0.0)) {
        *azimuth = K(0.0);
        *colatitude = K(0.0);
        return;
    }
    R az = ATAN2(c.y, c.x);
    if (az < K(0.0))
        az += K(2.0) * PI_R;
    *azimuth = az;
    *colatitude = ACOS(FMIN(FMAX(c.z / radius, K(-1.0)), K(1.0)));
}

/* core.cl:101-109 */
static inline VEC FN(exps2)(VEC p, VEC x, R t)
{
    const R n = FN(length)(x);
    if (n <= K(0.0))
        return p;
    R s, c;
    SINCOS(t * n, &s, &c);
    const R q = s / n;
    return FN(mkvec)(c * p.x + x.x * q, c * p.y + x.y * q, c * p.z + x.z * q);
}

/* core.cl:111-120 with SOFTMAX_SCALE = 100 (local.py:17) */
static inline R FN(softmax)(R x, R k)
{
    if (x * k < K(30.0))
        return LOG(1 + EXP(k * x)) / k;
    return x;
}

static inline R FN(dsoftmax)(R x, R k) { return K(1.0) / (1 + EXP(-k * x)); }

/* core.cl:240-245 */
static inline R FN(randn)(uint32_t state[2])
{
    R n = (R)tgo_randu(state);
    while (n <= K(0.0))
        n = (R)tgo_randu(state);
    return SQRT(K(-2.0) * LOG(n)) * COS_OF((2 * PI_R) * (R)tgo_randu(state));
}

/* spharm.cl:216-413: real symmetric SH basis l <= 8 in MRtrix order, its
 * colatitude derivative and its azimuth derivative divided by sin(theta),
 * restated index by index. */
static void FN(sh45_one)(R phi, R theta, R ylm[45], R dp[45], R dt[45])
{
    R st, ct, sp, cp, s2p, c2p;
    SINCOS(theta, &st, &ct);
    SINCOS(phi, &sp, &cp);
    SINCOS(K(2.0) * phi, &s2p, &c2p);
    const R s3p = 2 * cp * s2p - sp, c3p = 2 * cp * c2p - cp;
    const R s4p = 2 * cp * s3p - s2p, c4p = 2 * cp * c3p - c2p;
    const R s5p = 2 * cp * s4p - s3p, c5p = 2 * cp * c4p - c3p;
    const R s6p = 2 * cp * s5p - s4p, c6p = 2 * cp * c5p - c4p;
    const R s7p = 2 * cp * s6p - s5p, c7p = 2 * cp * c6p - c5p;
    const R s8p = 2 * cp * s7p - s6p, c8p = 2 * cp * c7p - c6p;

    const R st2 = st * st, st3 = st2 * st, st4 = st3 * st, st5 = st4 * st;
    const R st6 = st5 * st, st7 = st6 * st, st8 = st7 * st;
    const R ct2 = ct * ct, ct3 = ct2 * ct, ct4 = ct3 * ct, ct5 = ct4 * ct;
    const R ct6 = ct5 * ct, ct7 = ct6 * ct, ct8 = ct7 * ct;
    const R stct = st * ct;

    ylm[0] = K(0.28209479177387814);
    ylm[1] = K(0.54627421529) * st2 * s2p;
    ylm[2] = K(-1.0925484305920792) * stct * sp;
    ylm[3] = K(0.31539156525252005) * (K(3.0) * ct2 - K(1.0));
    ylm[4] = K(-1.0925484305920792) * stct * cp;
    ylm[5] = K(0.54627421529) * st2 * c2p;
    ylm[6] = K(0.62583573544) * st4 * s4p;
    ylm[7] = K(-1.77013076978) * st3 * ct * s3p;
    ylm[8] = K(0.47308734787) * st2 * (K(7.0) * ct2 - K(1.0)) * s2p;
    ylm[9] = K(-0.66904654355) * st * (K(7.0) * ct2 - K(3.0)) * ct * sp;
    ylm[10] = K(0.10578554691520431) * (K(35.0) * ct2 * ct2 - K(30.0) * ct2 + K(3.0));
    ylm[11] = K(-0.66904654355) * st * (K(7.0) * ct2 - K(3.0)) * ct * cp;
    ylm[12] = K(0.47308734787) * st2 * (K(7.0) * ct2 - K(1.0)) * c2p;
    ylm[13] = K(-1.77013076978) * st3 * ct * c3p;
    ylm[14] = K(0.62583573544) * st4 * c4p;
    ylm[15] = K(0.68318410519) * st6 * s6p;
    ylm[16] = K(-2.36661916223) * st5 * ct * s5p;
    ylm[17] = K(0.50456490072) * st4 * (K(11.0) * ct2 - K(1.0)) * s4p;
    ylm[18] = K(-0.92120525951) * st3 * (K(11.0) * ct3 - K(3.0) * ct) * s3p;
    ylm[19] = K(0.46060262975) * st2 * (K(33.0) * ct4 - K(18.0) * ct2 + K(1.0)) * s2p;
    ylm[20] = K(-0.58262136251) * st * (K(33.0) * ct5 - K(30.0) * ct3 + K(5.0) * ct) * sp;
    ylm[21] = K(0.06356920226) * (K(231.0) * ct6 - K(315.0) * ct4 + K(105.0) * ct2 - K(5.0));
    ylm[22] = K(-0.58262136251) * st * (K(33.0) * ct5 - K(30.0) * ct3 + K(5.0) * ct) * cp;
    ylm[23] = K(0.46060262975) * st2 * (K(33.0) * ct4 - K(18.0) * ct2 + K(1.0)) * c2p;
    ylm[24] = K(-0.92120525951) * st3 * (K(11.0) * ct3 - K(3.0) * ct) * c3p;
    ylm[25] = K(0.50456490072) * st4 * (K(11.0) * ct2 - K(1.0)) * c4p;
    ylm[26] = K(-2.36661916223) * st5 * ct * c5p;
    ylm[27] = K(0.68318410519) * st6 * c6p;
    ylm[28] = K(0.72892666017) * st8 * s8p;
    ylm[29] = K(-2.9157066407) * st7 * ct * s7p;
    ylm[30] = K(0.53233276606) * st6 * (K(15.0) * ct2 - K(1.0)) * s6p;
    ylm[31] = K(-3.4499106221) * st5 * (K(5.0) * ct3 - ct) * s5p;
    ylm[32] = K(0.47841652475) * st4 * (K(65.0) * ct4 - K(26.0) * ct2 + K(1.0)) * s4p;
    ylm[33] = K(-1.2352661553) * st3 * (K(39.0) * ct5 - K(26.0) * ct3 + 3 * ct) * s3p;
    ylm[34] = K(0.45615225843) * st2 * (K(143.0) * ct6 - K(143.0) * ct4 + K(33.0) * ct2 - K(1.0)) * s2p;
    ylm[35] = K(-0.10904124589) * st * (K(715.0) * ct7 - K(1001.0) * ct5 + K(385.0) * ct3 - K(35.0) * ct) * sp;
    ylm[36] = K(0.00908677049) * (K(6435.0) * ct8 - K(12012.0) * ct6 + K(6930.0) * ct4 - K(1260.0) * ct2 + K(35.0));
    ylm[37] = K(-0.10904124589) * st * (K(715.0) * ct7 - K(1001.0) * ct5 + K(385.0) * ct3 - K(35.0) * ct) * cp;
    ylm[38] = K(0.45615225843) * st2 * (K(143.0) * ct6 - K(143.0) * ct4 + K(33.0) * ct2 - K(1.0)) * c2p;
    ylm[39] = K(-1.2352661553) * st3 * (K(39.0) * ct5 - K(26.0) * ct3 + 3 * ct) * c3p;
    ylm[40] = K(0.47841652475) * st4 * (K(65.0) * ct4 - K(26.0) * ct2 + K(1.0)) * c4p;
    ylm[41] = K(-3.4499106221) * st5 * (K(5.0) * ct3 - ct) * c5p;
    ylm[42] = K(0.53233276606) * st6 * (K(15.0) * ct2 - K(1.0)) * c6p;
    ylm[43] = K(-2.9157066407) * st7 * ct * c7p;
    ylm[44] = K(0.72892666017) * st8 * c8p;

    dp[0] = K(0.0);
    dp[1] = K(0.54627421529) * st * K(2.0) * c2p;
    dp[2] = K(-1.0925484305920792) * ct * cp;
    dp[3] = K(0.0);
    dp[4] = K(1.0925484305920792) * ct * sp;
    dp[5] = K(-0.54627421529) * st * K(2.0) * s2p;
    dp[6] = K(0.62583573544) * st3 * K(4.0) * c4p;
    dp[7] = K(-1.77013076978) * st2 * ct * K(3.0) * c3p;
    dp[8] = K(0.47308734787) * st * (K(7.0) * ct2 - K(1.0)) * K(2.0) * c2p;
    dp[9] = K(-0.66904654355) * ct * (K(7.0) * ct2 - K(3.0)) * cp;
    dp[10] = K(0.0);
    dp[11] = K(0.66904654355) * ct * (K(7.0) * ct2 - K(3.0)) * sp;
    dp[12] = K(-0.47308734787) * st * (K(7.0) * ct2 - K(1.0)) * K(2.0) * s2p;
    dp[13] = K(1.77013076978) * st2 * ct * K(3.0) * s3p;
    dp[14] = K(-0.62583573544) * st3 * K(4.0) * s4p;
    dp[15] = K(0.68318410519) * st5 * K(6.0) * c6p;
    dp[16] = K(-2.36661916223) * st4 * ct * K(5.0) * c5p;
    dp[17] = K(0.50456490072) * st3 * (K(11.0) * ct2 - K(1.0)) * K(4.0) * c4p;
    dp[18] = K(-0.92120525951) * st2 * (K(11.0) * ct3 - K(3.0) * ct) * K(3.0) * c3p;
    dp[19] = K(0.46060262975) * st * (K(33.0) * ct4 - K(18.0) * ct2 + K(1.0)) * K(2.0) * c2p;
    dp[20] = K(-0.58262136251) * (K(33.0) * ct5 - K(30.0) * ct3 + K(5.0) * ct) * cp;
    dp[21] = K(0.0);
    dp[22] = K(0.58262136251) * (K(33.0) * ct5 - K(30.0) * ct3 + K(5.0) * ct) * sp;
    dp[23] = K(-0.46060262975) * st * (K(33.0) * ct4 - K(18.0) * ct2 + K(1.0)) * K(2.0) * s2p;
    dp[24] = K(0.92120525951) * st2 * (K(11.0) * ct3 - K(3.0) * ct) * K(3.0) * s3p;
    dp[25] = K(-0.50456490072) * st3 * (K(11.0) * ct2 - K(1.0)) * K(4.0) * s4p;
    dp[26] = K(2.36661916223) * st4 * ct * K(5.0) * s5p;
    dp[27] = K(-0.68318410519) * st5 * K(6.0) * s6p;
    dp[28] = K(0.72892666017) * st7 * K(8.0) * c8p;
    dp[29] = K(-2.9157066407) * st6 * ct * K(7.0) * c7p;
    dp[30] = K(0.53233276606) * st5 * (K(15.0) * ct2 - K(1.0)) * K(6.0) * c6p;
    dp[31] = K(-3.4499106221) * st4 * (K(5.0) * ct3 - ct) * K(5.0) * c5p;
    dp[32] = K(0.47841652475) * st3 * (K(65.0) * ct4 - K(26.0) * ct2 + K(1.0)) * K(4.0) * c4p;
    dp[33] = K(-1.2352661553) * st2 * (K(39.0) * ct5 - K(26.0) * ct3 + 3 * ct) * K(3.0) * c3p;
    dp[34] = K(0.45615225843) * st * (K(143.0) * ct6 - K(143.0) * ct4 + K(33.0) * ct2 - K(1.0)) * K(2.0) * c2p;
    dp[35] = K(-0.10904124589) * (K(715.0) * ct7 - K(1001.0) * ct5 + K(385.0) * ct3 - K(35.0) * ct) * cp;
    dp[36] = K(0.0);
    dp[37] = K(0.10904124589) * (K(715.0) * ct7 - K(1001.0) * ct5 + K(385.0) * ct3 - K(35.0) * ct) * sp;
    dp[38] = K(-0.45615225843) * st * (K(143.0) * ct6 - K(143.0) * ct4 + K(33.0) * ct2 - K(1.0)) * K(2.0) * s2p;
    dp[39] = K(1.2352661553) * st2 * (K(39.0) * ct5 - K(26.0) * ct3 + 3 * ct) * K(3.0) * s3p;
    dp[40] = K(-0.47841652475) * st3 * (K(65.0) * ct4 - K(26.0) * ct2 + K(1.0)) * K(4.0) * s4p;
    dp[41] = K(3.4499106221) * st4 * (K(5.0) * ct3 - ct) * K(5.0) * s5p;
    dp[42] = K(-0.53233276606) * st5 * (K(15.0) * ct2 - K(1.0)) * K(6.0) * s6p;
    dp[43] = K(2.9157066407) * st6 * ct * K(7.0) * s7p;
    dp[44] = K(-0.72892666017) * st7 * K(8.0) * s8p;

    dt[0] = K(0.0);
    dt[1] = K(0.54627421529) * K(2.0) * stct * s2p;
    dt[2] = K(-1.0925484305920792) * (ct2 - st2) * sp;
    dt[3] = K(-0.31539156525252005) * K(6.0) * ct * st;
    dt[4] = K(-1.0925484305920792) * (ct2 - st2) * cp;
    dt[5] = K(0.54627421529) * K(2.0) * stct * c2p;
    dt[6] = K(0.62583573544) * K(4.0) * st3 * ct * s4p;
    dt[7] = K(-1.77013076978) * (K(3.0) * st2 * ct2 - st4) * s3p;
    dt[8] = K(0.47308734787) * (K(2.0) * stct * (K(7.0) * ct2 - K(1.0)) - K(14.0) * st3 * ct) * s2p;
    dt[9] = K(-0.66904654355) * ((ct2 - st2) * (K(7.0) * ct2 - K(3.0)) - (K(14.0) * st2 * ct2)) * sp;
    dt[10] = K(0.10578554691520431) * (K(-35.0) * K(4.0) * ct2 * stct + K(30.0) * K(2.0) * stct);
    dt[11] = K(-0.66904654355) * ((ct2 - st2) * (K(7.0) * ct2 - K(3.0)) - (K(14.0) * st2 * ct2)) * cp;
    dt[12] = K(0.47308734787) * (K(2.0) * stct * (K(7.0) * ct2 - K(1.0)) - K(14.0) * st3 * ct) * c2p;
    dt[13] = K(-1.77013076978) * (K(3.0) * st2 * ct2 - st4) * c3p;
    dt[14] = K(0.62583573544) * K(4.0) * st3 * ct * c4p;
    dt[15] = K(0.68318410519) * K(6.0) * st5 * ct * s6p;
    dt[16] = K(-2.36661916223) * (K(5.0) * st4 * ct2 - st6) * s5p;
    dt[17] = K(0.50456490072) * (K(4.0) * st3 * ct * (K(11.0) * ct2 - K(1.0)) - K(22.0) * st5 * ct) * s4p;
    dt[18] = K(-0.92120525951) * (K(3.0) * st2 * ct * (K(11.0) * ct3 - K(3.0) * ct) - st4 * (K(33.0) * ct2 - K(3.0))) * s3p;
    dt[19] = K(0.46060262975) * (K(2.0) * st * ct * (K(33.0) * ct4 - K(18.0) * ct2 + K(1.0)) - st3 * (K(33.0) * K(4.0) * ct3 - K(36.0) * ct)) * s2p;
    dt[20] = K(-0.58262136251) * (ct * (K(33.0) * ct5 - K(30.0) * ct3 + K(5.0) * ct) - st2 * (K(33.0) * K(5.0) * ct4 - K(90.0) * ct2 + K(5.0))) * sp;
    dt[21] = K(-0.06356920226) * (K(1386.0) * ct5 - K(1260.0) * ct3 + K(210.0) * ct) * st;
    dt[22] = K(-0.58262136251) * (ct * (K(33.0) * ct5 - K(30.0) * ct3 + K(5.0) * ct) - st2 * (K(33.0) * K(5.0) * ct4 - K(90.0) * ct2 + K(5.0))) * cp;
    dt[23] = K(0.46060262975) * (K(2.0) * st * ct * (K(33.0) * ct4 - K(18.0) * ct2 + K(1.0)) - st3 * (K(33.0) * K(4.0) * ct3 - K(36.0) * ct)) * c2p;
    dt[24] = K(-0.92120525951) * (K(3.0) * st2 * ct * (K(11.0) * ct3 - K(3.0) * ct) - st4 * (K(33.0) * ct2 - K(3.0))) * c3p;
    dt[25] = K(0.50456490072) * (K(4.0) * st3 * ct * (K(11.0) * ct2 - K(1.0)) - K(22.0) * st5 * ct) * c4p;
    dt[26] = K(-2.36661916223) * (K(5.0) * st4 * ct2 - st6) * c5p;
    dt[27] = K(0.68318410519) * K(6.0) * st5 * ct * c6p;
    dt[28] = K(0.72892666017) * K(8.0) * st7 * ct * s8p;
    dt[29] = K(-2.9157066407) * (K(7.0) * st6 * ct2 - st8) * s7p;
    dt[30] = K(0.53233276606) * (K(6.0) * st5 * ct * (K(15.0) * ct2 - K(1.0)) - st7 * K(30.0) * ct) * s6p;
    dt[31] = K(-3.4499106221) * (K(5.0) * st4 * ct * (K(5.0) * ct3 - ct) - st6 * (K(15.0) * ct2 - K(1.0))) * s5p;
    dt[32] = K(0.47841652475) * (K(4.0) * st3 * ct * (K(65.0) * ct4 - K(26.0) * ct2 + K(1.0)) - st5 * (K(260.0) * ct3 - K(52.0) * ct)) * s4p;
    dt[33] = K(-1.2352661553) * (K(3.0) * st2 * ct * (K(39.0) * ct5 - K(26.0) * ct3 + 3 * ct) - st4 * (K(195.0) * ct4 - K(78.0) * ct2 + K(3.0))) * s3p;
    dt[34] = K(0.45615225843) * (K(2.0) * st * ct * (K(143.0) * ct6 - K(143.0) * ct4 + K(33.0) * ct2 - K(1.0)) - st3 * (K(858.0) * ct5 - K(572.0) * ct3 + K(66.0) * ct)) * s2p;
    dt[35] = K(-0.10904124589) * (ct * (K(715.0) * ct7 - K(1001.0) * ct5 + K(385.0) * ct3 - K(35.0) * ct) - st2 * (K(5005.0) * ct6 - K(5005.0) * ct4 + K(1155.0) * ct2 - K(35.0))) * sp;
    dt[36] = K(-0.00908677049) * (K(8.0) * K(6435.0) * ct7 - K(6.0) * K(12012.0) * ct5 + K(4.0) * K(6930.0) * ct3 - K(2.0) * K(1260.0) * ct) * st;
    dt[37] = K(-0.10904124589) * (ct * (K(715.0) * ct7 - K(1001.0) * ct5 + K(385.0) * ct3 - K(35.0) * ct) - st2 * (K(5005.0) * ct6 - K(5005.0) * ct4 + K(1155.0) * ct2 - K(35.0))) * cp;
    dt[38] = K(0.45615225843) * (K(2.0) * st * ct * (K(143.0) * ct6 - K(143.0) * ct4 + K(33.0) * ct2 - K(1.0)) - st3 * (K(858.0) * ct5 - K(572.0) * ct3 + K(66.0) * ct)) * c2p;
    dt[39] = K(-1.2352661553) * (K(3.0) * st2 * ct * (K(39.0) * ct5 - K(26.0) * ct3 + 3 * ct) - st4 * (K(195.0) * ct4 - K(78.0) * ct2 + K(3.0))) * c3p;
    dt[40] = K(0.47841652475) * (K(4.0) * st3 * ct * (K(65.0) * ct4 - K(26.0) * ct2 + K(1.0)) - st5 * (K(260.0) * ct3 - K(52.0) * ct)) * c4p;
    dt[41] = K(-3.4499106221) * (K(5.0) * st4 * ct * (K(5.0) * ct3 - ct) - st6 * (K(15.0) * ct2 - K(1.0))) * c5p;
    dt[42] = K(0.53233276606) * (K(6.0) * st5 * ct * (K(15.0) * ct2 - K(1.0)) - st7 * K(30.0) * ct) * c6p;
    dt[43] = K(-2.9157066407) * (K(7.0) * st6 * ct2 - st8) * c7p;
    dt[44] = K(0.72892666017) * K(8.0) * st7 * ct * c8p;
}

typedef struct {
    R value, dtheta, dphi;
} FN(model_value);
#define MV FN(model_value)

/* local.py:91-120 (SH, K <= 45; the FP64 instance also takes even orders
 * past 8, up to the generic basis's limit) */
static inline MV FN(sh_model)(const R *model, int n_parameters, VEC o)
{
#ifdef TGO_GENERIC_SH
    R ylm[(TGO_LMAX_GENERIC + 1) * (TGO_LMAX_GENERIC + 2) / 2], dp[sizeof ylm / sizeof ylm[0]],
        dt[sizeof ylm / sizeof ylm[0]];
#else
    R ylm[45], dp[45], dt[45];
#endif
    R az, col;
    FN(cart2sph_one)(o, &az, &col);
#ifdef TGO_GENERIC_SH
    if (n_parameters > 45) {
        int lmax = 0;
        while ((lmax + 1) * (lmax + 2) / 2 < n_parameters)
            lmax += 2;
        tgo_sh_generic64(lmax, az, col, ylm, dp, dt);
    } else
#endif
        FN(sh45_one)(az, col, ylm, dp, dt);
    MV r;
    r.value = K(0.0);
    r.dtheta = K(0.0);
    r.dphi = K(0.0);
    for (int i = 0; i < n_parameters; i++) {
        const R c = model[i];
        r.value += c * ylm[i];
        r.dtheta += c * dt[i];
        r.dphi += c * dp[i];
    }
    const R d = FN(dsoftmax)(r.value, K(100.0));
    r.value = FN(softmax)(r.value, K(100.0));
    r.dtheta = r.dtheta * d;
    r.dphi = r.dphi * d;
    return r;
}

/* local.py:212-230 (DTI, coefficients Dxx Dyy Dzz Dxy Dxz Dyz) */
static inline MV FN(dti_model)(const R *m, VEC o)
{
    const R t0 = (m[0] * o.x + m[3] * o.y) + m[4] * o.z;
    const R t1 = (m[3] * o.x + m[1] * o.y) + m[5] * o.z;
    const R t2 = (m[4] * o.x + m[5] * o.y) + m[2] * o.z;
    R az, col, sp, cp, st, ct;
    FN(cart2sph_one)(o, &az, &col);
    SINCOS(az, &sp, &cp);
    SINCOS(col, &st, &ct);
    const R etx = ct * cp, ety = ct * sp, etz = -st;
    const R epx = -sp, epy = cp, epz = K(0.0);
    MV r;
    r.value = (o.x * t0 + o.y * t1) + o.z * t2;
    r.dtheta = K(2.0) * ((t0 * etx + t1 * ety) + t2 * etz);
    r.dphi = K(2.0) * ((t0 * epx + t1 * epy) + t2 * epz);
    return r;
}

static inline MV FN(model_evaluate)(int model, int n_parameters, const R *params, VEC o)
{
    if (model == TGO_MODEL_DTI)
        return FN(dti_model)(params, o);
    return FN(sh_model)(params, n_parameters, o);
}

/* algorithm.py:83-112 (Transport) and 198-228 (Diffusion). */
static inline VEC FN(cross)(VEC a, VEC b)
{
    return FN(mkvec)(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline R FN(dot)(VEC a, VEC b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
/* The OpenCL normalize builtin, taken as v / length(v). */
static inline VEC FN(normalize)(VEC v)
{
    const R n = FN(length)(v);
    return FN(mkvec)(v.x / n, v.y / n, v.z / n);
}

/* algorithm.py:340-343: a direction uniform on the cap of cos >= cos_max
 * around o (two uniform draws: the cosine, then the azimuth). */
static inline VEC FN(cap_candidate)(VEC o, VEC t1, VEC t2, R cos_max, uint32_t rng[2])
{
    const R cos_angle = K(1.0) - (R)tgo_randu(rng) * (K(1.0) - cos_max);
    const R sin_angle = SQRT(FMAX(K(0.0), K(1.0) - cos_angle * cos_angle));
    const R azimuth = (2 * PI_R) * (R)tgo_randu(rng);
    R sa, ca;
    SINCOS(azimuth, &sa, &ca);
    return FN(mkvec)(cos_angle * o.x + sin_angle * (ca * t1.x + sa * t2.x),
                     cos_angle * o.y + sin_angle * (ca * t1.y + sa * t2.y),
                     cos_angle * o.z + sin_angle * (ca * t1.z + sa * t2.z));
}

/* algorithm.py:315-376 (Probabilistic): weights max(value, 0) (0 if not
 * finite) of n_samples cap candidates, target = randu * total, then the
 * noise state replayed and the first candidate whose running weight exceeds
 * the target returned; the normalised input when the total is not positive
 * or no candidate exceeds the target. */
static VEC FN(probabilistic_one)(const tgo_algorithm *a, const R *params, VEC o_in, uint32_t rng[2])
{
    const R norm = FN(length)(o_in);
    const VEC o = norm <= K(1e-8) ? FN(mkvec)(K(0.0), K(0.0), K(1.0))
                                  : FN(mkvec)(o_in.x / norm, o_in.y / norm, o_in.z / norm);
    const VEC helper = FABS(o.z) < K(0.9) ? FN(mkvec)(K(0.0), K(0.0), K(1.0)) : FN(mkvec)(K(0.0), K(1.0), K(0.0));
    const VEC t1 = FN(normalize)(FN(cross)(helper, o));
    const VEC t2 = FN(cross)(o, t1);
    const R cos_max = COS_OF((R)a->max_angle);
    const uint32_t initial[2] = {rng[0], rng[1]};
    R total = K(0.0);
    for (int i = 0; i < a->n_samples; i++) {
        const R v = FN(model_evaluate)(a->model, a->n_parameters, params,
                                       FN(cap_candidate)(o, t1, t2, cos_max, rng)).value;
        total += isfinite(v) ? FMAX(v, K(0.0)) : K(0.0);
    }
    if (total <= K(0.0))
        return o;
    const R target = (R)tgo_randu(rng) * total;
    rng[0] = initial[0];
    rng[1] = initial[1];
    R cumulative = K(0.0);
    for (int i = 0; i < a->n_samples; i++) {
        const VEC c = FN(cap_candidate)(o, t1, t2, cos_max, rng);
        const R v = FN(model_evaluate)(a->model, a->n_parameters, params, c).value;
        cumulative += isfinite(v) ? FMAX(v, K(0.0)) : K(0.0);
        if (cumulative > target)
            return c;
    }
    return o;
}

/* algorithm.py:500-513: model gradient on the sphere (dphi already carries
 * the 1 / sin(theta)). */
static inline VEC FN(sphere_gradient)(const tgo_algorithm *a, const R *params, VEC o)
{
    const MV m = FN(model_evaluate)(a->model, a->n_parameters, params, o);
    R az, col, sp, cp, st, ct;
    FN(cart2sph_one)(o, &az, &col);
    SINCOS(az, &sp, &cp);
    SINCOS(col, &st, &ct);
    return FN(mkvec)(m.dtheta * (ct * cp) + m.dphi * (-sp), m.dtheta * (ct * sp) + m.dphi * cp,
                     m.dtheta * (-st) + m.dphi * K(0.0));
}

/* algorithm.py:535-598 (Deterministic): Newton ascent with a central-
 * difference Hessian (h = 0.01), 0.1 gradient steps off negative-definite
 * regions, steps capped at max_step_size, 8-halving line search on
 * value > f0 + 1e-7. */
static VEC FN(deterministic_one)(const tgo_algorithm *a, const R *params, VEC o_in)
{
    const R h = K(1e-2);
    VEC o = FN(normalize)(o_in);
    for (int iteration = 0; iteration < a->max_iterations; iteration++) {
        R az, col, sp, cp, st, ct;
        FN(cart2sph_one)(o, &az, &col);
        SINCOS(az, &sp, &cp);
        SINCOS(col, &st, &ct);
        const VEC e1 = FN(mkvec)(ct * cp, ct * sp, -st);
        const VEC e2 = FN(mkvec)(-sp, cp, K(0.0));
        const MV current = FN(model_evaluate)(a->model, a->n_parameters, params, o);
        const R f0 = current.value, g1 = current.dtheta, g2 = current.dphi;
        if (SQRT(g1 * g1 + g2 * g2) <= (R)a->tolerance)
            break;
        const VEC step1 = FN(mkvec)(h * e1.x, h * e1.y, h * e1.z);
        const VEC step2 = FN(mkvec)(h * e2.x, h * e2.y, h * e2.z);
        const VEC gp1 = FN(sphere_gradient)(a, params, FN(exps2)(o, step1, K(1.0)));
        const VEC gm1 = FN(sphere_gradient)(a, params, FN(exps2)(o, FN(mkvec)(-step1.x, -step1.y, -step1.z), K(1.0)));
        const VEC gp2 = FN(sphere_gradient)(a, params, FN(exps2)(o, step2, K(1.0)));
        const VEC gm2 = FN(sphere_gradient)(a, params, FN(exps2)(o, FN(mkvec)(-step2.x, -step2.y, -step2.z), K(1.0)));
        const VEC dg1 = FN(mkvec)(gp1.x - gm1.x, gp1.y - gm1.y, gp1.z - gm1.z);
        const VEC dg2 = FN(mkvec)(gp2.x - gm2.x, gp2.y - gm2.y, gp2.z - gm2.z);
        const R h11 = FN(dot)(dg1, e1) / (K(2.0) * h);
        const R h22 = FN(dot)(dg2, e2) / (K(2.0) * h);
        const R h12a = FN(dot)(dg2, e1) / (K(2.0) * h);
        const R h12b = FN(dot)(dg1, e2) / (K(2.0) * h);
        const R h12 = K(0.5) * (h12a + h12b);
        const R determinant = h11 * h22 - h12 * h12;
        R d1, d2;
        if (h11 < K(-1e-7) && h22 < K(-1e-7) && determinant > K(1e-12)) {
            d1 = (-h22 * g1 + h12 * g2) / determinant;
            d2 = (h12 * g1 - h11 * g2) / determinant;
        } else {
            const R gradient_norm = SQRT(g1 * g1 + g2 * g2);
            d1 = K(0.1) * g1 / gradient_norm;
            d2 = K(0.1) * g2 / gradient_norm;
        }
        const R step_norm = SQRT(d1 * d1 + d2 * d2);
        if (step_norm > (R)a->max_step_size) {
            const R scale = (R)a->max_step_size / step_norm;
            d1 *= scale;
            d2 *= scale;
        }
        if (SQRT(d1 * d1 + d2 * d2) <= K(1e-7))
            break;
        const VEC tangent = FN(mkvec)(d1 * e1.x + d2 * e2.x, d1 * e1.y + d2 * e2.y, d1 * e1.z + d2 * e2.z);
        int accepted = 0;
        R scale = K(1.0);
        for (int search = 0; search < 8; search++) {
            const VEC c = FN(exps2)(o, FN(mkvec)(scale * tangent.x, scale * tangent.y, scale * tangent.z), K(1.0));
            if (FN(model_evaluate)(a->model, a->n_parameters, params, c).value > f0 + K(1e-7)) {
                accepted = 1;
                o = c;
                break;
            }
            scale *= K(0.5);
        }
        if (!accepted)
            break;
    }
    return o;
}

static inline VEC FN(update_one)(const tgo_algorithm *a, const R *params, VEC o, uint32_t rng[2])
{
    if (a->algorithm == TGO_ALGO_PROBABILISTIC)
        return FN(probabilistic_one)(a, params, o, rng);
    if (a->algorithm == TGO_ALGO_DETERMINISTIC)
        return FN(deterministic_one)(a, params, o);
    const MV mv = FN(model_evaluate)(a->model, a->n_parameters, params, o);
    R az, col, sp, cp, st, ct;
    FN(cart2sph_one)(o, &az, &col);
    SINCOS(az, &sp, &cp);
    SINCOS(col, &st, &ct);
    const R etx = ct * cp, ety = ct * sp, etz = -st;
    const R epx = -sp, epy = cp, epz = K(0.0);
    const R m = FMAX(mv.value, K(1e-6));
    const R dx = (mv.dtheta * etx + mv.dphi * epx) / m;
    const R dy = (mv.dtheta * ety + mv.dphi * epy) / m;
    const R dz = (mv.dtheta * etz + mv.dphi * epz) / m;
    const R step = (R)a->step_size;
    const R gamma = (R)a->gamma;
    const R scale_drift = gamma * step;
    if (a->algorithm == TGO_ALGO_TRANSPORT) {
        const VEC tangent = FN(mkvec)(scale_drift * dx, scale_drift * dy, scale_drift * dz);
        return FN(exps2)(o, tangent, K(1.0));
    }
    const R nv = (R)a->noise_variance;
    R nx = K(0.0), ny = K(0.0), nz = K(0.0);
    if (nv > K(0.0)) {
        const R r1 = FN(randn)(rng); /* left operand first, as clang orders it */
        const R r2 = FN(randn)(rng);
        nx = r1 * etx + r2 * epx;
        ny = r1 * ety + r2 * epy;
        nz = r1 * etz + r2 * epz;
    }
    const R scale_noise = SQRT(nv * gamma * step);
    const VEC tangent = FN(mkvec)(scale_drift * dx + scale_noise * nx, scale_drift * dy + scale_noise * ny,
                                  scale_drift * dz + scale_noise * nz);
    return FN(exps2)(o, tangent, K(1.0));
}

/* World to voxel: row dots of the float32 inverse affine with (p, 1)
 * (field.py:91-96, domain.py:361-366). */
static inline void FN(to_voxel)(const float A[16], VEC p, R v[3])
{
    for (int i = 0; i < 3; i++)
        v[i] = (((R)A[4 * i] * p.x + (R)A[4 * i + 1] * p.y) + (R)A[4 * i + 2] * p.z) + (R)A[4 * i + 3];
}

/* field.py:88-114: nearest voxel by rint (ties to even); false outside. */
static inline int FN(field_index)(const tgo_scene *s, VEC p, int64_t *base, R *margin)
{
    R v[3];
    FN(to_voxel)(s->field_inverse_affine, p, v);
    R r[3];
    for (int i = 0; i < 3; i++) {
        r[i] = RINT(v[i]);
        if (margin) {
            const R d = K(0.5) - fabs((double)(v[i] - r[i]));
            if (d < *margin)
                *margin = d;
        }
    }
    for (int i = 0; i < 3; i++)
        if (!(r[i] >= K(0.0) && r[i] < (R)s->field_dims[i]))
            return 0;
    *base = (((int64_t)r[0] * s->field_dims[1] + (int64_t)r[1]) * s->field_dims[2] + (int64_t)r[2]) *
            (int64_t)s->field_dims[3];
    return 1;
}

/* domain.py:358-382: half-open cells via floor(v + 0.5). */
static inline int FN(mask_contains)(const tgo_scene *s, VEC p, R *margin)
{
    R v[3];
    FN(to_voxel)(s->mask_inverse_affine, p, v);
    R r[3];
    for (int i = 0; i < 3; i++) {
        r[i] = FLOOR(v[i] + K(0.5));
        if (margin) {
            const R d = K(0.5) - fabs((double)(v[i] - r[i]));
            if (d < *margin)
                *margin = d;
        }
    }
    for (int i = 0; i < 3; i++)
        if (!(r[i] >= K(0.0) && r[i] < (R)s->mask_dims[i]))
            return 0;
    const int64_t idx = (int64_t)r[0] * ((int64_t)s->mask_dims[1] * s->mask_dims[2]) +
                        (int64_t)r[1] * s->mask_dims[2] + (int64_t)r[2];
    return s->mask[idx] != 0;
}

/* bvh.cl:16-46 */
static inline int FN(bvh_bounds)(const tgo_scene *s, uint32_t node, const R start[3], const R dir[3], R max_t)
{
    R t_min = K(0.0), t_max = max_t;
    for (int axis = 0; axis < 3; axis++) {
        const R lo = (R)s->bvh_min[4 * node + axis];
        const R hi = (R)s->bvh_max[4 * node + axis];
        if (FABS(dir[axis]) < K(1e-12)) {
            if (start[axis] < lo || start[axis] > hi)
                return 0;
        } else {
            const R inv = K(1.0) / dir[axis];
            R near_t = (lo - start[axis]) * inv;
            R far_t = (hi - start[axis]) * inv;
            if (near_t > far_t) {
                const R tmp = near_t;
                near_t = far_t;
                far_t = tmp;
            }
            t_min = FMAX(t_min, near_t);
            t_max = FMIN(t_max, far_t);
            if (t_min > t_max)
                return 0;
        }
    }
    return 1;
}

/* bvh.cl:48-80 (Moller-Trumbore) */
static inline int FN(bvh_triangle)(const tgo_scene *s, uint32_t tri, const R start[3], const R dir[3], R max_t,
                                   R *hit_t, R *hit_u, R *hit_v)
{
    const uint32_t *ix = s->triangles + 4 * (size_t)tri;
    R a[3], b[3], c[3];
    for (int i = 0; i < 3; i++) {
        a[i] = (R)s->vertices[4 * (size_t)ix[0] + i];
        b[i] = (R)s->vertices[4 * (size_t)ix[1] + i];
        c[i] = (R)s->vertices[4 * (size_t)ix[2] + i];
    }
    const R e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const R e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const R p[3] = {dir[1] * e2[2] - dir[2] * e2[1], dir[2] * e2[0] - dir[0] * e2[2], dir[0] * e2[1] - dir[1] * e2[0]};
    const R det = (e1[0] * p[0] + e1[1] * p[1]) + e1[2] * p[2];
    if (FABS(det) < K(1e-10))
        return 0;
    const R inv = K(1.0) / det;
    const R off[3] = {start[0] - a[0], start[1] - a[1], start[2] - a[2]};
    const R u = ((off[0] * p[0] + off[1] * p[1]) + off[2] * p[2]) * inv;
    if (u < K(-1e-6) || u > K(1.0) + K(1e-6))
        return 0;
    const R q[3] = {off[1] * e1[2] - off[2] * e1[1], off[2] * e1[0] - off[0] * e1[2], off[0] * e1[1] - off[1] * e1[0]};
    const R v = ((dir[0] * q[0] + dir[1] * q[1]) + dir[2] * q[2]) * inv;
    if (v < K(-1e-6) || u + v > K(1.0) + K(1e-6))
        return 0;
    const R t = ((e2[0] * q[0] + e2[1] * q[1]) + e2[2] * q[2]) * inv;
    if (t <= K(1e-6) || t > max_t)
        return 0;
    *hit_t = t;
    *hit_u = u;
    *hit_v = v;
    return 1;
}

/* bvh.cl:83-131: closest hit with t in (1e-6, 1]; a later-visited triangle
 * at an equal t replaces the earlier one; left child visited first. */
/* bvh.cl:83-131; bary receives (1 - u - v, u, v) of the closest hit. */
static inline int FN(bvh_hit_bary)(const tgo_scene *s, VEC start, VEC end, uint32_t *hit, R *hit_t, R bary[3])
{
    const R st[3] = {start.x, start.y, start.z};
    const R dir[3] = {end.x - start.x, end.y - start.y, end.z - start.z};
    int32_t stack[64];
    uint32_t size = 1;
    stack[0] = 0;
    *hit_t = K(1.0);
    *hit = 0;
    int found = 0;
    while (size > 0) {
        const uint32_t node = (uint32_t)stack[--size];
        if (!FN(bvh_bounds)(s, node, st, dir, *hit_t))
            continue;
        const int32_t *nd = s->bvh_nodes + 4 * (size_t)node;
        if (nd[3] > 0) {
            for (int i = 0; i < nd[3]; i++) {
                const uint32_t tri = s->bvh_order[nd[2] + i];
                R t, u, v;
                if (FN(bvh_triangle)(s, tri, st, dir, *hit_t, &t, &u, &v)) {
                    *hit_t = t;
                    *hit = tri;
                    bary[0] = K(1.0) - u - v;
                    bary[1] = u;
                    bary[2] = v;
                    found = 1;
                }
            }
        } else {
            if (size + 2 > 64)
                return found;
            stack[size++] = nd[1];
            stack[size++] = nd[0];
        }
    }
    return found;
}

static inline int FN(bvh_hit)(const tgo_scene *s, VEC start, VEC end, uint32_t *hit, R *hit_t)
{
    R bary[3];
    return FN(bvh_hit_bary)(s, start, end, hit, hit_t, bary);
}

/* tracking.py:112-148 for one seed. Mesh domains: the loop's domain test is
 * "the segment just traversed crossed no boundary triangle", checked at the
 * top of each iteration after the first. */
static void FN(track_one)(const tgo_scene *s, const tgo_algorithm *a, const float *seed, uint32_t rng_seed,
                          uint32_t stream_id, uint32_t n_points, float save_at, uint32_t *length_out,
                          uint32_t *steps_out, uint32_t *hit_out, R *points, R *margins)
{
    VEC point = FN(mkvec)((R)seed[0], (R)seed[1], (R)seed[2]);
    VEC orientation = FN(mkvec)((R)seed[4], (R)seed[5], (R)seed[6]);
    VEC previous = point;
    uint32_t rng[2];
    tgo_rng_init(rng_seed, stream_id, rng);
    const int K_ = a->n_parameters;
    R params[160];
    const R step = (R)a->step_size;
    const R save = (R)save_at;
    R margin = K(0.5);

    points[0] = point.x;
    points[1] = point.y;
    points[2] = point.z;
    if (margins)
        margins[0] = margin;
    uint32_t n = 1, steps = 0, hit = 0xffffffffu;
    R time = K(0.0);
    int first = 1;
    while (n < n_points) {
        if (s->domain_kind == TGO_DOMAIN_MASK) {
            if (!FN(mask_contains)(s, point, margins ? &margin : 0))
                break;
        } else if (!first) {
            uint32_t tri;
            R t;
            if (FN(bvh_hit)(s, previous, point, &tri, &t)) {
                hit = tri;
                break;
            }
        }
        int64_t base;
        if (!FN(field_index)(s, point, &base, margins ? &margin : 0))
            break;
        for (int i = 0; i < K_; i++)
            params[i] = (R)s->field[base + i];
        orientation = FN(update_one)(a, params, orientation, rng);
        steps++;
        previous = point;
        point.x += step * orientation.x;
        point.y += step * orientation.y;
        point.z += step * orientation.z;
        first = 0;
        time += step;
        if (time >= save) {
            time -= save;
            points[3 * n + 0] = point.x - time * orientation.x;
            points[3 * n + 1] = point.y - time * orientation.y;
            points[3 * n + 2] = point.z - time * orientation.z;
            if (margins)
                margins[n] = margin;
            n++;
        }
    }
    *length_out = n;
    *steps_out = steps;
    *hit_out = hit;
}

void FN(track)(const tgo_scene *s, const tgo_algorithm *a, const float *seeds, int64_t n, const uint32_t *seed_rng,
               const uint32_t *seed_stream, uint32_t n_points, float save_at, uint32_t *lengths, uint32_t *steps,
               uint32_t *hits, R *points TRACK_EXTRA_PARAM, int n_threads)
{
#pragma omp parallel for schedule(dynamic, 16) num_threads(n_threads > 0 ? n_threads : 1)
    for (int64_t g = 0; g < n; g++)
        FN(track_one)(s, a, seeds + 8 * g, seed_rng[g], seed_stream[g], n_points, save_at, lengths + g, steps + g,
                      hits + g, points + (size_t)3 * n_points * (size_t)g, TRACK_MARGINS(g));
}

/* tracking.py:457-558 (kernel connectome) for seeds 0..n-1: the boundary
 * sample of rng_init(rng_seed, g), whose state then drives the noise; mask,
 * field, update, orientation check (finite, length >= 0.5), move; then the
 * last segment against the boundary BVH. rows/cols/weights: 9 per seed, row
 * -1 when the track contributes nothing. */
void FN(connectome)(const tgo_scene *s, const tgo_algorithm *a, const float *bvertices, const uint32_t *btriangles,
                    const float *bcumulative, uint32_t bn, uint32_t rng_seed, int64_t n, uint32_t n_steps,
                    float min_length, int32_t *rows, int32_t *cols, float *weights, int n_threads)
{
    const int K_ = a->n_parameters;
#pragma omp parallel for schedule(dynamic, 16) num_threads(n_threads > 0 ? n_threads : 1)
    for (int64_t g = 0; g < n; g++) {
        int32_t *row = rows + 9 * g;
        int32_t *col = cols + 9 * g;
        float *w = weights + 9 * g;
        for (int i = 0; i < 9; i++) {
            row[i] = -1;
            col[i] = -1;
            w[i] = 0.0f;
        }
        uint32_t rng[2];
        tgo_rng_init(rng_seed, (uint32_t)g, rng);
        float p0[3], o0[3], sbary[3];
        uint32_t stri;
        tgo_sample_boundary(bvertices, btriangles, bcumulative, bn, rng, p0, o0, &stri, sbary);
        VEC point = FN(mkvec)((R)p0[0], (R)p0[1], (R)p0[2]);
        VEC orientation = FN(mkvec)((R)o0[0], (R)o0[1], (R)o0[2]);
        VEC previous = point;
        const R step = (R)a->step_size;
        R params[160];
        uint32_t tracked = 0;
        for (uint32_t k = 0; k < n_steps; k++) {
            if (!FN(mask_contains)(s, point, 0))
                break;
            int64_t base;
            if (!FN(field_index)(s, point, &base, 0))
                break;
            for (int i = 0; i < K_; i++)
                params[i] = (R)s->field[base + i];
            orientation = FN(update_one)(a, params, orientation, rng);
            if (!(isfinite(orientation.x) && isfinite(orientation.y) && isfinite(orientation.z)) ||
                SQRT((orientation.x * orientation.x + orientation.y * orientation.y) +
                     orientation.z * orientation.z) < K(0.5))
                break;
            previous = point;
            point.x += step * orientation.x;
            point.y += step * orientation.y;
            point.z += step * orientation.z;
            tracked++;
        }
        if ((R)tracked * step < (R)min_length)
            continue;
        uint32_t htri;
        R t, hbary[3];
        if (!FN(bvh_hit_bary)(s, previous, point, &htri, &t, hbary))
            continue;
        const uint32_t *si = btriangles + 4 * (size_t)stri;
        const uint32_t *ei = s->triangles + 4 * (size_t)htri;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                row[3 * i + j] = (int32_t)si[i];
                col[3 * i + j] = (int32_t)ei[j];
                w[3 * i + j] = (float)((R)sbary[i] * hbary[j]);
            }
    }
}

/* One flush of a voxel run into the sums (the reference's atomic add,
 * core.cl:80-99, here in seed order). */
static inline void FN(tod_flush)(int64_t cell, const R acc[45], R *raw, double *abs_sum, uint32_t *flushes)
{
    for (int i = 0; i < 45; i++) {
        raw[45 * cell + i] += acc[i];
        if (abs_sum)
            abs_sum[45 * cell + i] += fabs((double)acc[i]);
    }
    if (flushes)
        flushes[cell] += 1;
}

/* tracking.py:262-357 (kernel tod) for seeds 0..n-1, serially in seed order:
 * the boundary sample of rng_init(rng_seed, g), whose state then drives the
 * noise; per step mask, field, then the SH basis (spharm.cl:10-207) of
 * cart2sph(orientation) on the output cell floor(out_inverse_affine p + 0.5)
 * (summed over a run of steps in one cell, flushed on leaving it, on leaving
 * the grid and at the end), then update, orientation check, move. raw
 * (X, Y, Z, 45) receives the sums and must start at zero; abs_sum (same
 * shape, float64) and flushes (X, Y, Z), when given, receive per cell the
 * sum of |flushed value| and the number of flushes, which bound how far a
 * sum of the same flushes in another order can fall from raw. */
void FN(tod)(const tgo_scene *s, const tgo_algorithm *a, const float *bvertices, const uint32_t *btriangles,
             const float *bcumulative, uint32_t bn, uint32_t rng_seed, int64_t n, uint32_t n_steps,
             const uint32_t out_dims[3], const float out_inverse_affine[16], R *raw, double *abs_sum,
             uint32_t *flushes)
{
    const int K_ = a->n_parameters;
    for (int64_t g = 0; g < n; g++) {
        uint32_t rng[2];
        tgo_rng_init(rng_seed, (uint32_t)g, rng);
        float p0[3], o0[3], sbary[3];
        uint32_t stri;
        tgo_sample_boundary(bvertices, btriangles, bcumulative, bn, rng, p0, o0, &stri, sbary);
        VEC point = FN(mkvec)((R)p0[0], (R)p0[1], (R)p0[2]);
        VEC orientation = FN(mkvec)((R)o0[0], (R)o0[1], (R)o0[2]);
        const R step = (R)a->step_size;
        R params[160];
        R acc[45], ylm[45], dp[45], dt[45];
        int64_t previous = -1;
        for (uint32_t k = 0; k < n_steps; k++) {
            if (!FN(mask_contains)(s, point, 0))
                break;
            int64_t base;
            if (!FN(field_index)(s, point, &base, 0))
                break;
            R v[3];
            FN(to_voxel)(out_inverse_affine, point, v);
            int inside = 1;
            int64_t idx[3];
            for (int i = 0; i < 3; i++) {
                const R r = FLOOR(v[i] + K(0.5));
                if (!(r >= K(0.0) && r < (R)out_dims[i]))
                    inside = 0;
                idx[i] = inside ? (int64_t)r : 0;
            }
            if (inside) {
                const int64_t cell = (idx[0] * out_dims[1] + idx[1]) * out_dims[2] + idx[2];
                R az, col;
                FN(cart2sph_one)(orientation, &az, &col);
                FN(sh45_one)(az, col, ylm, dp, dt);
                if (cell == previous) {
                    for (int i = 0; i < 45; i++)
                        acc[i] += ylm[i];
                } else {
                    if (previous >= 0)
                        FN(tod_flush)(previous, acc, raw, abs_sum, flushes);
                    for (int i = 0; i < 45; i++)
                        acc[i] = ylm[i];
                    previous = cell;
                }
            } else if (previous >= 0) {
                FN(tod_flush)(previous, acc, raw, abs_sum, flushes);
                previous = -1;
            }
            for (int i = 0; i < K_; i++)
                params[i] = (R)s->field[base + i];
            orientation = FN(update_one)(a, params, orientation, rng);
            if (!(isfinite(orientation.x) && isfinite(orientation.y) && isfinite(orientation.z)) ||
                SQRT((orientation.x * orientation.x + orientation.y * orientation.y) + orientation.z * orientation.z) <
                    K(0.5))
                break;
            point.x += step * orientation.x;
            point.y += step * orientation.y;
            point.z += step * orientation.z;
        }
        if (previous >= 0)
            FN(tod_flush)(previous, acc, raw, abs_sum, flushes);
    }
}

void FN(sh45)(const R *phi, const R *theta, int64_t n, R *ylm, R *dp, R *dt)
{
    for (int64_t i = 0; i < n; i++)
        FN(sh45_one)(phi[i], theta[i], ylm + 45 * i, dp + 45 * i, dt + 45 * i);
}

void FN(evaluate)(int model, int n_parameters, const R *params, const R *orientations, int64_t n, R *out)
{
    for (int64_t i = 0; i < n; i++) {
        const VEC o = FN(mkvec)(orientations[3 * i], orientations[3 * i + 1], orientations[3 * i + 2]);
        const MV mv = FN(model_evaluate)(model, n_parameters, params + (size_t)n_parameters * i, o);
        out[3 * i] = mv.value;
        out[3 * i + 1] = mv.dtheta;
        out[3 * i + 2] = mv.dphi;
    }
}

void FN(update)(const tgo_algorithm *a, uint32_t rng_seed, const R *params, const R *orientations, int64_t n,
                     R *out)
{
    for (int64_t i = 0; i < n; i++) {
        uint32_t rng[2];
        tgo_rng_init(rng_seed, (uint32_t)i, rng);
        const VEC o = FN(mkvec)(orientations[3 * i], orientations[3 * i + 1], orientations[3 * i + 2]);
        const VEC r = FN(update_one)(a, params + (size_t)a->n_parameters * i, o, rng);
        out[3 * i] = r.x;
        out[3 * i + 1] = r.y;
        out[3 * i + 2] = r.z;
    }
}

/* One update of one orientation from an explicit noise state, which is
 * advanced in place (the state a track carries from step to step). */
void FN(update_state)(const tgo_algorithm *a, const R *params, const R orientation[3], uint32_t state[2], R out[3])
{
    const VEC o = FN(mkvec)(orientation[0], orientation[1], orientation[2]);
    const VEC r = FN(update_one)(a, params, o, state);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

#undef VEC
#undef MV
