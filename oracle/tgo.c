/*
 * tgo.c -- host restatement of the reference tracking path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py as the checker; the product library never
 * links or calls it. Build with oracle/Makefile (gcc -O2 -ffp-contract=off).
 *
 * The FP32 build shares only the elementary functions of
 * tractography_amd/csrc/tgk_math.h with the kernels (bit-identical libm on
 * both sides is a precondition of bit-exact parity); everything else here is
 * restated independently from the reference sources.
 */
#include "tgo.h"

#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include "../tractography_amd/csrc/tgk_math.h"

/* ---------------------------------------------------------------- RNG --- */

/* core.cl:180-190: multiply-with-carry, A = 4294883355. */
static inline uint32_t tgo_mwc64x(uint32_t state[2])
{
    const uint32_t A = 4294883355u;
    uint32_t x = state[0], c = state[1];
    const uint32_t res = x ^ c;
    const uint32_t hi = (uint32_t)(((uint64_t)x * A) >> 32);
    x = x * A + c;
    c = hi + (x < c);
    state[0] = x;
    state[1] = c;
    return res;
}

/* core.cl:193-201 */
static inline uint32_t tgo_rng_mix32(uint32_t v)
{
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

/* core.cl:203-208 */
void tgo_rng_init(uint32_t seed, uint32_t stream_id, uint32_t state[2])
{
    uint32_t x = tgo_rng_mix32(seed ^ (stream_id + 0x9e3779b9u));
    if (x < 1u)
        x = 1u;
    state[0] = x;
    state[1] = tgo_rng_mix32(seed + stream_id * 0x85ebca6bu) % 4294883355u;
}

/* core.cl:212-214: exact, (res >> 8) * 2^-24. */
static inline float tgo_randu(uint32_t state[2]) { return (float)(tgo_mwc64x(state) >> 8) * 0x1p-24f; }

/* core.cl:227-229 */
static inline uint32_t tgo_randi(uint32_t state[2], uint32_t max) { return (uint32_t)(tgo_randu(state) * (float)max); }

/* domain.py:131-166 (domain_sample_boundary): area-weighted triangle by
 * binary search of the cumulative areas, uniform point from sqrt(randu) and
 * randu, inward normal as orientation (normalize(v) restated as
 * v / sqrt(dot(v, v)) per component), the point moved 1e-4 inward; the
 * triangle and the barycentric weights of the sample are returned too. */
static void tgo_sample_boundary(const float *vertices, const uint32_t *triangles, const float *cumulative,
                                uint32_t n_triangles, uint32_t rng[2], float point[3], float orientation[3],
                                uint32_t *triangle, float bary[3])
{
    const float target = tgo_randu(rng);
    uint32_t lower = 0, upper = n_triangles;
    while (lower < upper) {
        const uint32_t middle = lower + (upper - lower) / 2;
        if (cumulative[middle] <= target)
            lower = middle + 1;
        else
            upper = middle;
    }
    const uint32_t tri = lower < n_triangles - 1 ? lower : n_triangles - 1;
    const uint32_t *ix = triangles + 4 * (size_t)tri;
    float a[3], b[3], c[3];
    for (int i = 0; i < 3; i++) {
        a[i] = vertices[4 * (size_t)ix[0] + i];
        b[i] = vertices[4 * (size_t)ix[1] + i];
        c[i] = vertices[4 * (size_t)ix[2] + i];
    }
    const float root_u = tgk_sqrtf(tgo_randu(rng));
    const float v = tgo_randu(rng);
    const float wa = 1.0f - root_u;
    const float wb = root_u * (1.0f - v);
    const float wc = root_u * v;
    bary[0] = wa;
    bary[1] = wb;
    bary[2] = wc;
    const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const float nrm[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const float len = tgk_sqrtf((nrm[0] * nrm[0] + nrm[1] * nrm[1]) + nrm[2] * nrm[2]);
    for (int i = 0; i < 3; i++) {
        const float sampled = (wa * a[i] + wb * b[i]) + wc * c[i];
        orientation[i] = -(nrm[i] / len);
        point[i] = sampled + 1e-4f * orientation[i];
    }
    *triangle = tri;
}

/* ------------------------------------------------------ FP32 instance --- */

static inline void tgo_sincos32(float x, float *s, float *c) { tgk_sincosf(x, s, c); }

#define R float
#define FN(x) tgo32_##x
#define K(c) c##f
#define PI_R 3.14159265359f
#define SQRT tgk_sqrtf
#define SINCOS tgo_sincos32
#define COS_OF tgk_cosf
#define ATAN2 tgk_atan2f
#define ACOS tgk_acosf
#define EXP tgk_expf
#define LOG tgk_logf
#define FLOOR floorf
#define RINT rintf
#define FMAX fmaxf
#define FMIN fminf
#define FABS fabsf
#define TRACK_EXTRA_PARAM
#define TRACK_MARGINS(g) ((float *)0)
#include "tgo_impl.h"
#undef R
#undef FN
#undef K
#undef PI_R
#undef SQRT
#undef SINCOS
#undef COS_OF
#undef ATAN2
#undef ACOS
#undef EXP
#undef LOG
#undef FLOOR
#undef RINT
#undef FMAX
#undef FMIN
#undef FABS
#undef TRACK_EXTRA_PARAM
#undef TRACK_MARGINS

/* ------------------------------------------------------ FP64 instance --- */

static inline void tgo_sincos64(double x, double *s, double *c)
{
    *s = sin(x);
    *c = cos(x);
}

/* Real symmetric SH of any even order lmax <= TGO_LMAX_GENERIC with the
 * colatitude derivative and the azimuth derivative over sin(colatitude), in
 * the convention of core.py:36-69 (MRtrix order, index l(l+1)/2 + m; m > 0
 * sqrt(2) N P_l^m cos(m phi), m < 0 sqrt(2) N P_l^|m| sin(|m| phi), P with
 * the Condon-Shortley phase). No reference kernel goes past l = 8: this
 * extends the FP64 oracle to order 12 (91 coefficients) for the fast path.
 * P_l^m by the three-term recurrence in l; P_l^m / sin(theta) (m >= 1) by the
 * same recurrence from (-1)^m (2m-1)!! sin^(m-1); d P_l^m / d theta =
 * (P_l^(m+1) - (l+m)(l-m+1) P_l^(m-1)) / 2, and P_l^1 for m = 0. */
#define TGO_LMAX_GENERIC 16
static void tgo_sh_generic64(int lmax, double phi, double theta, double *ylm, double *dp, double *dt)
{
    enum { L = TGO_LMAX_GENERIC + 2 };
    double P[L][L], Q[L][L];
    const double st = sin(theta), ct = cos(theta);
    for (int l = 0; l < L; l++)
        for (int m = 0; m < L; m++)
            P[l][m] = Q[l][m] = 0.0;
    const int top = lmax + 1; /* P_l^(m+1) needs m + 1 <= l */
    double pmm = 1.0, qmm = 1.0;
    for (int m = 0; m <= top; m++) {
        if (m > 0) {
            pmm *= -(2.0 * m - 1.0) * st;
            qmm = (m == 1) ? -1.0 : qmm * -(2.0 * m - 1.0) * st;
        }
        P[m][m] = pmm;
        Q[m][m] = m > 0 ? qmm : 0.0;
        if (m + 1 <= top) {
            P[m + 1][m] = (2.0 * m + 1.0) * ct * pmm;
            Q[m + 1][m] = (2.0 * m + 1.0) * ct * Q[m][m];
        }
        for (int l = m + 2; l <= top; l++) {
            P[l][m] = ((2.0 * l - 1.0) * ct * P[l - 1][m] - (l + m - 1.0) * P[l - 2][m]) / (double)(l - m);
            Q[l][m] = ((2.0 * l - 1.0) * ct * Q[l - 1][m] - (l + m - 1.0) * Q[l - 2][m]) / (double)(l - m);
        }
    }
    for (int l = 0; l <= lmax; l += 2) {
        const int centre = l * (l + 1) / 2;
        for (int m = 0; m <= l; m++) {
            double ratio = 1.0; /* (l - m)! / (l + m)! */
            for (int k = l - m + 1; k <= l + m; k++)
                ratio /= (double)k;
            const double N = sqrt((2.0 * l + 1.0) / (4.0 * 3.141592653589793) * ratio);
            const double dP = m == 0 ? P[l][1] : 0.5 * (P[l][m + 1] - (double)(l + m) * (l - m + 1.0) * P[l][m - 1]);
            if (m == 0) {
                ylm[centre] = N * P[l][0];
                dt[centre] = N * dP;
                dp[centre] = 0.0;
            } else {
                const double F = sqrt(2.0) * N, cm = cos(m * phi), sm = sin(m * phi);
                ylm[centre + m] = F * P[l][m] * cm;
                dt[centre + m] = F * dP * cm;
                dp[centre + m] = -F * m * Q[l][m] * sm;
                ylm[centre - m] = F * P[l][m] * sm;
                dt[centre - m] = F * dP * sm;
                dp[centre - m] = F * m * Q[l][m] * cm;
            }
        }
    }
}

void tgo64_sh_basis(int lmax, const double *phi, const double *theta, int64_t n, double *ylm, double *dp, double *dt)
{
    const int K = (lmax + 1) * (lmax + 2) / 2;
    for (int64_t i = 0; i < n; i++)
        tgo_sh_generic64(lmax, phi[i], theta[i], ylm + (size_t)K * i, dp + (size_t)K * i, dt + (size_t)K * i);
}

#define TGO_GENERIC_SH 1
#define R double
#define FN(x) tgo64_##x
#define K(c) c
#define PI_R 3.141592653589793
#define SQRT sqrt
#define SINCOS tgo_sincos64
#define COS_OF cos
#define ATAN2 atan2
#define ACOS acos
#define EXP exp
#define LOG log
#define FLOOR floor
#define RINT rint
#define FMAX fmax
#define FMIN fmin
#define FABS fabs
#define TRACK_EXTRA_PARAM , double *margins
#define TRACK_MARGINS(g) (margins ? margins + (size_t)n_points * (size_t)(g) : (double *)0)
#include "tgo_impl.h"

/* ------------------------------------------------------- one-op tools --- */

void tgo_math32(int fn, const float *x, const float *y, float *out, int64_t n)
{
    for (int64_t i = 0; i < n; i++) {
        switch (fn) {
        case 0: out[i] = tgk_sinf(x[i]); break;
        case 1: out[i] = tgk_cosf(x[i]); break;
        case 2: out[i] = tgk_atan2f(x[i], y[i]); break;
        case 3: out[i] = tgk_acosf(x[i]); break;
        case 4: out[i] = tgk_expf(x[i]); break;
        case 5: out[i] = tgk_logf(x[i]); break;
        default: out[i] = tgk_sqrtf(x[i]); break;
        }
    }
}

void tgo_rng_draws(uint32_t state[2], int kind, uint32_t aux, int64_t n, void *out)
{
    for (int64_t i = 0; i < n; i++) {
        switch (kind) {
        case 0: ((uint32_t *)out)[i] = tgo_mwc64x(state); break;
        case 1: ((float *)out)[i] = tgo_randu(state); break;
        case 2: ((float *)out)[i] = tgo32_randn(state); break;
        case 3: ((double *)out)[i] = tgo64_randn(state); break;
        default: ((uint32_t *)out)[i] = tgo_randi(state, aux); break;
        }
    }
}

void tgo32_cart2sph(const float *xyz, int64_t n, float *out)
{
    for (int64_t i = 0; i < n; i++)
        tgo32_cart2sph_one(tgo32_mkvec(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]), out + 2 * i, out + 2 * i + 1);
}

void tgo_field_evaluate(const tgo_scene *s, const float *points, int64_t n, float *out)
{
    const int K_ = (int)s->field_dims[3];
    for (int64_t i = 0; i < n; i++) {
        int64_t base = 0;
        const tgo32_vec p = tgo32_mkvec(points[3 * i], points[3 * i + 1], points[3 * i + 2]);
        const int inside = tgo32_field_index(s, p, &base, 0);
        for (int k = 0; k < K_; k++)
            out[(size_t)K_ * i + k] = inside ? s->field[base + k] : 0.0f;
    }
}

void tgo_domain_contains(const tgo_scene *s, const float *points, int64_t n, uint8_t *out)
{
    for (int64_t i = 0; i < n; i++)
        out[i] = (uint8_t)tgo32_mask_contains(s, tgo32_mkvec(points[3 * i], points[3 * i + 1], points[3 * i + 2]), 0);
}

void tgo_bvh_segment_hit(const tgo_scene *s, const float *starts, const float *ends, int64_t n, uint8_t *found,
                         uint32_t *hit, float *t)
{
    for (int64_t i = 0; i < n; i++) {
        const tgo32_vec a = tgo32_mkvec(starts[3 * i], starts[3 * i + 1], starts[3 * i + 2]);
        const tgo32_vec b = tgo32_mkvec(ends[3 * i], ends[3 * i + 1], ends[3 * i + 2]);
        found[i] = (uint8_t)tgo32_bvh_hit(s, a, b, hit + i, t + i);
    }
}

/* ---------------------------------------------------- boundary seeding --- */

/* domain.py:494-575: boundary seed g from rng_init(rng_seed, g). */
void tgo_domain_seed(const float *vertices, const uint32_t *triangles, const float *cumulative, uint32_t n_triangles,
                     uint32_t rng_seed, int64_t n, float *out)
{
    for (int64_t g = 0; g < n; g++) {
        uint32_t rng[2];
        tgo_rng_init(rng_seed, (uint32_t)g, rng);
        float point[3], orientation[3], bary[3];
        uint32_t tri;
        tgo_sample_boundary(vertices, triangles, cumulative, n_triangles, rng, point, orientation, &tri, bary);
        float *row = out + 8 * g;
        for (int i = 0; i < 3; i++) {
            row[i] = point[i];
            row[4 + i] = orientation[i];
        }
        row[3] = 1.0f;
        row[7] = 0.0f;
    }
}

/* ------------------------------------------------------- FOD seeding --- */

/* seeds.cl:12-25 */
static inline tgo32_vec tgo_fibonacci(uint32_t n, uint32_t n_points)
{
    const float phi = (tgk_sqrtf(5.0f) + 1.0f) * 0.5f;
    const float golden_angle = 2.0f * 3.14159265359f / phi;
    const float theta = (float)n * golden_angle;
    const float z = 1.0f - (2.0f * (float)n / (float)(n_points - 1));
    const float radius = tgk_sqrtf(1.0f - z * z);
    float st, ct;
    tgk_sincosf(theta, &st, &ct);
    return tgo32_mkvec(radius * ct, radius * st, z);
}

/* seeds.cl:28-62 */
static tgo32_vec tgo_sample_fod(int model, int K_, const float *fod, uint32_t state[2])
{
    float sum = 0.0f;
    for (uint32_t i = 0; i < 1000; i++) {
        const tgo32_model_value mv = tgo32_model_evaluate(model, K_, fod, tgo_fibonacci(i, 1000));
        sum += fmaxf(mv.value, 0.0f);
    }
    const float threshold = tgo_randu(state) * sum;
    float cs = 0.0f;
    uint32_t selected = 1000;
    for (uint32_t i = 0; i < 1000; i++) {
        const tgo32_model_value mv = tgo32_model_evaluate(model, K_, fod, tgo_fibonacci(i, 1000));
        cs += fmaxf(mv.value, 0.0f);
        if (selected == 1000 && cs > threshold)
            selected = i;
    }
    if (selected < 1000)
        return tgo_fibonacci(selected, 1000);
    return tgo32_mkvec(0.0f, 0.0f, 1.0f);
}

/* seeds.cl:68-118 */
void tgo_seeds_from_fod(int model, int K_, const float *fod, const float *voxels, uint32_t nnz, const float *affine,
                        uint32_t *states, int64_t n, float *out, int n_threads)
{
#pragma omp parallel for schedule(dynamic, 16) num_threads(n_threads > 0 ? n_threads : 1)
    for (int64_t g = 0; g < n; g++) {
        uint32_t *state = states + 2 * g;
        const uint32_t index = tgo_randi(state, nnz);
        float voxel[4];
        for (int i = 0; i < 4; i++)
            voxel[i] = voxels[4 * (size_t)index + i];
        const float jx = tgo_randu(state) - 0.5f;
        const float jy = tgo_randu(state) - 0.5f;
        const float jz = tgo_randu(state) - 0.5f;
        voxel[0] += jx;
        voxel[1] += jy;
        voxel[2] += jz;
        const tgo32_vec o = tgo_sample_fod(model, K_, fod + (size_t)K_ * index, state);
        float *row = out + 8 * g;
        for (int i = 0; i < 4; i++)
            row[i] = ((affine[4 * i] * voxel[0] + affine[4 * i + 1] * voxel[1]) + affine[4 * i + 2] * voxel[2]) +
                     affine[4 * i + 3] * voxel[3];
        row[4] = o.x;
        row[5] = o.y;
        row[6] = o.z;
        row[7] = 0.0f;
    }
}
