// tgk_plain_model.h -- the reference-semantics (plain FP32) local models and
// orientation updates, device side.
//
// Results are bit-identical to a literal FP32 reading of the reference OpenCL
// (every product and sum rounded in source order, no contraction) given the
// elementary functions of tgk_math.h. Include only from translation units
// compiled with -ffp-contract=off.
#pragma once

#include <type_traits>

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

// SH model contraction (local.py:91-120): value, d/dtheta and
// (d/dphi)/sin(theta) of sum_i c_i Y_i. Terms are accumulated in index order,
// so sums round exactly as the reference's `for (i < K)` loop; coefficients
// past K arrive as +0 and leave the sums unchanged.
__device__ __forceinline__ void sh45_contract(const float c[45], float phi, float theta, float &v, float &th,
                                              float &ph)
{
    float sv = 0.0f, sth = 0.0f, sph = 0.0f;
#define TGK_ACC(i, Y, DP, DT)   \
    {                           \
        const float ci_ = c[i]; \
        sv += ci_ * (Y);        \
        sth += ci_ * (DT);      \
        sph += ci_ * (DP);      \
    }
#include "tgk_sh45_body.inc"
#undef TGK_ACC
    v = sv;
    th = sth;
    ph = sph;
}

// No use for the basis values beyond the contraction.
struct NoTerms {
    __device__ __forceinline__ void operator()(int, float) const {}
};

// The same contraction handing each basis value to terms(i, Y_i) as well
// (all 45, whatever K).
template <class Terms>
__device__ __forceinline__ void sh45_contract_terms(const float c[45], float phi, float theta, float &v, float &th,
                                                    float &ph, Terms &&terms)
{
    float sv = 0.0f, sth = 0.0f, sph = 0.0f;
#define TGK_ACC(i, Y, DP, DT)   \
    {                           \
        const float ci_ = c[i]; \
        const float y_ = (Y);   \
        sv += ci_ * y_;         \
        sth += ci_ * (DT);      \
        sph += ci_ * (DP);      \
        terms(i, y_);           \
    }
#include "tgk_sh45_body.inc"
#undef TGK_ACC
    v = sv;
    th = sth;
    ph = sph;
}

// The basis values alone, handed to visit(i, Y_i) in index order.
template <class Visit>
__device__ __forceinline__ void sh45_values(float phi, float theta, Visit &&visit)
{
#define TGK_ACC(i, Y, DP, DT) visit(i, (Y));
#include "tgk_sh45_body.inc"
#undef TGK_ACC
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
    if constexpr (std::is_same_v<std::decay_t<Terms>, NoTerms>)
        sh45_contract(c, az, col, v, th, ph);
    else
        sh45_contract_terms(c, az, col, v, th, ph, terms);
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
