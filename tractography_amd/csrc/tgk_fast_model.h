// tgk_fast_model.h -- model evaluation and orientation update of the fast
// path (device side). Include only from tgk_fast.hip (contraction on).
//
// Real symmetric SH to even order L (8, or 12 past the reference's basis) in
// polynomial form. For a unit direction with z = cos(theta), s = sin(theta)
// and u = s e^{i phi} = x + i y:
//
//   f     = A_0(z) + sum_{m>=1} A_m(z) Re u^m + B_m(z) Im u^m
//   df/dt = -s A_0'(z) + sum_{m>=1} m z (A_m Re v_m + B_m Im v_m)
//                                  - s (A_m' Re u^m + B_m' Im u^m)
//   (df/dphi)/sin(theta) = sum_{m>=1} m (B_m Re v_m - A_m Im v_m)
//
// with v_m = u^{m-1} e^{i phi} and A_m, B_m polynomials in z whose monomial
// coefficients are a fixed linear map of the voxel's SH coefficients (A_m
// collects the order +m harmonics, B_m the order -m ones; each includes the
// normalisation, the sqrt(2) of the real basis and the Condon-Shortley
// sign). The map is applied once per field upload (launch_field_transform),
// so a step costs ~2K polynomial FMAs and two complex power sequences
// instead of the trigonometric basis.
#pragma once

#include "tgk_device.h"

namespace tgk {
namespace fast {

// Layout of the (L+1)(L+2)/2 transformed coefficients per voxel: for
// m = 0..L, the coefficients of A_m (then B_m for m > 0) by ascending power
// of z, powers m mod 2, m mod 2 + 2, ..., L - m.
template <int L> struct Order {
    static constexpr int TERMS = (L + 1) * (L + 2) / 2;
    static constexpr int STRIDE = (TERMS + 3) / 4 * 4;
    __host__ __device__ static constexpr int poly_terms(int m) { return (L - m - (m & 1)) / 2 + 1; }
    __host__ __device__ static constexpr int a_offset(int m)
    {
        int off = 0;
        for (int k = 0; k < m; k++)
            off += poly_terms(k) * (k == 0 ? 1 : 2);
        return off;
    }
    __host__ __device__ static constexpr int b_offset(int m) { return a_offset(m) + poly_terms(m); }
};

// q(w) and q'(w) for q = sum_j c[j] w^j (Horner; the derivative's first
// step is peeled, 0 * w + v not being foldable under IEEE rules).
template <int N>
__device__ __forceinline__ void horner(const float *c, float w, float &v, float &d)
{
    v = c[N - 1];
    if constexpr (N == 1) {
        d = 0.0f;
    } else {
        d = v;
        v = v * w + c[N - 2];
#pragma unroll
        for (int j = N - 3; j >= 0; j--) {
            d = d * w + v;
            v = v * w + c[j];
        }
    }
}

// A(z) and A'(z) for A(z) = z^p q(z^2), p = m mod 2, q with N terms.
template <int M, int N>
__device__ __forceinline__ void poly_m(const float *c, float z, float w, float &a, float &ad)
{
    float v, d;
    horner<N>(c, w, v, d);
    if constexpr ((M & 1) == 0) {
        a = v;
        ad = 2.0f * z * d;
    } else {
        a = z * v;
        ad = v + 2.0f * w * d;
    }
}

struct Direction {
    float z, s;   // cos and sin of the colatitude
    float c1, s1; // cos and sin of the azimuth
};

// Angles of a (not necessarily unit) orientation, as the reference's
// cart2sph: the zero vector maps to colatitude 0, azimuth 0.
__device__ __forceinline__ Direction direction_of(V3 o)
{
    Direction d;
    const float rho2 = o.x * o.x + o.y * o.y;
    const float r2 = rho2 + o.z * o.z;
    if (r2 > 0.0f) {
        const float inv_r = __builtin_amdgcn_rsqf(r2);
        d.z = fminf(fmaxf(o.z * inv_r, -1.0f), 1.0f);
        d.s = rho2 * __builtin_amdgcn_rsqf(r2 * rho2 > 0.0f ? r2 * rho2 : 1.0f);
    } else {
        d.z = 1.0f;
        d.s = 0.0f;
    }
    if (rho2 > 0.0f) {
        const float inv_rho = __builtin_amdgcn_rsqf(rho2);
        d.c1 = o.x * inv_rho;
        d.s1 = o.y * inv_rho;
    } else {
        // On the axis the azimuth is atan2(y, x) of signed zeros: 0 for
        // x = +0, +-pi for x = -0 (axis-aligned orientations, such as the
        // inward normals of boundary seeds, carry such zeros); 0 for the
        // zero vector.
        d.c1 = r2 > 0.0f && __builtin_signbit(o.x) ? -1.0f : 1.0f;
        d.s1 = 0.0f;
    }
    return d;
}

struct Raw {
    float value, dtheta, dphi;
};

// Running state of the order loop: u^m and v_m.
struct Powers {
    float ur, ui, vr, vi;
};

template <int L, int M>
__device__ __forceinline__ void sh_order_terms(const float *q, float z, float w, float s, float u1r, float u1i,
                                               Powers &pw, Raw &r)
{
    // Fused as the compiler finds best, as in the rest of this file before
    // contraction became per expression: the Deterministic Hessian divides
    // the rounding of these gradients by 2h (2.7e-3 voxel from FP64 with
    // per-expression fusion, 1.1e-3 with this). Only the unshaped
    // Deterministic kernels use the theta/phi form.
    #pragma clang fp contract(fast)
    if constexpr (M <= L) {
        using O = Order<L>;
        if constexpr (M > 1) {
            const float nvr = pw.vr * u1r - pw.vi * u1i;
            pw.vi = pw.vr * u1i + pw.vi * u1r;
            pw.vr = nvr;
        }
        const float nur = pw.ur * u1r - pw.ui * u1i;
        pw.ui = pw.ur * u1i + pw.ui * u1r;
        pw.ur = nur;
        float a, ad, b, bd;
        poly_m<M, O::poly_terms(M)>(q + O::a_offset(M), z, w, a, ad);
        poly_m<M, O::poly_terms(M)>(q + O::b_offset(M), z, w, b, bd);
        r.value += a * pw.ur + b * pw.ui;
        r.dtheta += (float)M * z * (a * pw.vr + b * pw.vi) - s * (ad * pw.ur + bd * pw.ui);
        r.dphi += (float)M * (b * pw.vr - a * pw.vi);
        sh_order_terms<L, M + 1>(q, z, w, s, u1r, u1i, pw, r);
    }
}

template <int L>
__device__ __forceinline__ Raw sh_evaluate(const float *q, const Direction &dir)
{
    #pragma clang fp contract(fast)
    using O = Order<L>;
    const float z = dir.z, s = dir.s, w = z * z;
    Raw r;
    float a, ad;
    poly_m<0, O::poly_terms(0)>(q + O::a_offset(0), z, w, a, ad);
    r.value = a;
    r.dtheta = -s * ad;
    r.dphi = 0.0f;
    Powers pw{1.0f, 0.0f, dir.c1, dir.s1};
    sh_order_terms<L, 1>(q, z, w, s, s * dir.c1, s * dir.s1, pw, r);
    return r;
}

// Softplus with k = 100 and its derivative (core.cl:111-120) from one
// hardware exp2: e = exp(100 x), value = log(1 + e) / 100, d = e / (1 + e).
__device__ __forceinline__ void softplus(float x, float &value, float &d)
{
    if (x * 100.0f >= 30.0f) {
        value = x;
        d = 1.0f;
        return;
    }
    const float e = __builtin_amdgcn_exp2f(x * 144.26950408889634f);
    const float one_e = 1.0f + e;
    value = __builtin_amdgcn_logf(one_e) * 0.006931471805599453f;
    d = e * __builtin_amdgcn_rcpf(one_e);
}

struct ModelOut {
    float value, dtheta, dphi;
};

template <int L>
__device__ __forceinline__ ModelOut sh_model(const float *q, const Direction &dir)
{
    #pragma clang fp contract(fast)
    const Raw raw = sh_evaluate<L>(q, dir);
    ModelOut m;
    float d;
    softplus(raw.value, m.value, d);
    m.dtheta = raw.dtheta * d;
    m.dphi = raw.dphi * d;
    return m;
}

// The SH value and its gradient on the sphere without angles. The
// tangential gradient of a function on the unit sphere is the ambient
// gradient G of any extension, minus its normal part: G - (G.n) n. The
// extension taken is P(x, y, z) = A_0(z) + sum_m A_m(z) Re (x+iy)^m +
// B_m(z) Im (x+iy)^m, whose partial derivatives are
//   dP/dx = sum_m m (A_m Re w_m + B_m Im w_m),
//   dP/dy = sum_m m (B_m Re w_m - A_m Im w_m),
//   dP/dz = A_0'(z) + sum_m A_m'(z) Re u^m + B_m'(z) Im u^m,
// with u = x + iy and w_m = u^(m-1): one complex power per order and no
// sin(theta) factors, so it is also regular at the poles.
struct Cartesian {
    float value, gx, gy, gz;
};

template <int L, int M>
__device__ __forceinline__ void sh_cartesian_terms(const float *q, float z, float w, float x, float y, float &pr,
                                                   float &pi, Cartesian &g)
{
    if constexpr (M <= L) {
        using O = Order<L>;
        float ur, ui;
        if constexpr (M == 1) {
            ur = x;
            ui = y;
        } else {
            ur = pr * x - pi * y;
            ui = pr * y + pi * x;
        }
        float a, ad, b, bd;
        poly_m<M, O::poly_terms(M)>(q + O::a_offset(M), z, w, a, ad);
        poly_m<M, O::poly_terms(M)>(q + O::b_offset(M), z, w, b, bd);
        g.value += a * ur + b * ui;
        g.gz += ad * ur + bd * ui;
        if constexpr (M == 1) {
            g.gx += a;
            g.gy += b;
        } else {
            g.gx += (float)M * (a * pr + b * pi);
            g.gy += (float)M * (b * pr - a * pi);
        }
        pr = ur;
        pi = ui;
        sh_cartesian_terms<L, M + 1>(q, z, w, x, y, pr, pi, g);
    }
}

// Raw (pre-softplus) value and ambient gradient of P at the unit vector n.
template <int L>
__device__ __forceinline__ Cartesian sh_cartesian(const float *q, V3 n)
{
    using O = Order<L>;
    const float w = n.z * n.z;
    Cartesian g;
    poly_m<0, O::poly_terms(0)>(q + O::a_offset(0), n.z, w, g.value, g.gz);
    g.gx = 0.0f;
    g.gy = 0.0f;
    float pr = 1.0f, pi = 0.0f;
    sh_cartesian_terms<L, 1>(q, n.z, w, n.x, n.y, pr, pi, g);
    return g;
}

// The direction cart2sph evaluates the model at: o / |o|, and the zero
// vector at colatitude 0.
__device__ __forceinline__ V3 unit_direction(V3 o)
{
    const float r2 = o.x * o.x + o.y * o.y + o.z * o.z;
    if (!(r2 > 0.0f))
        return v3(0.0f, 0.0f, 1.0f);
    const float inv = __builtin_amdgcn_rsqf(r2);
    return v3(o.x * inv, o.y * inv, o.z * inv);
}

// local.py:212-230: t = D o, f = o.t, df/dtheta = 2 t.e_theta,
// (df/dphi)/sin(theta) = 2 t.e_phi (o as given, unit frame vectors).
__device__ __forceinline__ ModelOut dti_model(const float m[8], V3 o, const Direction &dir)
{
    const float t0 = m[0] * o.x + m[3] * o.y + m[4] * o.z;
    const float t1 = m[3] * o.x + m[1] * o.y + m[5] * o.z;
    const float t2 = m[4] * o.x + m[5] * o.y + m[2] * o.z;
    ModelOut r;
    r.value = o.x * t0 + o.y * t1 + o.z * t2;
    r.dtheta = 2.0f * (dir.z * (t0 * dir.c1 + t1 * dir.s1) - t2 * dir.s);
    r.dphi = 2.0f * (t1 * dir.c1 - t0 * dir.s1);
    return r;
}

// Box-Muller normal from the MWC64X stream (core.cl:240-245) with the
// hardware log2 and cos(2 pi x) (v_cos_f32 takes revolutions).
__device__ __forceinline__ float randn(uint2 &rng)
{
    float u = randu(rng);
    while (u <= 0.0f)
        u = randu(rng);
    const float radius = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u));
    return radius * __builtin_amdgcn_cosf(randu(rng));
}

// exps2(p, x, 1) (core.cl:101-109) with cos(n) and sin(n)/n as Taylor
// polynomials in w = |x|^2 for |x| < 0.5 (truncation below 2e-8); larger
// turns, rare and where trajectories are most sensitive, use the accurate
// sincos of tgk_math.h rather than the hardware approximation.
__device__ __forceinline__ V3 exps2(V3 p, V3 x)
{
    const float w = x.x * x.x + x.y * x.y + x.z * x.z;
    if (!(w > 0.0f))
        return p;
    float c, sn;
    if (w < 0.25f) {
        c = 1.0f + w * (-0.5f + w * (1.0f / 24.0f + w * (-1.0f / 720.0f + w * (1.0f / 40320.0f))));
        sn = 1.0f + w * (-1.0f / 6.0f + w * (1.0f / 120.0f + w * (-1.0f / 5040.0f + w * (1.0f / 362880.0f))));
    } else {
        const float n = __builtin_sqrtf(w);
        float s;
        tgk_sincosf(n, &s, &c);
        sn = s / n;
    }
    return v3(c * p.x + sn * x.x, c * p.y + sn * x.y, c * p.z + sn * x.z);
}

// The two normal draws of a Diffusion update (algorithm.py:217-220), in the
// reference's order; none for Transport or a zero noise variance, which
// consume no random numbers. Drawn apart from the update so that a step can
// draw them before its memory-dependent work: the draws depend only on the
// lane's stream, and a lane whose streamline stops in that step discards
// its stream.
struct Noise {
    float r1, r2;
};

template <int ALGO> __device__ __forceinline__ Noise draw_noise(const AlgoParams &a, uint2 &rng)
{
    Noise n{0.0f, 0.0f};
    if constexpr (ALGO == ALGO_DIFFUSION) {
        if (a.noise_variance > 0.0f) {
            n.r1 = randn(rng);
            n.r2 = randn(rng);
        }
    }
    return n;
}

template <int ALGO>
__device__ __forceinline__ V3 update(const AlgoParams &a, const ModelOut &mv, const Direction &dir, V3 o,
                                     const Noise &nz)
{
    const float etx = dir.z * dir.c1, ety = dir.z * dir.s1, etz = -dir.s;
    const float epx = -dir.s1, epy = dir.c1;
    const float inv = __builtin_amdgcn_rcpf(fmaxf(mv.value, 1e-6f));
    const float scale = a.gamma * a.step_size * inv;
    float tx = scale * (mv.dtheta * etx + mv.dphi * epx);
    float ty = scale * (mv.dtheta * ety + mv.dphi * epy);
    float tz = scale * (mv.dtheta * etz);
    if constexpr (ALGO == ALGO_DIFFUSION) {
        if (a.noise_variance > 0.0f) {
            const float r1 = nz.r1, r2 = nz.r2;
            const float sn = __builtin_sqrtf(a.noise_variance * a.gamma * a.step_size);
            tx += sn * (r1 * etx + r2 * epx);
            ty += sn * (r1 * ety + r2 * epy);
            tz += sn * (r1 * etz);
        }
    }
    return exps2(o, v3(tx, ty, tz));
}

// Transport / Diffusion with the SH model: the same update as above, the
// model gradient taken in Cartesian form (sh_cartesian). The tangent frame
// (e_theta, e_phi) is formed only for the Diffusion noise, which is drawn
// in it.
template <int ALGO, int L>
__device__ __forceinline__ V3 sh_update(const AlgoParams &a, const float *q, V3 o, const Noise &nz)
{
    const V3 n = unit_direction(o);
    const Cartesian g = sh_cartesian<L>(q, n);
    float value, d;
    softplus(g.value, value, d);
    const float normal = g.gx * n.x + g.gy * n.y + g.gz * n.z;
    const float scale = a.gamma * a.step_size * __builtin_amdgcn_rcpf(fmaxf(value, 1e-6f)) * d;
    float tx = scale * (g.gx - normal * n.x);
    float ty = scale * (g.gy - normal * n.y);
    float tz = scale * (g.gz - normal * n.z);
    if constexpr (ALGO == ALGO_DIFFUSION) {
        if (a.noise_variance > 0.0f) {
            const float rho2 = n.x * n.x + n.y * n.y;
            // On the axis, the azimuth of signed zeros (see direction_of).
            float c1 = __builtin_signbit(n.x) ? -1.0f : 1.0f, s1 = 0.0f, sin_theta = 0.0f;
            if (rho2 > 0.0f) {
                const float inv_rho = __builtin_amdgcn_rsqf(rho2);
                c1 = n.x * inv_rho;
                s1 = n.y * inv_rho;
                sin_theta = rho2 * inv_rho;
            }
            const float z = fminf(fmaxf(n.z, -1.0f), 1.0f);
            const float r1 = nz.r1, r2 = nz.r2;
            const float sn = __builtin_sqrtf(a.noise_variance * a.gamma * a.step_size);
            tx += sn * (r1 * (z * c1) - r2 * s1);
            ty += sn * (r1 * (z * s1) + r2 * c1);
            tz += sn * (r1 * -sin_theta);
        }
    }
    return exps2(o, v3(tx, ty, tz));
}

// ------------------------------------------ Probabilistic, Deterministic ---

// Model value, d/dtheta and (d/dphi)/sin(theta) at a (not necessarily unit)
// direction, from a cell's coefficients q (SH polynomial form or DTI).
template <int MODEL, int LMAX>
__device__ __forceinline__ ModelOut model_at(const float *q, V3 c, const Direction &dir)
{
    if constexpr (MODEL == MODEL_SH)
        return sh_model<LMAX>(q, dir);
    else
        return dti_model(q, c, dir);
}

// The SH value alone at a unit direction u: f = A_0(z) + sum_m A_m(z) Re u^m
// + B_m(z) Im u^m with u = x + i y (= sin(theta) e^{i phi}), so no angles,
// no normalisation and no derivative terms.
template <int N>
__device__ __forceinline__ float horner_value(const float *c, float w)
{
    float v = c[N - 1];
#pragma unroll
    for (int j = N - 2; j >= 0; j--)
        v = v * w + c[j];
    return v;
}

template <int L, int M>
__device__ __forceinline__ void sh_value_terms(const float *q, float z, float w, float u1r, float u1i, float &ur,
                                               float &ui, float &f)
{
    if constexpr (M <= L) {
        using O = Order<L>;
        const float nur = ur * u1r - ui * u1i;
        ui = ur * u1i + ui * u1r;
        ur = nur;
        float a = horner_value<O::poly_terms(M)>(q + O::a_offset(M), w);
        float b = horner_value<O::poly_terms(M)>(q + O::b_offset(M), w);
        if constexpr ((M & 1) == 1) {
            a *= z;
            b *= z;
        }
        f += a * ur + b * ui;
        sh_value_terms<L, M + 1>(q, z, w, u1r, u1i, ur, ui, f);
    }
}

template <int L>
__device__ __forceinline__ float sh_value_unit(const float *q, V3 u)
{
    using O = Order<L>;
    const float w = u.z * u.z;
    float f = horner_value<O::poly_terms(0)>(q + O::a_offset(0), w);
    float ur = 1.0f, ui = 0.0f;
    sh_value_terms<L, 1>(q, u.z, w, u.x, u.y, ur, ui, f);
    return f;
}

// The model value alone at a direction of unit length up to rounding (cap
// candidates, exps2 results): the reference normalises by the radius, which
// here differs from 1 by ~1e-7.
template <int MODEL, int LMAX>
__device__ __forceinline__ float value_at(const float *q, V3 c)
{
    if constexpr (MODEL == MODEL_SH) {
        const float raw = sh_value_unit<LMAX>(q, c);
        if (raw * 100.0f >= 30.0f)
            return raw;
        return __builtin_amdgcn_logf(1.0f + __builtin_amdgcn_exp2f(raw * 144.26950408889634f)) *
               0.006931471805599453f;
    } else {
        return model_at<MODEL, LMAX>(q, c, direction_of(c)).value;
    }
}

__device__ __forceinline__ V3 unit_or_z(V3 o)
{
    const float n2 = o.x * o.x + o.y * o.y + o.z * o.z;
    if (!(__builtin_sqrtf(n2) > 1e-8f))
        return v3(0.0f, 0.0f, 1.0f);
    const float inv = __builtin_amdgcn_rsqf(n2);
    return v3(o.x * inv, o.y * inv, o.z * inv);
}

// algorithm.py:305-376 (Probabilistic), the plain path's weighted choice
// (choose_candidate) on the polynomial model: candidates are Cartesian
// already, so a candidate costs one polynomial and one softplus (hardware
// exp2/log2), with no cart2sph and no trigonometric basis. The azimuth
// 2 pi u is handed to the hardware sine and cosine as u revolutions.
template <int MODEL, int LMAX>
__device__ __forceinline__ V3 probabilistic(const AlgoParams &a, const float *q, V3 o_in, uint2 &rng)
{
    const V3 o = unit_or_z(o_in);
    const bool z_helper = fabsf(o.z) < 0.9f;
    // cross(helper, o) for helper e_z or e_y, then normalised.
    const V3 c = z_helper ? v3(-o.y, o.x, 0.0f) : v3(o.z, 0.0f, -o.x);
    const float ic = __builtin_amdgcn_rsqf(c.x * c.x + c.y * c.y + c.z * c.z);
    const V3 t1 = v3(c.x * ic, c.y * ic, c.z * ic);
    const V3 t2 = v3(o.y * t1.z - o.z * t1.y, o.z * t1.x - o.x * t1.z, o.x * t1.y - o.y * t1.x);
    const float one_minus_cos_max = 1.0f - tgk_cosf(a.max_angle);
    auto candidate = [&](uint2 &state) {
        const float cos_angle = 1.0f - randu(state) * one_minus_cos_max;
        const float sin_angle = __builtin_sqrtf(fmaxf(0.0f, 1.0f - cos_angle * cos_angle));
        const float turns = randu(state);
        const float ca = __builtin_amdgcn_cosf(turns), sa = __builtin_amdgcn_sinf(turns);
        return v3(cos_angle * o.x + sin_angle * (ca * t1.x + sa * t2.x),
                  cos_angle * o.y + sin_angle * (ca * t1.y + sa * t2.y),
                  cos_angle * o.z + sin_angle * (ca * t1.z + sa * t2.z));
    };
    auto weight = [&](V3 c) {
        const float v = value_at<MODEL, LMAX>(q, c);
        return __builtin_isfinite(v) ? fmaxf(v, 0.0f) : 0.0f;
    };
    V3 chosen;
    return choose_candidate(a.n_samples, rng, candidate, weight, chosen) ? chosen : o;
}

// Gradient on the sphere: dtheta e_theta + (dphi / sin theta) e_phi, from
// the angular derivatives. (Not the Cartesian form of sh_cartesian: its
// normal part, about L f, cancels to rounding that the central differences
// of the Hessian divide by 2h.)
template <int MODEL, int LMAX>
__device__ __forceinline__ V3 sphere_gradient(const float *q, V3 p)
{
    #pragma clang fp contract(fast)
    const Direction d = direction_of(p);
    const ModelOut m = model_at<MODEL, LMAX>(q, p, d);
    return v3(m.dtheta * (d.z * d.c1) - m.dphi * d.s1, m.dtheta * (d.z * d.s1) + m.dphi * d.c1, -m.dtheta * d.s);
}

// algorithm.py:515-598 (Deterministic): the plain path's Newton ascent with
// the polynomial model, hardware rsqrt frames and the Taylor exps2.
template <int MODEL, int LMAX>
__device__ __forceinline__ V3 deterministic(const AlgoParams &a, const float *q, V3 o_in)
{
    #pragma clang fp contract(fast)
    const float h = 1e-2f, inv_2h = 50.0f;
    const float n2 = o_in.x * o_in.x + o_in.y * o_in.y + o_in.z * o_in.z;
    const float inv = __builtin_amdgcn_rsqf(n2);
    V3 o = v3(o_in.x * inv, o_in.y * inv, o_in.z * inv);
    for (int iteration = 0; iteration < a.max_iterations; iteration++) {
        const Direction d = direction_of(o);
        const V3 e1 = v3(d.z * d.c1, d.z * d.s1, -d.s);
        const V3 e2 = v3(-d.s1, d.c1, 0.0f);
        const ModelOut current = model_at<MODEL, LMAX>(q, o, d);
        const float f0 = current.value, g1 = current.dtheta, g2 = current.dphi;
        const float gradient_norm = __builtin_sqrtf(g1 * g1 + g2 * g2);
        if (gradient_norm <= a.tolerance)
            break;
        const V3 s1 = v3(h * e1.x, h * e1.y, h * e1.z), s2 = v3(h * e2.x, h * e2.y, 0.0f);
        const V3 gp1 = sphere_gradient<MODEL, LMAX>(q, exps2(o, s1));
        const V3 gm1 = sphere_gradient<MODEL, LMAX>(q, exps2(o, v3(-s1.x, -s1.y, -s1.z)));
        const V3 gp2 = sphere_gradient<MODEL, LMAX>(q, exps2(o, s2));
        const V3 gm2 = sphere_gradient<MODEL, LMAX>(q, exps2(o, v3(-s2.x, -s2.y, 0.0f)));
        const V3 dg1 = v3(gp1.x - gm1.x, gp1.y - gm1.y, gp1.z - gm1.z);
        const V3 dg2 = v3(gp2.x - gm2.x, gp2.y - gm2.y, gp2.z - gm2.z);
        const float h11 = (dg1.x * e1.x + dg1.y * e1.y + dg1.z * e1.z) * inv_2h;
        const float h22 = (dg2.x * e2.x + dg2.y * e2.y) * inv_2h;
        const float h12 = 0.5f * ((dg2.x * e1.x + dg2.y * e1.y + dg2.z * e1.z) + (dg1.x * e2.x + dg1.y * e2.y)) *
                          inv_2h;
        const float determinant = h11 * h22 - h12 * h12;
        float d1, d2;
        if (h11 < -1e-7f && h22 < -1e-7f && determinant > 1e-12f) {
            const float r = __builtin_amdgcn_rcpf(determinant);
            d1 = (-h22 * g1 + h12 * g2) * r;
            d2 = (h12 * g1 - h11 * g2) * r;
        } else {
            const float r = 0.1f * __builtin_amdgcn_rcpf(gradient_norm);
            d1 = g1 * r;
            d2 = g2 * r;
        }
        const float step_norm = __builtin_sqrtf(d1 * d1 + d2 * d2);
        if (step_norm > a.max_step_size) {
            const float scale = a.max_step_size / step_norm;
            d1 *= scale;
            d2 *= scale;
        }
        if (__builtin_sqrtf(d1 * d1 + d2 * d2) <= 1e-7f)
            break;
        const V3 tangent = v3(d1 * e1.x + d2 * e2.x, d1 * e1.y + d2 * e2.y, d1 * e1.z);
        bool accepted = false;
        float scale = 1.0f;
        for (int search = 0; search < 8; search++) {
            const V3 c = exps2(o, v3(scale * tangent.x, scale * tangent.y, scale * tangent.z));
            if (value_at<MODEL, LMAX>(q, c) > f0 + 1e-7f) {
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

// The reference's 45-term real symmetric SH basis (order 8, MRtrix order:
// index l(l+1)/2 + m for even l; spharm.cl:10-207, core.py:36-69) at the
// direction of o, without angles. With z = cos(theta) and u = (x + iy) / r =
// sin(theta) e^{i phi}, Y_{l,0} = L_l^0(z), Y_{l,+-m} = sqrt(2) L_l^m(z)
// (Re, Im) u^m, where L_l^m = N_lm P_l^m / sin^m(theta) (N_lm the
// normalisation, P_l^m with the Condon-Shortley phase) is a polynomial in z
// of degree l - m, run up in l for each m by the normalised recurrence
//   L_m^m = N_mm (-1)^m (2m-1)!!,
//   L_l^m = alpha_lm z L_{l-1}^m - beta_lm L_{l-2}^m,
//   alpha_lm = (2l-1)/(l-m) N_lm/N_{l-1,m}, beta_lm = (l+m-1)/(l-m) N_lm/N_{l-2,m},
// from P_l^m = ((2l-1) z P_{l-1}^m - (l+m-1) P_{l-2}^m) / (l-m). Every value
// is scaled by w; the zero vector maps to the pole, as cart2sph does.
// visit(index, value) receives each of the 45 values.
namespace basis {

constexpr double csqrt(double x)
{
    if (!(x > 0.0))
        return 0.0;
    double r = x > 1.0 ? x : 1.0;
    for (int i = 0; i < 200; i++)
        r = 0.5 * (r + x / r);
    return r;
}
constexpr double factorial(int n)
{
    double f = 1.0;
    for (int i = 2; i <= n; i++)
        f *= i;
    return f;
}
constexpr double norm(int l, int m)
{
    return csqrt((2.0 * l + 1.0) / (4.0 * 3.14159265358979323846) * factorial(l - m) / factorial(l + m));
}
// L_m^m, with the sqrt(2) of the m > 0 harmonics folded in.
constexpr double start(int m)
{
    double dfact = 1.0;
    for (int k = 2 * m - 1; k > 1; k -= 2)
        dfact *= k;
    return (m > 0 ? csqrt(2.0) : 1.0) * norm(m, m) * ((m & 1) ? -dfact : dfact);
}
constexpr double alpha(int l, int m) { return (2.0 * l - 1.0) / (l - m) * norm(l, m) / norm(l - 1, m); }
constexpr double beta(int l, int m) { return l - m >= 2 ? (l + m - 1.0) / (l - m) * norm(l, m) / norm(l - 2, m) : 0.0; }

template <int M, int L, class Visit>
__device__ __forceinline__ void order_step(float z, float ur, float ui, float &prev2, float &prev, Visit &visit)
{
    if constexpr (L <= 8) {
        constexpr float a = (float)alpha(L, M);
        constexpr float b = (float)beta(L, M);
        const float cur = L - M >= 2 ? a * z * prev - b * prev2 : a * z * prev;
        prev2 = prev;
        prev = cur;
        if constexpr ((L & 1) == 0) {
            constexpr int centre = L * (L + 1) / 2;
            if constexpr (M == 0) {
                visit(centre, cur);
            } else {
                visit(centre + M, cur * ur);
                visit(centre - M, cur * ui);
            }
        }
        order_step<M, L + 1>(z, ur, ui, prev2, prev, visit);
    }
}

template <int M, class Visit>
__device__ __forceinline__ void orders(float z, float ux, float uy, float w, float ur, float ui, Visit &visit)
{
    if constexpr (M <= 8) {
        constexpr float lmm = (float)start(M);
        float prev2 = 0.0f, prev = w * lmm;
        if constexpr ((M & 1) == 0) {
            constexpr int centre = M * (M + 1) / 2;
            if constexpr (M == 0) {
                visit(centre, prev);
            } else {
                visit(centre + M, prev * ur);
                visit(centre - M, prev * ui);
            }
        }
        order_step<M, M + 1>(z, ur, ui, prev2, prev, visit);
        // u^(M+1) from u^M.
        orders<M + 1>(z, ux, uy, w, ur * ux - ui * uy, ur * uy + ui * ux, visit);
    }
}

}  // namespace basis

template <class Visit> __device__ __forceinline__ void sh45_basis(V3 o, float w, Visit &&visit)
{
    const float r2 = o.x * o.x + o.y * o.y + o.z * o.z;
    float z = 1.0f, ux = 0.0f, uy = 0.0f;
    if (r2 > 0.0f) {
        const float inv = __builtin_amdgcn_rsqf(r2);
        z = fminf(fmaxf(o.z * inv, -1.0f), 1.0f);
        ux = o.x * inv;
        uy = o.y * inv;
    }
    basis::orders<0>(z, ux, uy, w, 1.0f, 0.0f, visit);
}

}  // namespace fast
}  // namespace tgk
