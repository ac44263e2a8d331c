// The guided vertex of the path integrator with learned guiding (no counterpart in
// the reference): the vMF-mixture proposal the guiding network outputs, the
// continuation drawn from the one-sample MIS density p_G = alpha q + (1 - alpha)
// p_bsdf, next-event estimation against it, and the training records the
// continuation directions become.  Shared by the pass schedule (tracer.hip) and
// the batched persistent-slot schedule (batch.hip), which differ only in where a
// vertex's surface record and network outputs are stored.
#pragma once

#include "guiding.hpp"
#include "path.hpp"
#include "shading.hpp"

namespace rbw {

// the vMF-mixture head of the guiding MLP (same parameterisation as the training loss),
// evaluated with the hardware exponential, logarithm, sine, cosine, reciprocal and square
// root (v_exp_f32, v_log_f32, v_sin_f32, v_cos_f32, v_rcp_f32, v_sqrt_f32; a few ulp): the
// sampler and the pdf share the lobes' parameters, and the estimator divides by the
// density it evaluates, so the rounding changes the variance a little, not the mean
__device__ __forceinline__ float vrcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float vsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
struct vmf_mixture {
    f3 mu[guide::N_LOBES];
    float kappa[guide::N_LOBES], w[guide::N_LOBES], norm[guide::N_LOBES];
    __device__ void load(const float* o) {
        float amax = -3.4e38f;
        for (int j = 0; j < guide::N_LOBES; j++) amax = fmaxf(amax, o[5 * j + 4]);
        float ws = 0.0f;
        for (int j = 0; j < guide::N_LOBES; j++) {
            w[j] = __expf(o[5 * j + 4] - amax);
            ws += w[j];
            const f3 m = f3{o[5 * j], o[5 * j + 1], o[5 * j + 2]};
            const float inm = vrcp(fmaxf(vsqrt(m.x * m.x + m.y * m.y + m.z * m.z), 1e-12f));
            mu[j] = f3{m.x * inm, m.y * inm, m.z * inm};
            const float x = o[5 * j + 3];
            const float sp = x > 20.0f ? x : __logf(1.0f + __expf(x));
            kappa[j] = fminf(sp + guide::KAPPA_MIN, guide::KAPPA_MAX);
            // 1 - e^(-2 kappa) in FP32 keeps ~5e-6 relative accuracy down to kappa = KAPPA_MIN
            norm[j] = kappa[j] * vrcp(2.0f * 3.14159265f * (1.0f - __expf(-2.0f * kappa[j])));
        }
        const float iws = vrcp(ws);
        for (int j = 0; j < guide::N_LOBES; j++) w[j] *= iws;
    }
    __device__ float pdf(f3 d) const {
        float q = 0.0f;
        for (int j = 0; j < guide::N_LOBES; j++)
            q += w[j] * norm[j] * __expf(kappa[j] * ((mu[j].x * d.x + mu[j].y * d.y + mu[j].z * d.z) - 1.0f));
        return q;
    }
    // pick a lobe with u.x, then Jakob's vMF sampler around its mean
    __device__ f3 sample(f2 u) const {
        int j = 0;
        float acc = 0.0f;
        for (; j < guide::N_LOBES - 1; j++) {
            if (u.x < acc + w[j]) break;
            acc += w[j];
        }
        const float ux = fminf(fmaxf((u.x - acc) * vrcp(fmaxf(w[j], 1e-20f)), 0.0f), 0.99999994f);
        const float k = kappa[j];
        const float W = 1.0f + __logf(ux + (1.0f - ux) * __expf(-2.0f * k)) * vrcp(k);
        const float sr = vsqrt(fmaxf(0.0f, 1.0f - W * W));
        const float phi = 2.0f * 3.14159265f * u.y;
        const frame fr = frame_from_normal(mu[j]);
        return local_to_world(fr, f3{sr * __cosf(phi), sr * __sinf(phi), W});
    }
};

// sample from the guided technique: the vMF mixture with probability alpha, else
// the BSDF; returns false when no direction results
__device__ __forceinline__ bool guided_sample(const bsdf& fn, const surf& si, const vmf_mixture& mix, float alpha,
                                              float sel, f2 u, f3& wi) {
    if (sel < alpha) {
        wi = mix.sample(u);
        return true;
    }
    f3 f0;
    float p0;
    return fn.sample(si, u, wi, f0, p0);
}

// the network query of a surface: position in the scene box, shading normal and wo mapped to [0,1]
__device__ __forceinline__ void guide_query(const surf& si, f3 smin, f3 sinv, float q[9]) {
    const f3 pn = (si.point - smin) * sinv;
    q[0] = gclamp(pn.x, 0.0f, 1.0f);
    q[1] = gclamp(pn.y, 0.0f, 1.0f);
    q[2] = gclamp(pn.z, 0.0f, 1.0f);
    q[3] = 0.5f * (si.sh.z.x + 1.0f);
    q[4] = 0.5f * (si.sh.z.y + 1.0f);
    q[5] = 0.5f * (si.sh.z.z + 1.0f);
    q[6] = 0.5f * (si.wo.x + 1.0f);
    q[7] = 0.5f * (si.wo.y + 1.0f);
    q[8] = 0.5f * (si.wo.z + 1.0f);
}

// Guided vertex: next-event estimation whose second technique is the path's own
// continuation, drawn from the guided density p_G = alpha q + (1 - alpha) p_bsdf.
// The light sample is weighted by the power heuristic against p_G (the structure
// of integrator.cpp:101-161 with the BSDF technique replaced); the continuation
// ray, when it reaches an emitter, adds the emission weighted against the light
// pdf (shade_path, bounce > 0).  At the last vertex there is no continuation and
// the light sample takes weight 1.  Russian roulette as path_integrator.cpp:35-43.
// The continuation direction becomes a training record (query row `qrow`), closed
// at the next vertex: its target is everything the continuation brings back
// (emission weighted by its MIS weight, plus all later bounces), so the guide
// learns the radiance that light sampling does not already cover.  `mix` holds the
// network's raw outputs for this vertex; light_tag is handed to the connect stage
// when light-sample training records are on.
__device__ __forceinline__ void guided_vertex(const dscene& S, const dpaths& P, const dparams& prm, uint32_t slot,
                                              const surf& si, uint32_t mat, const float* mix_row,
                                              const float* qrow, uint32_t light_tag, bool& alive,
                                              bool& need_connect) {
    const dguide& G = prm.g;
    const bsdf fn = make_bsdf(S, mat, RBW_NONE);
    vmf_mixture mix;
    mix.load(mix_row);
    sampler_state smp = load_sampler(P, prm, slot);
    int bounce = (int)P.U(U_BOUNCE, slot);
    f3 beta = P.F3(F_BX, slot);
    const f3 wo_s = world_to_local(si.sh, si.wo);
    const float a = G.alpha;
    const bool last = (uint32_t)bounce + 1 >= prm.max_depth;
    auto pdf_g = [&](f3 w, f3 w_s) { return a * mix.pdf(w) + (1.0f - a) * fn.pdf(wo_s, w_s); };
    // light sample
    const uint32_t ne = S.num_emitters;
    const float u1 = smp.next();
    uint32_t which = (uint32_t)floorf(u1 * (float)ne);
    if (which > ne - 1) which = ne - 1;
    const float pick = 1.0f / (float)ne;
    const uint32_t lid = S.emitters[which];
    const dentity L = S.ents[lid];
    light_sample ls = sample_light(S, L, si.point, si.normal, smp.next2());
    ls.pdf = ls.pdf * pick;
    if (!is_black(ls.L) && ls.pdf > 0.0f) {
        const f3 wi = world_to_local(si.sh, ls.wi);
        f3 f = fn.evaluate(wo_s, wi);
        f = f * fabsf(dot(ls.wi, si.sh.z));
        if (!is_black(f)) {
            f3 so, sd;
            float slen;
            spawn_ray_to(si.point, si.normal, ls.point, so, sd, slen);
            const float w = last ? 1.0f : power_heuristic(ls.pdf, pdf_g(ls.wi, wi));
            P.setF3(F_SOX, slot, so);
            P.setF3(F_SDX, slot, sd);
            P.F(F_SLEN, slot) = slen;
            P.setF3(F_SFX, slot, f * ls.L * w / ls.pdf);
            if (G.train && G.light_records) {
                // training record of the light sample: its MIS-weighted radiance / pdf
                const float fl = G.product ? 0.2126f * f.x + 0.7152f * f.y + 0.0722f * f.z : 1.0f;
                P.F(F_MFX, slot) = fl * w * (0.2126f * ls.L.x + 0.7152f * ls.L.y + 0.0722f * ls.L.z) / ls.pdf;
                P.F(F_MFY, slot) = __uint_as_float(light_tag);
            }
            P.U(U_SENT, slot) = lid;
            P.setF3(F_CBX, slot, beta);
            P.U(U_CFLAGS, slot) = C_SHADOW | ((uint32_t)bounce << 8);
            need_connect = true;
        }
    }
    // continuation from the guided technique
    bool cont = !last;
    f3 wi = f3{0.0f, 0.0f, 0.0f};
    float pdf = 0.0f, rec_pdf = 0.0f;
    if (cont) {
        const float sel = smp.next();
        const f2 u = smp.next2();
        cont = guided_sample(fn, si, mix, a, sel, u, wi);
    }
    if (cont) {
        const f3 wi_s = world_to_local(si.sh, wi);
        const f3 f = fn.evaluate(wo_s, wi_s);
        pdf = pdf_g(wi, wi_s);
        if (is_black(f) || !(pdf > 0.0f)) {
            cont = false;
        } else {
            const f3 fc = f * fabsf(dot(wi, si.sh.z));
            // the record's weight is L_i / rec_pdf: with product targets rec_pdf carries 1 / (f |cos|)
            rec_pdf = G.product ? pdf / fmaxf(0.2126f * fc.x + 0.7152f * fc.y + 0.0722f * fc.z, 1e-30f) : pdf;
            beta = beta * (fc / pdf);
            f3 no, nd;
            spawn_ray(si.point, si.normal, wi, no, nd);
            P.setF3(F_OX, slot, no);
            P.setF3(F_DX, slot, nd);
            // what the next vertex needs to weight emission it finds against light sampling
            P.setF3(F_MPX, slot, si.point);
            P.setF3(F_MNX, slot, si.normal);
            P.setF3(F_MWX, slot, wi);
            P.F(F_MPDF, slot) = pdf;
        }
    }
    if (cont) {
        const float mc = spectrum_max(beta * 1.0f);
        if (mc < prm.rr_threshold && bounce > 3) {
            const float qq = gmax(0.05f, 1.0f - mc);
            if (smp.next() < qq) cont = false;
            else beta = beta / (1.0f - qq);
        }
    }
    const int b0 = bounce;
    bounce += 1;
    alive = cont;
    if (alive && G.train) {
        const size_t ri = (size_t)slot * G.rec_depth + (uint32_t)b0;
        float4* r = reinterpret_cast<float4*>(G.rec + ri * R_STRIDE);
        r[0] = make_float4(qrow[0], qrow[1], qrow[2], qrow[3]);
        r[1] = make_float4(qrow[4], qrow[5], qrow[6], qrow[7]);
        r[2] = make_float4(qrow[8], wi.x, wi.y, wi.z);
        r[3] = make_float4(rec_pdf, beta.x, beta.y, beta.z);
        G.rec_flag[ri] = 1;
        P.U(U_PENDING, slot) = (uint32_t)b0;
    }
    P.setF3(F_BX, slot, beta);
    P.U(U_BOUNCE, slot) = (uint32_t)bounce;
    store_sampler(P, slot, smp);
}

// The training weight of a closed record (b the vertex's bounce) of the path in `slot`
// once the path has ended: the luminance of the radiance the record's continuation
// brought back (final value minus the value at the next vertex's start, over the
// throughput after the vertex) over the continuation's pdf; 0 when not usable.
__device__ __forceinline__ float record_target(const dpaths& P, const dguide& G, uint32_t slot, size_t ri) {
    const f3 vf = P.F3(F_VX, slot);
    const float4 pb = *reinterpret_cast<const float4*>(G.rec + ri * R_STRIDE + R_PDF);
    const float4 mm = *reinterpret_cast<const float4*>(G.rec + ri * R_STRIDE + R_MX);
    const f3 m = f3{mm.x, mm.y, mm.z};
    const f3 b = f3{pb.y, pb.z, pb.w};
    const f3 d = vf - m;
    const float Lx = b.x > 0.0f ? d.x / b.x : 0.0f, Ly = b.y > 0.0f ? d.y / b.y : 0.0f,
                Lz = b.z > 0.0f ? d.z / b.z : 0.0f;
    const float lum = 0.2126f * Lx + 0.7152f * Ly + 0.0722f * Lz;
    const float pdf = pb.x;
    const float c = pdf > 0.0f ? lum / pdf : 0.0f;
    return c > 0.0f && isfinite(c) ? c : 0.0f;
}

// copies a record's query and direction with its weight c into training row k
__device__ __forceinline__ void record_to_row(const dpaths& P, const dguide& G, size_t ri, float c, float* t_in,
                                              float* t_dir, float* t_c, size_t k) {
    const float4* r = reinterpret_cast<const float4*>(G.rec + ri * R_STRIDE);
    const float4 r0 = r[0], r1 = r[1], r2 = r[2];
    const float q[9] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x};
    for (int j = 0; j < 9; j++) t_in[k * 9 + j] = q[j];
    t_dir[k * 3 + 0] = r2.y;
    t_dir[k * 3 + 1] = r2.z;
    t_dir[k * 3 + 2] = r2.w;
    t_c[k] = c;
}

}  // namespace rbw
