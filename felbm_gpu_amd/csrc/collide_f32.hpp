// collide_f32.hpp -- the FP32 form of the fused collision (streamed by gather).
//
// The same physics as k_collide (kernels.hpp), re-associated where FP32
// allows it (the FP32 engine is held to 1e-5 against the FP64 oracle, not to
// bit identity):
//  - central differences are antisymmetric in a direction pair (m, opp m), so
//    9 numerators serve 18 directions and the vector gradients sum 9 pairs;
//  - eq_h + F_h collapses to G_k (K0 + a2 bd_c + a4 bd_p + a5 e_k.bf + mob avg_k):
//    the central-difference terms of eq_h (device_engine.cuh:586-588) and F_h
//    (:598) cancel exactly;
//  - the MRT transform (:617-638) splits each pair into even (y_k + y_opp) and
//    odd (y_k - y_opp) parts, and Minv = M^T diag(1/|row|^2) folds into the
//    relaxation rates;
//  - divisions use the hardware reciprocal;
//  - g and p are deviations from the rest pressure (g_k - w_k, p - 1; see
//    p_ref in kernels.hpp): eq_g and F_g are affine in (g, p) with the w_k
//    offset cancelling, so the arithmetic below yields the stored deviation.
#pragma once
#include "kernels.hpp"

namespace fm {

constexpr bool EVEN_ROW[19] = { true, true, true, false, false, false, false, false, false,
                                true, true, true, true, true, true, true, false, false, false };
constexpr bool CONS_ROW[19] = { true, false, false, true, false, true, false, true, false,
                                false, false, false, false, false, false, false, false, false, false };

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

template <bool MRT, bool HBB, bool OPEN>
__device__ __forceinline__ void collide_f32_body(const Inst<float> *__restrict__ D, const int2 bm, unsigned T, int stats)
{
  if (bm.x < 0) return;
  const Inst<float> &d = D[bm.x];
  const unsigned cur = (T ^ (unsigned)d.par0) & 1u;
  const int i = bm.y + (int)threadIdx.x;
  const bool live = i < d.n;
  Fx s0 = {0,0,0}, s1 = {0,0,0}, s2 = {0,0,0}, s3 = {0,0,0}, s4 = {0,0,0};
  unsigned long long vmax = 0;
  // this step's mass-correction lambda, which k_mass_lambda left in the
  // descriptor: it arrives with the descriptor's scalar loads
  float lam = 0.0f;
  if (d.mass_on) {
    lam = d.lam_on ? (float)d.lam : 0.0f;
  }
  if (threadIdx.x == 0 && bm.y == 0) {
    if (d.mass_on) stg(d.acc, ACC_FLAGS, 1ull);
    if (OPEN && d.obc) stg(d.acc, ACC_FLAGS + 1, ldg(d.acc, ACC_FLAGS + 1) + 1ull);   // ++open_step
    if (d.trk) stg(d.acc, ACC_FLAGS + 2, ldg(d.acc, ACC_FLAGS + 2) + 1ull);            // ++tracer step
  }
  if (live) {
    constexpr unsigned RB = 4;
    const unsigned nsb = (unsigned)d.ns*RB, io = (unsigned)i*RB;
    const rsrc_t rF = make_rsrc(d.craw, (unsigned)F_COUNT*nsb), rW = make_rsrc(d.words, d.nwords*8u);
    const rsrc_t rGi = make_rsrc(cur ? d.g1 : d.g0, 19u*nsb), rGo = make_rsrc(cur ? d.g0 : d.g1, 19u*nsb);
    const rsrc_t rH = make_rsrc(d.h, 19u*nsb);
    int x, y, z; decode(d, i, x, y, z);
    Nbhd nb; make_nbhd(d, x, y, z, nb);
    unsigned vo[Q], fl, so, fl2;
    ring_f<float, OPEN, 1, false>(d, rW, nb, io, vo, fl, so);
    // The streamed g, read with the non-temporal hint: its lines are not read
    // again this step, and evicting them first keeps the neighbour records of
    // the +-2 planes in L2 (0.826 -> 0.766 ms).  Issued as soon as the ring-1
    // offsets are known, so the HBM latency of the gather overlaps the ring-2
    // row windows (0.724 -> 0.720 ms); issuing it together with the record
    // loads instead cost 10%.
    float gk[Q];
    #pragma unroll
    for (int k = 0; k < Q; ++k) gk[k] = bld_nt<float>(rGi, pull_off(k, vo, fl, io, nsb), pull_soff(k, nsb));
    // open-boundary node; a buffer node does not collide and stores the
    // equilibrium the reference writes into it (k_open_bnd_buf)
    int eb = -1;
    const float *ovp = nullptr;
    if constexpr (OPEN) {
      if (d.obc) { eb = ldg(d.ent, i); if (eb >= 0) ovp = d.ov + 8*(long)(eb >> 4); }
    }
    const bool nostream = eb >= 0 && (eb & ENT_NOSTREAM);
    if (nostream) {
      const float oc = ldg(ovp, 0), orho = ldg(ovp, 1), op = ldg(ovp, 2);
      const float oux = ldg(ovp, 3), ouy = ldg(ovp, 4), ouz = ldg(ovp, 5);
      #pragma unroll
      for (int k = 0; k < Q; ++k) {
        float hv, gv;
        plain_eq<float>(k, oux, ouy, ouz, op, orho, oc, hv, gv);     // op = p - 1: gv = g_eq - w_k
        store_post<float, HBB, OPEN>(hv, rH, k, vo[k], (fl >> k) & 1u, (so >> k) & 1u, io, nsb);
        store_post<float, HBB, OPEN>(g_flush(gv), rGo, k, vo[k], (fl >> k) & 1u, (so >> k) & 1u, io, nsb, -W<float>(k));
      }
      if (stats) {
        double vx = (double)oux, vy = (double)ouy, vz = (double)ouz;
        double v2 = vx*vx + vy*vy + vz*vz;
        fx_add(s0, vx); fx_add(s1, vy); fx_add(s2, vz); fx_add(s3, v2); fx_add(s4, (double)oc);
        vmax = (unsigned long long)__double_as_longlong(v2);
      }
    } else {
      float cN[Q], pN[Q], c2[Q], p2[Q], lN[Q];
      const Rec4<float> r0 = bld_r3(rF, io, F_REC*nsb);
      const float mu_own = bld_nt<float>(rF, io, F_MU*nsb);
      const float craw = bld_nt<float>(rF, io, F_CRAW*nsb);      // read once
      const float clo = bld_nt<float>(rF, io, F_CLO*nsb);        // Sigma h - c_raw (moments pass)
      {
        // a neighbour that is not a site takes the own record's values
        unsigned vo2[Q];
        { unsigned so2; ring_f<float, OPEN, 2, false>(d, rW, nb, io, vo2, fl2, so2); }
        #pragma unroll
        for (int m = 1; m < Q; ++m) {
          const Rec4<float> r1 = bld_r3(rF, vo[m], F_REC*nsb);
          const Rec2<float> r2 = bld_r2(rF, vo2[m], F_REC*nsb);
          const bool s1 = (fl >> m) & 1u, s2 = (fl2 >> m) & 1u;
          cN[m] = s1 ? r1.c : r0.c; pN[m] = s1 ? r1.p : r0.p; lN[m] = s1 ? r1.l : r0.l;
          c2[m] = s2 ? r2.c : r0.c; p2[m] = s2 ? r2.p : r0.p;
        }
      }
      const float c = r0.c, p = r0.p, mu = mu_own, lmu = r0.l;
      // The mass correction c = c_raw - lambda phi+ (applied by the c pass) is
      // an FP32 value, and near the interface the correction per step (1e-7 to
      // 1e-6) is carried only to half an ulp of c (3e-8 at c = 0.5): the same
      // rounding at the same sites every step, 2e-4 in c after 1000 steps at a
      // 139-degree contact angle.  What it dropped ((c_raw - c) is exact) joins
      // the small terms of h, so the populations carry the correction, and
      // stream it, as the FP64 state does.  0 when no correction applies.
      const float phi_raw = craw*(1.0f - craw);        // as corrected() forms it
      // with the moments pass's remainder: Sigma h of the next step starts from
      // the exact sum of this step's populations, corrected, not from the FP32 c
      // (open-boundary nodes carry none: their c_raw may hold an imposed value
      // or a mirrored vertex's, which the remainder does not belong to)
      const float cres = (craw - c) - lam*(phi_raw > 0.0f ? phi_raw : 0.0f) + (eb >= 0 ? 0.0f : clo);
      const float drho = d.drho;
      const float CS2 = 1.0f/3.0f;
      const float rho = ovp ? ldg(ovp, 1) : d.rho1 + drho*c;
      const float irho = frcp(rho);
      const float crl = ovp ? ldg(ovp, 6) : c;     // relax from the raw c
      const float cp = fminf(fmaxf(crl, 0.0f), 1.0f);
      const float tau = frcp(cp*frcp(d.tau0) + (1.0f - cp)*frcp(d.tau1));
      const float relax = frcp(tau + 0.5f);
      const float c_over = 3.0f*c*irho;
      const float a2 = 0.5f*(1.0f + c_over*mu), a4 = -0.5f*c_over, a5 = 0.5f*c_over;
      const float hmob = 0.5f*d.mobility;
      // per direction, once the stencils are in: the h-side terms
      // Ph_k = a2 bd_c + a4 bd_p + mob/2 lap(mu)(x+e_k), and the g-side
      // central (cd_c) and combined (cd_c + bd_c) differences of c
      float Ph[Q], cdc[Q], dc2[Q];
      float e_dcx = 0, e_dcy = 0, e_dcz = 0, e_dpx = 0, e_dpy = 0, e_dpz = 0;
      float b_cx = 0, b_cy = 0, b_cz = 0, b_px = 0, b_py = 0, b_pz = 0;
      Ph[0] = hmob*lmu; cdc[0] = 0.0f; dc2[0] = 0.0f;
      #pragma unroll
      for (int m = 1; m < Q; m += 2) {
        const int o = m + 1;
        const bool hin = (fl >> m) & 1u, hip = (fl >> o) & 1u, h2 = (fl2 >> m) & 1u, hp2 = (fl2 >> o) & 1u;
        bool use;
        const float nc = cd_num<float, HBB>(hin, hip, cN[m], cN[o], use);
        const float np = cd_num<float, HBB>(hin, hip, pN[m], pN[o], use);
        const float c3w = 3.0f*W<float>(m);
        acc_vec<float>(m, use ? c3w*nc : 0.0f, e_dcx, e_dcy, e_dcz);
        acc_vec<float>(m, use ? c3w*np : 0.0f, e_dpx, e_dpy, e_dpz);
        cdc[m] = use ? 0.5f*nc : 0.0f; cdc[o] = -cdc[m];
        const float bcm = bd_num<float, HBB>(hin, h2, hip, hp2, cN[m], c2[m], cN[o], c2[o], c);
        const float bpm = bd_num<float, HBB>(hin, h2, hip, hp2, pN[m], p2[m], pN[o], p2[o], p);
        const float bco = bd_num<float, HBB>(hip, hp2, hin, h2, cN[o], c2[o], cN[m], c2[m], c);
        const float bpo = bd_num<float, HBB>(hip, hp2, hin, h2, pN[o], p2[o], pN[m], p2[m], p);
        acc_vec<float>(m, c3w*(bcm - bco), b_cx, b_cy, b_cz);
        acc_vec<float>(m, c3w*(bpm - bpo), b_px, b_py, b_pz);
        Ph[m] = a2*bcm + a4*bpm + hmob*lN[m];
        Ph[o] = a2*bco + a4*bpo + hmob*lN[o];
        dc2[m] = cdc[m] + bcm; dc2[o] = cdc[o] + bco;
      }
      float jx = 0, jy = 0, jz = 0;
      #pragma unroll
      for (int k = 0; k < Q; ++k) acc_vec<float>(k, gk[k], jx, jy, jz);
      const float f_dcx = 0.5f*(e_dcx + b_cx), f_dcy = 0.5f*(e_dcy + b_cy), f_dcz = 0.5f*(e_dcz + b_cz);
      const float f_dpx = 0.5f*(e_dpx + b_px), f_dpy = 0.5f*(e_dpy + b_py), f_dpz = 0.5f*(e_dpz + b_pz);
      const float inv = 3.0f*irho;
      const float mor = mu*irho;
      const float ffr = d.ff*irho;
      float ux = jx*inv + 0.5f*(d.gx + d.fx*ffr + mor*e_dcx);
      float uy = jy*inv + 0.5f*(d.gy + d.fy*ffr + mor*e_dcy);
      float uz = jz*inv + 0.5f*(d.gz + d.fz*ffr + mor*e_dcz);
      if (eb >= 0 && (eb & ENT_BCPROC)) { ux = ldg(ovp, 3); uy = ldg(ovp, 4); uz = ldg(ovp, 5); }
      if (eb >= 0 && (eb & ENT_EQ)) {
        #pragma unroll
        for (int k = 0; k < Q; ++k) { float hv; plain_eq<float>(k, ux, uy, uz, p, rho, c, hv, gk[k]); }
      }
      if (stats) {
        double vx = (double)ux, vy = (double)uy, vz = (double)uz;
        double v2 = vx*vx + vy*vy + vz*vz;
        fx_add(s0, vx); fx_add(s1, vy); fx_add(s2, vz); fx_add(s3, v2); fx_add(s4, (double)c);
        vmax = (unsigned long long)__double_as_longlong(v2);
      }
      const float bfx = rho*d.gx + d.fx*d.ff, bfy = rho*d.gy + d.fy*d.ff, bfz = rho*d.gz + d.fz*d.ff;
      const float drho_cs2 = drho*CS2, rho_cs2 = rho*CS2;
      const float e_fgx = mu*e_dcx + bfx, e_fgy = mu*e_dcy + bfy, e_fgz = mu*e_dcz + bfz;
      const float e_fhx = e_dcx - c_over*(e_dpx - e_fgx), e_fhy = e_dcy - c_over*(e_dpy - e_fgy), e_fhz = e_dcz - c_over*(e_dpz - e_fgz);
      const float f_fgx = mu*f_dcx + bfx, f_fgy = mu*f_dcy + bfy, f_fgz = mu*f_dcz + bfz;
      const float f_fhx = f_dcx - c_over*(f_dpx - f_fgx), f_fhy = f_dcy - c_over*(f_dpy - f_fgy), f_fhz = f_dcz - c_over*(f_dpz - f_fgz);
      const float u_sq = ux*ux + uy*uy + uz*uz;
      const float u_term_h = ux*e_fhx + uy*e_fhy + uz*e_fhz;
      const float u_force_g = ux*e_fgx + uy*e_fgy + uz*e_fgz;
      const float u_drho = drho_cs2*(ux*e_dcx + uy*e_dcy + uz*e_dcz);
      const float u_fh_f = ux*f_fhx + uy*f_fhy + uz*f_fhz;
      const float u_drm = drho_cs2*(ux*f_dcx + uy*f_dcy + uz*f_dcz);
      const float u_fg_f = ux*f_fgx + uy*f_fgy + uz*f_fgz;
      // h_k = G_k (K0 + a2 bd_c + a4 bd_p + a5 e_k.bf + mob avg_k)
      //     = G_k (K0h + Ph_k + a5 e_k.bf),  K0h = K0 + mob/2 lap(mu)
      // c is kept apart from the small terms s0h = K0h - c and Ph_k: summed with
      // c first, terms below ulp(c) (6e-8 near c = 1) would round away, a bias
      // in the mass budget that the mass correction then moves onto the interface
      const float s0h = 0.5f*u_term_h - u_fh_f + hmob*lmu + cres;
      // eq_g = w p + wut (B0 - Bc cd_c) + G (u_force_g/2 +- (fp + mu cd_c)/2)
      // F_g  = wut (Bc (cd_c + bd_c) - u_drm) + G (mu/2 (cd_c + bd_c) + fp - u_fg_f)
      const float B0 = rho_cs2 + 0.5f*u_drho, Bc = 0.5f*drho_cs2, hmu = 0.5f*mu, hufg = 0.5f*u_force_g;
      const float gamma_u_sq = 1.5f*u_sq;
      auto storek = [&](float v, rsrc_t r, int k) {
        store_post<float, HBB, OPEN>(v, r, k, vo[k], (fl >> k) & 1u, (so >> k) & 1u, io, nsb);
      };
      // g slots hold g_k - w_k; a zero population is stored as -w_k
      auto storeg = [&](float v, int k) {
        store_post<float, HBB, OPEN>(g_flush(v), rGo, k, vo[k], (fl >> k) & 1u, (so >> k) & 1u, io, nsb, -W<float>(k));
      };
      float rr[Q], yv[Q];
      float relx = 0.0f;
      if constexpr (!MRT) relx = relax;
      float ukp = 0, fpp = 0;        // e_k.u and e_k.bf of the odd member of the pair
      #pragma unroll
      for (int k = 0; k < Q; ++k) {
        const float w = W<float>(k);
        // the even member k = opp(k - 1) takes the negated projections (exact)
        if (k == 0 || (k & 1)) {
          ukp = 0; fpp = 0;
          if (CX[k] > 0) { ukp += ux; fpp += bfx; } else if (CX[k] < 0) { ukp -= ux; fpp -= bfx; }
          if (CY[k] > 0) { ukp += uy; fpp += bfy; } else if (CY[k] < 0) { ukp -= uy; fpp -= bfy; }
          if (CZ[k] > 0) { ukp += uz; fpp += bfz; } else if (CZ[k] < 0) { ukp -= uz; fpp -= bfz; }
        }
        const float uk = (k == 0 || (k & 1)) ? ukp : -ukp, fp = (k == 0 || (k & 1)) ? fpp : -fpp;
        const float u_term = uk*(3.0f + 4.5f*uk) - gamma_u_sq;
        const float wut = w*u_term;
        const float G = w + wut;
        // Ph_k carries mob/2 lap(mu)(x+e_k); a non-site neighbour fetched lmu itself
        // one rounding at ulp(h_k): w c + wut c + G (s0h + Ph_k + a5 e_k.bf).
        // G = w + wut is not rounded on this path: the rest population's
        // G_0 = w_0 (1 - 1.5 u^2) differs from w_0 by a fixed fraction of an
        // ulp at every site of a given |u|, and rounding it first biased the
        // mass of every bulk site the same way (+5e-9 per site at |u| = 1e-4)
        storek(__builtin_fmaf(w, c, __builtin_fmaf(wut, c, G*(s0h + Ph[k] + a5*fp))), rH, k);
        const float cdk = cdc[k], dc2k = dc2[k];
        // BGK eq_g carries -(mu cd_c + fp)/2 (device_engine.cuh:516), MRT +(...)/2 (:598)
        const float eg = MRT ? w*p + wut*(B0 - Bc*cdk) + G*(0.5f*fp + hufg + hmu*cdk)
                             : w*p + wut*(B0 - Bc*cdk) + G*(hufg - 0.5f*fp - hmu*cdk);
        const float fg = wut*(Bc*dc2k - u_drm) + G*(hmu*dc2k + fp - u_fg_f);
        if constexpr (!MRT) {
          storeg(gk[k] + relx*(eg - gk[k]) + fg, k);
        } else {
          yv[k] = eg - gk[k];
          rr[k] = gk[k] + fg;
        }
      }
      if constexpr (MRT) {
        const float s0r = relax, taur = tau;   // 1/relax - 1/2 == tau
        const float s1r = 2.0f*taur*frcp(2.0f*d.mrt_lambda + taur);
        float sp[Q], dp[Q];
        #pragma unroll
        for (int m = 1; m < Q; m += 2) { sp[m] = yv[m] + yv[m + 1]; dp[m] = yv[m] - yv[m + 1]; }
        float xm[Q];
        #pragma unroll
        for (int m = 0; m < Q; ++m) {
          xm[m] = 0.0f;
          if (CONS_ROW[m]) continue;
          float a = 0.0f;
          if (EVEN_ROW[m] && MRTTAB.M[m][0] != 0.0) a += (float)MRTTAB.M[m][0]*yv[0];
          #pragma unroll
          for (int k = 1; k < Q; k += 2)
            if (MRTTAB.M[m][k] != 0.0) a += (float)MRTTAB.M[m][k]*(EVEN_ROW[m] ? sp[k] : dp[k]);
          const bool odd = (m == 4 || m == 6 || m == 8 || m >= 16);
          xm[m] = (odd ? s1r : s0r)*((float)(1.0/MRTTAB.N[m])*a);
        }
        {
          float a0 = 0.0f;
          #pragma unroll
          for (int m = 0; m < Q; ++m)
            if (!CONS_ROW[m] && EVEN_ROW[m] && MRTTAB.M[m][0] != 0.0) a0 += (float)MRTTAB.M[m][0]*xm[m];
          storeg(rr[0] + a0, 0);
        }
        #pragma unroll
        for (int k = 1; k < Q; k += 2) {
          float E = 0.0f, O = 0.0f;
          #pragma unroll
          for (int m = 0; m < Q; ++m) {
            if (CONS_ROW[m] || MRTTAB.M[m][k] == 0.0) continue;
            if (EVEN_ROW[m]) E += (float)MRTTAB.M[m][k]*xm[m];
            else O += (float)MRTTAB.M[m][k]*xm[m];
          }
          storeg(rr[k] + (E + O), k);
          storeg(rr[k + 1] + (E - O), k + 1);
        }
      }
    }
  }
  if (stats) {
    unsigned long long v[12] = { (unsigned long long)s0.hi, s0.lo, (unsigned long long)s1.hi, s1.lo,
                                 (unsigned long long)s2.hi, s2.lo, (unsigned long long)s3.hi, s3.lo,
                                 (unsigned long long)s4.hi, s4.lo, s0.bad + s3.bad + s4.bad, vmax };
    block_accumulate<12>(v, d.acc + ACC_STATS, NSLOT, (int)(blockIdx.x & (NSLOT - 1)), true);
  }
}

template <bool MRT, bool HBB, bool OPEN>
__global__ void __launch_bounds__(256) k_collide_f32(const Inst<float> *__restrict__ D,
                                                     const int2 *__restrict__ map, int sdi, unsigned T, int stats)
{
  const int2 bm = map_entry(map, sdi);
  if constexpr (OPEN) {
    if (bm.x >= 0 && (bm.x & MAP_INTERIOR)) { collide_f32_body<MRT, HBB, false>(D, make_int2(bm.x & ~MAP_INTERIOR, bm.y), T, stats); return; }
  }
  collide_f32_body<MRT, HBB, OPEN>(D, bm, T, stats);
}

}  // namespace fm
