// mc_kernel.h -- the Monte Carlo renderer's marches on the device: McMarch::run
// (src/mc.rs:488-977), as the reference's card batches them
// (gpu::mc_batch, src/gpu.rs:3013-3190).  Three kinds of task:
//
//   ray      CalcRay's march from a point: a random first step, the dynamic
//            fog's step count, the secant search or the bisection at the
//            surface (bin_search_it when the iteration limit stopped it),
//            the one-sided normal on the world axes with random sides;
//            a primary ray first looks whether it starts inside the set
//   shadow   CalcHSMC's ray towards a light; with `trans` the surface it
//            meets is refined, its normal taken and coloured
//   through  the way through a transparent object: from the far end back
//            in steps of at least one until the DE stop
//
// One task per lane, one f64 DE site per turn, as k_rays: every phase of the
// task (march, refinement, the normal's stencil) asks for the DE at the
// lane's point and the loop's head evaluates it.  The inside flag turns
// every DE around (invert_inside, src/calc.rs:1457-1472).  The march's
// bookkeeping is compiled without FMA contraction, in the reference's
// operation order.
#pragma once

#include "march_kernels.h"

namespace mb3d {

enum McPhase : int { MP_START, MP_PROBE_IN, MP_MARCH, MP_SEC, MP_BIS, MP_BSI, MP_NRM, MP_SHADOW, MP_SSEC, MP_THROUGH };

// GetRand (rand_of, src/mc.rs:470-476)
__device__ __forceinline__ double mc_rand(int32_t &seed) {
    seed = (int32_t)((uint32_t)seed * 0x343FDu + 0x269EC3u);
    return (double)(seed & 0x7FFFFFFF) * (1.0 / 2147483647.0);
}

// not_zero (src/mc.rs:53-63)
__device__ __forceinline__ double mc_not_zero(double d) {
    if (fabs(d) < 1e-40) return d < 0.0 ? -1e-40 : 1e-40;
    return d;
}

// surface_colour (src/mc.rs:645-662) with colour option 0: the gradient of
// the smoothed iteration count, then the orbit trap of the last DE
__device__ inline void mc_colour(const DevP &P, const Ray64 &R, float nn, bool de_limited, mb3d_hip_mc_res &o) {
#pragma clang fp contract(off)
    float g;
    if (de_limited) {
        const float lg = (float)log((double)R.ms * P.step_width);
        g = 32767.0f - (nn + P.d_col_plus + P.col_var * lg) * P.mcts_m;
    } else {
        g = 32767.0f - nn * P.mcts_m;
    }
    o.si_grad = (uint16_t)clip15(g);
    o.otrap = (uint16_t)clip15((float)(R.otrap * 4096.0));
}

constexpr uint32_t kMcGuard = 1000000u;  // march steps of one task at most

// Two waves per SIMD: the kernel's state fits 256 VGPRs with a few spills,
// and the second wave hides the f64 DE's latency (measured on the 1080p
// bulb's primary batch: 6.9 ms against 10.5 ms at one wave; three waves
// spill too much, 13.3 ms)
template <int KIND, int DEM>
__global__ __launch_bounds__(256, 2) void k_mc(DevP P, const mb3d_hip_mc_task *__restrict__ tasks, long long n,
                                            int secant, mb3d_hip_mc_res *__restrict__ res,
                                            unsigned long long *__restrict__ ctr) {
#pragma clang fp contract(off)
    Cnt cnt;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        const mb3d_hip_mc_task T = tasks[t];
        mb3d_hip_mc_res o;
        for (int k = 0; k < 3; k++) o.pos[k] = o.normals[k] = 0.0;
        o.zz = o.zz2 = o.len = 0.0;
        o.dyn_fog = o.ms_de_stop = 0.0f;
        o.kind = 0;
        o.si_grad = o.otrap = 0;
        o.flags = 0;
        if (T.kind < MB3D_HIP_MC_RAY || T.kind > MB3D_HIP_MC_THROUGH) {  // not a task: kind -1, no march
            o.kind = -1;
            o.seed = T.seed;
            res[t] = o;
            continue;
        }
        int32_t seed = T.seed;
        const bool inside = (T.flags & MB3D_HIP_MC_INSIDE) != 0;
        const bool is_ray = T.kind == MB3D_HIP_MC_RAY, is_shadow = T.kind == MB3D_HIP_MC_SHADOW;
        const double dir[3] = {T.dir[0], T.dir[1], T.dir[2]};
        auto ms_at = [&](double z) { return (float)((double)P.de_stop * (1.0 + fabs(z) * (double)P.de_stop_factor)); };
        Ray64 R;
        R.reset(P);
        for (int k = 0; k < 3; k++) R.c[k] = T.pos[k];
        double zz = T.zz, zz2 = T.zz2, dt1 = T.len;
        const double zzplus = T.zzplus;
        double dtmp = 0.0, rlast_de = 0.0, rlast_sw = 0.0, ds = 0.0, bsi_from = 0.0;
        float rsf = is_shadow ? 2.0f : 1.0f, dyn_fog = 0.0f, sca = 0.0f, nn = 0.0f;
        bool de_limited = true, first_step = true;
        int itmp = 0, nq = 0, tries = 0;
        uint32_t guard = 0;
        double no = 0.0, off = 0.0, saved = 0.0, nv[3] = {0, 0, 0};
        Refine F;
        int phase;
        if (is_ray || is_shadow) {
            R.ms = ms_at(zz);
            phase = MP_START;
        } else {
            // through (src/mc.rs:958-977): to the far end, then a random first step back
            R.ms = T.ms_de_stop;
            for (int k = 0; k < 3; k++) R.c[k] += dir[k] * dt1;
            dtmp = mc_rand(seed) * 8.0 * fmax((double)R.ms, 1.0);
            dtmp = -1.0 - fabs(dtmp);
            dt1 += dtmp;
            for (int k = 0; k < 3; k++) R.c[k] += dir[k] * dtmp;
            phase = MP_THROUGH;
        }
        for (;;) {
            R.take(de64<KIND, DEM>(P, R.c[0], R.c[1], R.c[2], R.max_it, R.calc_sit, R.ms), cnt);
            if (inside) {  // invert_inside
                const double ms = (double)R.ms, v = ms * 4.0 - R.de * 3.0;
                if (v >= ms) R.it -= 1;
                R.de = v;
            }
            if (phase == MP_THROUGH) {
                if (R.de > (double)R.ms || !(dt1 > 0.0) || ++guard > kMcGuard) {
                    o.len = dt1;
                    break;
                }
                dtmp = -1.0 - fabs(R.de);
                dt1 += dtmp;
                for (int k = 0; k < 3; k++) R.c[k] += dir[k] * dtmp;
                continue;
            }
            if (phase == MP_NRM) {
                // McMarch::normals (src/mc.rs:531-590): the centre, then z, x, y one-sided
                if (nq == 0) {
                    no = R.de;
                    nn = R.smooth_it;
                    R.calc_sit = false;
                    sca = (seed & 128) == 0 ? 0.15f : -0.15f;
                    tries = 4;
                } else {
                    // the axis of this DE: z, x, y (selects, so that no array is indexed at run time)
                    const double g = (R.de - no) * (double)sca;
                    if (nq == 1) {
                        nv[2] = g;
                        R.c[2] = saved;
                    } else if (nq == 2) {
                        nv[0] = g;
                        R.c[0] = saved;
                    } else {
                        nv[1] = g;
                        R.c[1] = saved;
                    }
                    if (nq == 3) {
                        const double l = nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2];
                        const bool facing = l > 1e-100 && nv[0] * dir[0] + nv[1] * dir[1] + nv[2] * dir[2] < 0.0;
                        if (facing || --tries == 0) {
                            if (facing) {
                                const double d = 1.0 / sqrt(l);
                                for (int q = 0; q < 3; q++) o.normals[q] = nv[q] * d;
                            } else {  // norm_to(-1, raydir)
                                const double d = -1.0 / sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2] + 1e-100);
                                for (int q = 0; q < 3; q++) o.normals[q] = dir[q] * d;
                            }
                            for (int q = 0; q < 3; q++) o.pos[q] = R.c[q];
                            if (is_ray) {
                                mc_colour(P, R, nn, de_limited, o);
                                o.zz = zz;
                                if (de_limited) o.flags |= MB3D_HIP_MC_DE_LIMITED;
                            } else {
                                // the iteration count after the normals, as in MB3D's CalcHSMC
                                mc_colour(P, R, nn, R.it < R.max_its, o);
                                o.zz2 = zz2;
                                o.ms_de_stop = R.ms;
                            }
                            o.kind = 1;
                            cnt.hits++;
                            break;
                        }
                        sca = (float)((mc_rand(seed) - 0.5) * 0.5);
                        nq = 0;
                    }
                }
                if (nq == 0)
                    off = (double)fminf(P.de_stop, 1.0f) * (1.0 + R.mzz * (double)P.de_stop_factor) * (double)sca *
                          P.step_width;
                nq++;
                if (nq == 1) {
                    saved = R.c[2];
                    R.c[2] = saved + off;
                } else if (nq == 2) {
                    saved = R.c[0];
                    R.c[0] = saved + off;
                } else {
                    saved = R.c[1];
                    R.c[1] = saved + off;
                }
                continue;
            }
            bool refine = false;
            if (phase == MP_START) {
                dtmp = R.de;
                if (is_shadow) goto shadow_step;
                if ((T.flags & MB3D_HIP_MC_PROBE) && (R.it >= R.max_its || dtmp < (double)R.ms)) {
                    // the start is inside the set (src/mc.rs:691-727; no inside-and-outside turn-around)
                    o.otrap = (uint16_t)clip15((float)(R.otrap * 4096.0));
                    if (inside) {
                        R.calc_sit = true;
                        phase = MP_PROBE_IN;
                        continue;
                    }
                    const double tr = fmin(fmax(R.rout / P.d_rstop, 0.0), 1.0);
                    o.si_grad = (uint16_t)(32768u + (uint32_t)rte(32767.0 * tr));
                    goto probe_end;
                }
                rlast_sw = dtmp * (double)P.z_step_div;
                goto ray_step;
            }
            if (phase == MP_PROBE_IN) {
                o.si_grad = (uint16_t)(clip15(R.smooth_it * P.mcts_m) | 32768u);
                goto probe_end;
            }
            if (phase == MP_MARCH) {
                dtmp = R.de;
                if (R.it >= R.max_its) {
                    de_limited = false;
                    refine = true;
                } else if (!(dtmp >= (double)R.ms) && R.it >= P.min_it) {
                    refine = true;
                } else {
                    if (dtmp > rlast_de + rlast_sw) dtmp = rlast_de + rlast_sw;
                    if (rlast_de > dtmp + 1e-30) {
                        const double d = rlast_sw / (rlast_de - dtmp);
                        rsf = d < 1.0 ? (float)fmax(d, 0.5) : 1.0f;
                    } else {
                        rsf = 1.0f;
                    }
                    if (++guard > kMcGuard) goto ray_open_air;
                    goto ray_step;
                }
            }
            if (refine) {
                if (inside) dyn_fog *= 200.0f * P.de_stop / (float)P.width;
                o.dyn_fog = dyn_fog;
                o.zz2 = zz2;
                if (inside) o.flags |= MB3D_HIP_MC_INSIDE;
                if (de_limited) {
                    if (secant) {
                        itmp = 4;
                        goto sec_head;
                    }
                    itmp = 10;
                    ds = rlast_sw * -0.5;
                    goto bis_head;
                }
                goto bsi_start;
            }
            if (phase == MP_SEC) {
                dtmp = R.de;
                goto sec_head;
            }
            if (phase == MP_BIS) {
                dtmp = R.de;
                ds = dtmp < (double)R.ms ? fabs(ds) * -0.55 : fabs(ds) * 0.55;
                goto bis_head;
            }
            if (phase == MP_BSI) {
                if (F.after(R, P)) continue;
                if (is_ray)
                    zz += (R.mzz - bsi_from) * zzplus;
                else
                    zz2 += (R.mzz - bsi_from) * zzplus;
                goto normals;
            }
            if (phase == MP_SHADOW) {
                dtmp = R.de;
                if (R.it >= R.max_its || !(dtmp >= (double)R.ms) || !(dtmp < 1e30)) goto shadow_end;
                if (dtmp > rlast_de + rlast_sw) dtmp = rlast_de + rlast_sw;
                if (rlast_de > dtmp + 1e-30) {
                    const double s = rlast_sw / (rlast_de - dtmp);
                    rsf = s < 1.0 ? (float)fmax(s, 0.5) : 1.0f;
                } else {
                    rsf = 1.0f;
                }
                if (!(dt1 >= 0.0) || ++guard > kMcGuard) goto shadow_end;
                goto shadow_step;
            }
            // MP_SSEC: the shadow's secant search took its DE
            dtmp = R.de;
            itmp -= 1;
            goto ssec_head;

        ray_step: {
            // one step of CalcRay's march (src/mc.rs:733-790)
            rlast_de = dtmp;
            double dstep = fmax((dtmp - (double)P.ms_de_sub * (double)R.ms) * (double)P.z_step_div * (double)rsf, 0.11);
            const double d1 = (double)(fmaxf(R.ms, 0.4f) * P.mh04zsd);
            const bool count = P.dfog_on_it == 0 || R.it == P.dfog_on_it;
            if (d1 < dstep) {
                if (count) dyn_fog += (float)(d1 / dstep);
                dstep = d1;
            } else if (count) {
                dyn_fog += 1.0f;
            }
            if (first_step) {
                first_step = false;
                dstep *= mc_rand(seed);
            }
            rlast_sw = dstep;
            zz2 += dstep;
            cnt.steps++;
            if (zz2 > T.len) goto ray_open_air;
            for (int k = 0; k < 3; k++) R.c[k] += dir[k] * dstep;
            zz += dstep * zzplus;
            R.ms = ms_at(zz);
            phase = MP_MARCH;
            continue;
        }
        ray_open_air:
            if (inside) dyn_fog *= 200.0f * P.de_stop / (float)P.width;
            o.dyn_fog = dyn_fog;
            o.zz2 = zz2;
            if (inside) o.flags |= MB3D_HIP_MC_INSIDE;
            for (int k = 0; k < 3; k++) o.pos[k] = R.c[k];
            o.zz = zz;
            break;
        probe_end:
            o.kind = 2;
            for (int k = 0; k < 3; k++) o.pos[k] = R.c[k];
            if (inside) o.flags |= MB3D_HIP_MC_INSIDE;
            break;
        sec_head:
            // the ray's secant search (src/mc.rs:806-830)
            if (itmp > 0 && fabs(dtmp - (double)R.ms) > 0.001) {
                rlast_de = mc_not_zero(rlast_de - dtmp);
                const double ms = (double)R.ms, f = rlast_sw * (dtmp - ms) / rlast_de;
                if (dtmp < ms)
                    rlast_sw = (f >= 0.0 || f < fabs(rlast_sw) * -0.94) ? fabs(rlast_sw) * -0.5 : f;
                else
                    rlast_sw = (f <= 0.0 || f > fabs(rlast_sw) * 0.94) ? fabs(rlast_sw) * 0.5 : f;
                rlast_de = dtmp;
                zz += rlast_sw * zzplus;
                for (int k = 0; k < 3; k++) R.c[k] += dir[k] * rlast_sw;
                R.ms = ms_at(zz);
                itmp -= 1;
                if (itmp > 0) {
                    phase = MP_SEC;
                    continue;
                }
            }
            goto normals;
        bis_head:
            // the ray's bisection (src/mc.rs:831-846)
            if (fabs(dtmp - (double)R.ms) > 0.001) {
                zz += ds * zzplus;
                for (int k = 0; k < 3; k++) R.c[k] += dir[k] * ds;
                R.ms = ms_at(zz);
                itmp -= 1;
                if (itmp > 0) {
                    phase = MP_BIS;
                    continue;
                }
            }
            goto normals;
        bsi_start:
            // bin_search_it on the ray from the depth (src/mc.rs:592-643; 10 turns)
            bsi_from = is_ray ? zz : zz2;
            for (int k = 0; k < 3; k++) R.v[k] = dir[k];
            R.mzz = bsi_from;
            F.start_bsi(R, 10);
            phase = MP_BSI;
            continue;
        shadow_step:
            // one step towards the light (src/mc.rs:881-891)
            rlast_de = dtmp;
            dtmp = fmin(dtmp * (double)P.z_step_div * (double)rsf, (double)(fmaxf(R.ms, 0.4f) * P.mh04zsd));
            rlast_sw = dtmp;
            dt1 -= dtmp;
            for (int k = 0; k < 3; k++) R.c[k] += dir[k] * dtmp;
            zz2 += dtmp * zzplus;
            R.ms = ms_at(zz2);
            cnt.steps++;
            phase = MP_SHADOW;
            continue;
        shadow_end:
            o.len = dt1;
            if (dt1 <= 0.01 || !(T.flags & MB3D_HIP_MC_TRANS)) break;
            // a transparent object between the surface and the light (src/mc.rs:918-942)
            if (R.it < R.max_its) {
                itmp = 3;
                goto ssec_head;
            }
            goto bsi_start;
        ssec_head:
            if (itmp > 0 && fabs(dtmp - (double)R.ms) > 0.01) {
                rlast_de = mc_not_zero(rlast_de - dtmp);
                const double ms = (double)R.ms, f = rlast_sw * (dtmp - ms) / rlast_de;
                if (dtmp < ms)
                    rlast_sw = (f >= 0.0 || f < fabs(rlast_sw) * -0.94) ? fabs(rlast_sw) * -0.5 : f;
                else
                    rlast_sw = (f <= 0.0 || f > fabs(rlast_sw) * 0.94) ? fabs(rlast_sw) * 0.5 : f;
                rlast_de = dtmp;
                zz2 += rlast_sw * zzplus;
                for (int k = 0; k < 3; k++) R.c[k] += dir[k] * rlast_sw;
                R.ms = ms_at(zz2);
                phase = MP_SSEC;
                continue;
            }
            goto normals;
        normals:
            R.mzz = is_ray ? fabs(zz) : fabs(zz2);
            R.calc_sit = true;
            nq = 0;
            phase = MP_NRM;
        }
        o.seed = seed;
        res[t] = o;
    }
    flush_counters(cnt, ctr);
}

void launch_gen_mc(int grid, hipStream_t s, const DevP &d, const mb3d_hip_mc_task *tasks, long long n, int secant,
                   mb3d_hip_mc_res *res, unsigned long long *ctr, int dem);

}  // namespace mb3d
