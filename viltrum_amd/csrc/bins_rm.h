/*
 * The per-bin phase in region-major order, for regions of 4 to 6 dimensions (FP32,
 * uniform region choice and uniform points: the estimator of
 * …regions-integrator-parallel-variance-reduction.h:67-103 with rr_uniform_region and
 * region_sampling_uniform, and the alpha = 1 form of
 * …regions-integrator-parallel-control-variates.h:39-65).
 *
 * At these dimensions a residual sample reads its region's 3^D polynomial coefficients
 * (972 B at D = 5), and the bins of a tile draw their samples from a few thousand
 * regions, so gathering the coefficients sample by sample is bound by the L2 request
 * rate.  Here a block takes a tile of 256 bins (16 x 16 with two bin dims) and, per pass
 * of up to RM_SPB samples per bin:
 *   1. bin-major, one lane per bin: draws each sample's region (the same region words as
 *      k_bins1) and looks it up in the bin's footprint list;
 *   2. groups the tile's samples by region in LDS (one counting pass over 256 buckets of
 *      the region id), so the samples of a region sit together;
 *   3. evaluates them in that order, 256 consecutive samples per step: lanes of a wave
 *      that share a region read the same coefficients in the same instruction;
 *   4. writes each sample's weighted f and P back to its (bin, sample) slot, and each
 *      lane sums its bin's samples in sample order, as k_bins1 does.
 * The grouping order depends on LDS atomics, but every sample's value and every bin's
 * sum do not: results are deterministic and the same in a batch as alone.  A sample's
 * point is drawn from its own two Philox words (counter: sample, bin, stream ^ bin_hi,
 * "RM" + word), so step 3 needs nothing from step 1 but the region and the slot.
 */
#ifndef VMX_BINS_RM_H
#define VMX_BINS_RM_H

namespace vmx {

template <typename T, int D, int DB, typename F, typename M>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_bins_rm(const Job<T>* __restrict__ jobs, uint32_t J, RegionArrays<T> r,
                                                 FootArrays<T> fa, const uint32_t* __restrict__ off,
                                                 const uint32_t* __restrict__ lst, const F* __restrict__ fs,
                                                 JobRt* rt, ListPrefix lp, int xcd) {
    static_assert(sizeof(T) == 4 && D >= 4, "FP32 regions of 4..6 dims");
    // LDS for four blocks per CU (at most 40 KB each): the bins' boxes are recomputed
    // where step 3 needs them instead of being kept here
    __shared__ uint32_t s_first[RM_MAXF][256];
    __shared__ uint32_t s_cum[RM_MAXF][256];
    __shared__ alignas(16) unsigned char f_raw[sizeof(F)];
    __shared__ uint32_t buf[2][RM_NS];  // keys: as drawn (then the samples' f terms), grouped by region
    __shared__ float resA[RM_NS];       // the samples' P terms
    __shared__ uint32_t s_n[256];
    __shared__ uint32_t s_cnt[256];
    __shared__ uint32_t s_wtot[4];
    const int tid = threadIdx.x;
    // xcd: each XCD takes one contiguous run of tiles, whose regions then share its L2
    const uint32_t tile = xcd ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    uint32_t j = 0;
    {
        uint32_t lo = 0, hi = J;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (jobs[mid].tile_base2 <= tile)
                lo = mid;
            else
                hi = mid;
        }
        j = __builtin_amdgcn_readfirstlane(lo);
    }
    const Job<T>& jb = jobs[j];
    if (!rm_job(jb.form, jb.N)) return;  // block-uniform: k_bins1 runs this job
    {
        const unsigned char* src = reinterpret_cast<const unsigned char*>(&fs[j]);
        for (int i = tid; i < int(sizeof(F) / 4); i += blockDim.x)
            reinterpret_cast<uint32_t*>(f_raw)[i] = reinterpret_cast<const uint32_t*>(src)[i];
        if (tid == 0)
            for (int i = int(sizeof(F) / 4) * 4; i < int(sizeof(F)); ++i) f_raw[i] = src[i];
    }
    const F& f_lds = *reinterpret_cast<const F*>(f_raw);
    // this lane's bin
    const uint32_t tt = tile - jb.tile_base2;
    uint64_t b = 0;
    bool active;
    uint32_t p[3] = {0, 0, 0};
    if constexpr (DB == 2) {
        const uint32_t tx_n = uint32_t((jb.res[1] + 15) / 16);
        const uint32_t row = (tt / tx_n) * 16 + uint32_t(tid) / 16, col = (tt % tx_n) * 16 + uint32_t(tid) % 16;
        active = row < jb.res[0] && col < jb.res[1];
        p[0] = row;
        p[1] = col;
        b = uint64_t(row) * jb.res[1] + col;
    } else {
        b = uint64_t(tt) * 256 + uint32_t(tid);
        active = b < jb.nbins;
        uint32_t rem = uint32_t(b);
#pragma unroll
        for (int i = DB - 1; i >= 0; --i) {
            const uint32_t ri = uint32_t(jb.res[i]);
            p[i] = i > 0 ? rem % ri : rem;
            rem /= ri;
        }
    }
    T blo[D], bhi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        blo[d] = jb.lo[d];
        bhi[d] = jb.hi[d];
    }
#pragma unroll
    for (int i = 0; i < DB; ++i) {
        blo[i] = jb.lo[i] + T(p[i]) * jb.drange[i];
        bhi[i] = jb.lo[i] + T(p[i] + 1) * jb.drange[i];
    }
    const double factor = double(jb.nbins);
    T* bins = static_cast<T*>(jb.bins);
    // control variate: footprint terms in list order (as k_bins1).  The LDS cache holds
    // every st-th footprint (st = 1 while the list fits): entry m is footprint m * st,
    // its first region and the regions before it (mc entries).
    double approx = 0.0;
    uint32_t o = 0, nf = 0, n = 0, first0 = 0, st = 1, mc = 0;
    if (active) {
        const uint64_t gb = jb.bin_base + b;
        o = off[gb];
        nf = off[gb + 1] - o;
        if (jb.form == VMX_FORM_CV_ALPHA1) approx = double(bins[b]);
        st = nf <= uint32_t(RM_MAXF) ? 1u : (nf + uint32_t(RM_MAXF) - 1) / uint32_t(RM_MAXF);
        for (uint32_t k = 0; k < nf; ++k) {
            const uint32_t fi = lst[o + k];
            if (k % st == 0) {
                const uint32_t f0 = fa.first[fi];
                if (k == 0) first0 = f0;
                s_first[k / st][tid] = f0;
                s_cum[k / st][tid] = n;
                mc = k / st + 1;
            }
            n += fa.count[fi];
            double aa[DB], bb[DB];
            bool empty = false;
#pragma unroll
            for (int d = 0; d < DB; ++d) {
                const T rl = fa.lo[size_t(fi) * 3 + d], rh = fa.hi[size_t(fi) * 3 + d];
                const T il = vmax(blo[d], rl);
                const T ih = vmax(il, vmin(bhi[d], rh));
                if (il >= ih) empty = true;
                aa[d] = double(pos_in(il, rl, rh));
                bb[d] = double(pos_in(ih, rl, rh));
            }
            const double integ =
                empty ? 0.0 : fa.scale[fi] * marginal_subrange_n<double, DB, false>(fa.marg + size_t(fi) * jb.mb, aa, bb);
            approx += empty ? 0.0 : factor * integ;
        }
    }
    s_n[tid] = n;
    const uint64_t spp = jb.spp;
    const bool sampling = active && spp > 0 && n > 0;
    if (active && spp > 0 && n == 0) atomicOr(&rt[j].flags, JF_EMPTY_BIN);
    // region index -> region: the cached entry at or below idx, then (st > 1: the list is
    // longer than PREFIX_MIN, so its prefix counts are stored) the footprint within its
    // stretch of st by binary search over the stored counts
    auto region_of = [&](uint32_t idx) -> uint32_t {
        if (nf == 1) return first0 + idx;
        uint32_t lo = 0, hi = mc;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (s_cum[mid][tid] <= idx)
                lo = mid;
            else
                hi = mid;
        }
        if (st == 1) return s_first[lo][tid] + (idx - s_cum[lo][tid]);
        uint32_t a = lo * st, z = a + st < nf ? a + st : nf;
        while (z - a > 1) {
            const uint32_t mid = (a + z) / 2;
            if (lp.pcum[o + mid] <= idx)
                a = mid;
            else
                z = mid;
        }
        return lp.pfirst[o + a] + (idx - lp.pcum[o + a]);
    };
    constexpr int SR = samp_stride<D>();
    const bool alpha1 = jb.form == VMX_FORM_CV_ALPHA1;
    const double a1w = spp > 0 ? factor * double(n) / double(spp) : 0.0;
    KahanSum<T> sf, sa, mf, ma, maa, mfa;
    double acc1 = 0.0;
    T k_f = 0, k_a = 0;
    for (uint64_t s0 = 0; s0 < spp; s0 += RM_SPB) {  // spp is the job's: block-uniform
        const uint32_t cnt = uint32_t(spp - s0 < uint64_t(RM_SPB) ? spp - s0 : uint64_t(RM_SPB));
        // 1. regions of this pass's samples, bin-major
        static_assert(RM_SPB % 4 == 0, "whole groups of 4 samples per pass");
        for (uint32_t g = 0; g < uint32_t(RM_SPB) / 4; ++g) {
            uint32_t w[4] = {0, 0, 0, 0};
            if (sampling && 4 * g < cnt) {
                const U4 rw = philox4x32_10(U4{uint32_t((s0 >> 2) + g), uint32_t(b), uint32_t(b >> 32) ^ jb.stream, 0u},
                                            jb.seed_lo, jb.seed_hi);
                w[0] = rw.x;
                w[1] = rw.y;
                w[2] = rw.z;
                w[3] = rw.w;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t q = 4 * g + uint32_t(k);
                uint32_t key = RM_NONE;
                if (sampling && q < cnt) {
                    const uint32_t reg = region_of(uniform_index(w[k], n));
                    key = ((reg - jb.reg_base) << RM_SLOT_BITS) | (uint32_t(tid) * RM_SPB + q);
                }
                buf[0][uint32_t(tid) * RM_SPB + q] = key;
            }
        }
        // 2. group by region: count per bucket, scan, place
        s_cnt[tid] = 0;
        __syncthreads();
        for (int q = 0; q < RM_SPB; ++q) {
            const uint32_t key = buf[0][uint32_t(tid) * RM_SPB + q];
            atomicAdd(&s_cnt[key == RM_NONE ? 255u : ((key >> RM_SLOT_BITS) & 255u)], 1u);
        }
        __syncthreads();
        {
            // exclusive prefix of the 256 bucket counts: a scan within each wave, then
            // the totals of the waves before it
            const uint32_t c = s_cnt[tid];
            const int lane = tid & 63, wave = tid >> 6;
            uint32_t incl = c;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t v = __shfl_up(incl, off, 64);
                if (lane >= off) incl += v;
            }
            if (lane == 63) s_wtot[wave] = incl;
            __syncthreads();
            uint32_t base = 0;
            for (int w = 0; w < wave; ++w) base += s_wtot[w];
            s_cnt[tid] = base + incl - c;
        }
        __syncthreads();
        for (int q = 0; q < RM_SPB; ++q) {
            const uint32_t key = buf[0][uint32_t(tid) * RM_SPB + q];
            const uint32_t pos = atomicAdd(&s_cnt[key == RM_NONE ? 255u : ((key >> RM_SLOT_BITS) & 255u)], 1u);
            buf[1][pos] = key;
        }
        __syncthreads();
        // 3. evaluate in grouped order: step q takes grouped samples q*256 .. q*256+255;
        //    the f terms go where the drawn keys were (read for the last time in step 2)
        float* resF = reinterpret_cast<float*>(buf[0]);
#pragma unroll 1
        for (int q = 0; q < RM_SPB; ++q) {
            const uint32_t key = buf[1][uint32_t(q) * 256 + uint32_t(tid)];
            if (key == RM_NONE) continue;
            const uint32_t slot = key & (uint32_t(RM_NS) - 1), lb = slot / RM_SPB, sq = slot % RM_SPB;
            const uint32_t reg = jb.reg_base + (key >> RM_SLOT_BITS);
            // the slot's bin: job-local index (for the point's Philox words) and position
            uint64_t sb;
            uint32_t sp[3] = {0, 0, 0};
            if constexpr (DB == 2) {
                const uint32_t tx_n = uint32_t((jb.res[1] + 15) / 16);
                const uint32_t row = (tt / tx_n) * 16 + lb / 16, col = (tt % tx_n) * 16 + lb % 16;
                sb = uint64_t(row) * jb.res[1] + col;
                sp[0] = row;
                sp[1] = col;
            } else {
                sb = uint64_t(tt) * 256 + lb;
                uint32_t rem = uint32_t(sb);
#pragma unroll
                for (int i = DB - 1; i >= 0; --i) {
                    const uint32_t ri = uint32_t(jb.res[i]);
                    sp[i] = i > 0 ? rem % ri : rem;
                    rem /= ri;
                }
            }
            const uint64_t s = s0 + sq;
            const U4 c0 = philox4x32_10(U4{uint32_t(s), uint32_t(sb), uint32_t(sb >> 32) ^ jb.stream, 0x524D0000u},
                                        jb.seed_lo, jb.seed_hi);
            const U4 c1 = philox4x32_10(U4{uint32_t(s), uint32_t(sb), uint32_t(sb >> 32) ^ jb.stream, 0x524D0001u},
                                        jb.seed_lo, jb.seed_hi);
            const uint32_t cw[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
            const T* rec = r.samp + size_t(reg) * SR;  // lo[D], hi[D], 1/extent[D], coefficients
            T x[D], t[D], ilo[D], ihi[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const T rl = rec[d], rh = rec[D + d];
                // the bin's box, as the bin's own lane computed it
                const int dd = d < DB ? d : 0;
                const T sbl = d < DB ? jb.lo[dd] + T(sp[dd]) * jb.drange[dd] : jb.lo[d];
                const T sbh = d < DB ? jb.lo[dd] + T(sp[dd] + 1) * jb.drange[dd] : jb.hi[d];
                ilo[d] = __builtin_fmaxf(sbl, rl);
                ihi[d] = __builtin_fmaxf(ilo[d], __builtin_fminf(sbh, rh));
                x[d] = mul_add<true>(u01f(cw[d]), ihi[d] - ilo[d], ilo[d]);
                t[d] = (x[d] - rl) * rec[2 * D + d];
            }
            const T vol = box_volume<T, D>(ilo, ihi);
            const T fx = residual_eval<M, D>(f_lds, x);
            const T px = eval_monomials<T, D, true>(rec + 3 * D, t);
            if (alpha1) {
                resF[slot] = (fx - px) * vol;
            } else {
                const T w = T(factor * double(s_n[lb])) * vol;
                resF[slot] = fx * w;
                resA[slot] = px * w;
            }
        }
        __syncthreads();
        // 4. each bin's samples in sample order: groups of 4 summed plainly, compensated once
        if (sampling) {
            for (uint32_t g0 = 0; g0 < cnt; g0 += 4) {
                T gf = 0, ga = 0, gmf = 0, gma = 0, gmaa = 0, gmfa = 0;
                for (uint32_t k = g0; k < g0 + 4 && k < cnt; ++k) {
                    const uint32_t slot = uint32_t(tid) * RM_SPB + k;
                    if (alpha1) {
                        acc1 += double(resF[slot]) * a1w;
                        continue;
                    }
                    const T fsv = resF[slot], asv = resA[slot];
                    if (s0 == 0 && k == 0) {  // the bin's first sample shifts the moments
                        k_f = vabs(fsv);
                        k_a = vabs(asv);
                    }
                    gf += fsv;
                    ga += asv;
                    const T df = vabs(fsv) - k_f, da = vabs(asv) - k_a;
                    gmf += df;
                    gma += da;
                    gmaa += da * da;
                    gmfa += df * da;
                }
                if (!alpha1) {
                    sf.add(gf);
                    sa.add(ga);
                    mf.add(gmf);
                    ma.add(gma);
                    maa.add(gmaa);
                    mfa.add(gmfa);
                }
            }
        }
        __syncthreads();  // buf is rewritten by the next pass
    }
    if (!active) return;
    if (alpha1) {
        bins[b] = T(approx + acc1);
        return;
    }
    double result;
    const double nsd = double(spp);
    const bool fixed_w = jb.form == VMX_FORM_VR_FIXED;
    if (!sampling || (!fixed_w && spp < 2)) {
        result = approx;
    } else {
        const double Sf = double(sf.s) - double(sf.c);
        const double Sa = double(sa.s) - double(sa.c);
        if (fixed_w) {
            result = (Sf - jb.alpha * Sa) / nsd + jb.alpha * approx;
        } else {
            const double Ef = double(mf.s) - double(mf.c);
            const double Ea = double(ma.s) - double(ma.c);
            const double Eaa = double(maa.s) - double(maa.c);
            const double Efa = double(mfa.s) - double(mfa.c);
            const double cov = (Efa - (Ef * Ea) / nsd) / (nsd - 1);
            const double var = (Eaa - (Ea * Ea) / nsd) / (nsd - 1);
            double alpha = 1.0;
            const double v = var > 0.0 ? var : 0.0;
            if (v > 0.0) {
                const double c = cov > 0.0 ? (cov < v ? cov : v) : 0.0;
                alpha = c / v;
            }
            result = (Sf - alpha * Sa) / nsd + alpha * approx;
        }
    }
    bins[b] = T(result);
}

}  // namespace vmx

#endif
