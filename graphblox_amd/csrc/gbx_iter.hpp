// gbx_iter.hpp -- per-iteration state, the FP64 grid reduction and the
// PageRank update epilogue shared by the pull iteration (gbx_pr.hip) and the
// binned iteration (gbx_binned.hip).
#pragma once

#include "gbx_plan.hpp"

namespace gbx {

struct IterView {
    float *__restrict__ r;
    const void *__restrict__ c_in;
    void *__restrict__ c_out;
    const float *__restrict__ tele;
    Slot *slots;
    Partial *partials;
    uint32_t *counter;
    uint32_t t;
    uint32_t hot;  // vertices of the contribution vector held in LDS
    float damp;
    double eps;
    int stop;
};

// ------------------------------------------------------------------ reductions

// Cross-lane moves of the wave (DPP, no LDS round trip): row_shr:n within
// rows of 16 lanes, the last lane of a row broadcast to the following
// row(s).  Lanes without a source read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp32(uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROW_MASK, 0xF, true);
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp64(double x)
{
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    const uint64_t lo = dpp32<CTRL, ROW_MASK>((uint32_t)u), hi = dpp32<CTRL, ROW_MASK>((uint32_t)(u >> 32));
    return __longlong_as_double((long long)((hi << 32) | lo));
}

__device__ __forceinline__ double readlane64(double x, int lane)
{
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __longlong_as_double((long long)((hi << 32) | lo));
}

// One step of an inclusive scan of every field (sums; maxima of the
// non-negative 32-bit fields, for which a missing source's 0 is neutral).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void scan_step(Partial &p)
{
    p.err += dpp64<CTRL, ROW_MASK>(p.err);
    p.dang += dpp64<CTRL, ROW_MASK>(p.dang);
    p.sum += dpp64<CTRL, ROW_MASK>(p.sum);
    p.active += dpp32<CTRL, ROW_MASK>(p.active);
    p.cmax = max(p.cmax, dpp32<CTRL, ROW_MASK>(p.cmax));
    p.cnmin = max(p.cnmin, dpp32<CTRL, ROW_MASK>(p.cnmin));
}

// The wave's totals in every lane (all 64 lanes active): an inclusive scan
// leaves them in lane 63, which every lane then reads.  The additions follow
// a fixed order, so the result is deterministic.
__device__ __forceinline__ void wave_reduce(Partial &p)
{
    scan_step<0x111, 0xF>(p);  // row_shr:1
    scan_step<0x112, 0xF>(p);  // row_shr:2
    scan_step<0x114, 0xF>(p);  // row_shr:4
    scan_step<0x118, 0xF>(p);  // row_shr:8
    scan_step<0x142, 0xA>(p);  // row_bcast:15 into rows 1 and 3
    scan_step<0x143, 0xC>(p);  // row_bcast:31 into rows 2 and 3
    p.err = readlane64(p.err, 63);
    p.dang = readlane64(p.dang, 63);
    p.sum = readlane64(p.sum, 63);
    p.active = (uint32_t)__builtin_amdgcn_readlane((int)p.active, 63);
    p.cmax = (uint32_t)__builtin_amdgcn_readlane((int)p.cmax, 63);
    p.cnmin = (uint32_t)__builtin_amdgcn_readlane((int)p.cnmin, 63);
}

__device__ __forceinline__ void add_partial(Partial &q, const Partial &b)
{
    q.err += b.err;
    q.dang += b.dang;
    q.sum += b.sum;
    q.active += b.active;
    q.cmax = max(q.cmax, b.cmax);
    q.cnmin = max(q.cnmin, b.cnmin);
}

// Block partial -> partials[blockIdx]; the last block reduces all partials
// and calls fin(total) on thread 0.  Within a block the waves' partials are
// combined by wave 0 (lane w holds wave w's); the partials are handed over
// with the hand-off stores and loads of gbx_plan.hpp (no device-scope fence),
// each thread of the last block loading its blocks' words in one round trip,
// then reduced in a fixed tree order, so the result is deterministic.
template <class Fin>
__device__ __forceinline__ void grid_reduce(Partial p, Partial *partials, uint32_t *counter, Fin fin)
{
    constexpr int kWaves = kBlock / 64;
    __shared__ Partial sp[kWaves];
    __shared__ bool last;
    const Partial zero = {0.0, 0.0, 0.0, 0u, 0u, 0u, 0u};
    wave_reduce(p);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sp[wid] = p;
    __syncthreads();
    if (wid == 0) {
        Partial q = lane < kWaves ? sp[lane] : zero;
        wave_reduce(q);
        if (lane == 0) {
            handoff_store_words<uint64_t>(partials + blockIdx.x, q);
            stores_performed();
            const uint32_t ticket = atomicAdd(counter, 1u);
            last = ticket == gridDim.x - 1;
        }
    }
    __syncthreads();
    if (!last) return;
    Partial q = zero;
    for (uint32_t b = threadIdx.x; b < gridDim.x; b += kBlock) add_partial(q, handoff_load_words<uint64_t>(partials + b));
    wave_reduce(q);
    __syncthreads();
    if (lane == 0) sp[wid] = q;
    __syncthreads();
    if (wid == 0) {
        Partial tot = lane < kWaves ? sp[lane] : zero;
        wave_reduce(tot);
        if (lane == 0) {
            *counter = 0;
            fin(tot);
        }
    }
}

// ------------------------------------------------------------------ PageRank

template <int KIND, class T>
struct PrEpi {
    float *__restrict__ r;
    T *__restrict__ c_out;
    const uint32_t *__restrict__ outdeg;
    const float *__restrict__ tele;
    float base, damp, fa, fs, fd, uni;
    double invN, eps;
    Partial part;
    __device__ __forceinline__ float init(uint32_t) const { return 0.0f; }
    // Per-row inputs of the update, loadable ahead of the row sum.
    struct In {
        float r_old, ti;
        uint32_t od;
    };
    __device__ __forceinline__ In load(uint32_t i) const
    {
        In x;
        x.r_old = ld_stream(r + i);
        x.od = ld_stream(outdeg + i);
        if constexpr (KIND == K_ADJ) x.ti = ld_stream(tele + i);
        else if constexpr (KIND != K_PR_REF) x.ti = tele ? ld_stream(tele + i) : uni;
        else x.ti = 0.0f;
        return x;
    }
    __device__ __forceinline__ void operator()(uint32_t i, float acc) { apply(i, acc, load(i)); }
    __device__ __forceinline__ void apply(uint32_t i, float acc, const In &x)
    {
        const float r_old = x.r_old;
        const uint32_t od = x.od;
        float r_new;
        if constexpr (KIND == K_PR_REF) {
            r_new = base + damp * acc;  // pageRank.c:324 (no contraction: -ffp-contract=off)
        } else if constexpr (KIND == K_ADJ) {
            // lambda[u] = g[u] + d (sum_v P[u,v] lambda[v] + [u dangling] Lambda/N)
            r_new = x.ti + fs * (acc + (od ? 0.0f : fd));
        } else {
            r_new = fa * x.ti + fs * (acc + fd);
        }
        st_stream(r + i, r_new);
        float cn;
        if constexpr (KIND == K_ADJ) cn = r_new;
        else if constexpr (KIND == K_PR_STOCH_W) cn = od ? r_new : 0.0f;
        else cn = od ? r_new / (float)od : 0.0f;  // pageRank.c:297-301
        store_val(c_out, i, cn);
        const float diff = fabsf(r_new - r_old);
        if constexpr (KIND == K_PR_REF) part.err += (double)diff * invN;
        else part.err += (double)diff;
        part.active += ((double)diff >= eps) ? 1u : 0u;
        if constexpr (KIND == K_ADJ) part.dang += (double)fabsf(r_new);  // |lambda|_1 for the relative stop
        else if (!od) part.dang += (double)r_new;
        part.sum += (double)r_new;
    }
};

// SpMV on the binned iteration: y[i] += the row sum of w x (SPMV.c:171-189;
// y accumulates over the passes).  y is the state's rank array.
struct SpmvBinEpi {
    float *__restrict__ y;
    Partial part;
    struct In {
        float y;
    };
    __device__ __forceinline__ In load(uint32_t i) const { return In{ld_stream(y + i)}; }
    __device__ __forceinline__ void apply(uint32_t i, float acc, const In &x) { st_stream(y + i, x.y + acc); }
    __device__ __forceinline__ void operator()(uint32_t i, float acc) { apply(i, acc, load(i)); }
};

__device__ __forceinline__ SpmvBinEpi make_spmv_epi(const IterView &A)
{
    SpmvBinEpi epi;
    epi.y = A.r;
    epi.part = {0.0, 0.0, 0.0, 0u, 0u};
    return epi;
}

// ------------------------------------------------------------------ fixed point
// The conversions of fixedPoint.h as pageRank.c:496-1030 applies them (REF
// update).  c = r / outdeg in float; to fixed: 64-bit (u64)((double)c * 2^32
// + 0.5); 32-bit (u32)(c * 65536.0f + 0.5f); 16-bit (u16)(c * 16384.0f +
// 0.5f); 8-bit (u8)(c * 64.0f + 0.5f); float to u16/u8 through a 32-bit signed
// truncation keeping the low bits (what x86-64 code of the reference does).
// Back: 64-bit r' = (float)((double)base + (double)d * ((double)s / 2^32));
// others r' = base + d * ((float)s / 2^k) in float (k = 16, 14, 6).
// -ffp-contract=off keeps every multiply and add separately rounded.
template <int KIND>
__device__ __forceinline__ uint64_t fixed_conv(float x)
{
    if constexpr (KIND == K_FIX64) return (uint64_t)((double)x * 4294967296.0 + 0.5);
    else if constexpr (KIND == K_FIX32) return (uint64_t)(uint32_t)(long long)(x * 65536.0f + 0.5f);
    else if constexpr (KIND == K_FIX16) return (uint64_t)(uint16_t)(int)(x * 16384.0f + 0.5f);
    else return (uint64_t)(uint8_t)(int)(x * 64.0f + 0.5f);
}

template <int KIND>
__device__ __forceinline__ float fixed_update(float base, float damp, uint64_t s)
{
    if constexpr (KIND == K_FIX64) return (float)((double)base + (double)damp * ((double)s / 4294967296.0));
    else if constexpr (KIND == K_FIX32) return base + damp * ((float)s / 65536.0f);
    else if constexpr (KIND == K_FIX16) return base + damp * ((float)s / 16384.0f);
    else return base + damp * ((float)s / 64.0f);
}

// Update of a fixed-point kind from the exact integer row sum.
template <int KIND, class T>
struct FixEpi {
    float *__restrict__ r;
    T *__restrict__ c_out;
    const uint32_t *__restrict__ outdeg;
    float base, damp;
    double invN, eps;
    Partial part;
    struct In {
        float r_old;
        uint32_t od;
    };
    __device__ __forceinline__ In load(uint32_t i) const { return In{ld_stream(r + i), ld_stream(outdeg + i)}; }
    __device__ __forceinline__ void apply(uint32_t i, uint64_t s, const In &x)
    {
        const float r_new = fixed_update<KIND>(base, damp, s);  // pageRank.c:577,711,843,976
        st_stream(r + i, r_new);
        c_out[i] = x.od ? (T)fixed_conv<KIND>(r_new / (float)x.od) : (T)0;  // pageRank.c:543-548
        const float diff = fabsf(r_new - x.r_old);
        part.err += (double)diff * invN;
        part.active += ((double)diff >= eps) ? 1u : 0u;
        if (!x.od) part.dang += (double)r_new;
        part.sum += (double)r_new;
    }
    __device__ __forceinline__ void operator()(uint32_t i, uint64_t s) { apply(i, s, load(i)); }
};

template <int KIND, class T>
__device__ __forceinline__ FixEpi<KIND, T> make_fix_epi(const uint32_t *outdeg, uint32_t N, const IterView &A)
{
    FixEpi<KIND, T> epi;
    epi.r = A.r;
    epi.c_out = reinterpret_cast<T *>(A.c_out);
    epi.outdeg = outdeg;
    epi.base = 1.0f - A.damp;  // pageRank.c:55
    epi.damp = A.damp;
    epi.invN = 1.0 / (double)N;
    epi.eps = A.eps;
    epi.part = {0.0, 0.0, 0.0, 0u, 0u};
    return epi;
}

// ------------------------------------------------------------------ quantized
// The quantization of pageRank.c:1216-1231 with quantization.h's GetScale_k
// and quantize_k: scale = (max - min in double) / RANGE_MAX_k (1 when
// min == max), kept in double for 32 bits and rounded to float for 16 and 8
// bits; zero point 0; q = (u32)(c / scale + 0.5) in double, where the u32
// conversion is x86-64's (a 64-bit signed truncation keeping the low 32
// bits: 2^32 and above wrap, 2^63 and above give 0), then CLAMP to the
// width.  Back (pageRank.c:1245): next = (float)(scale * (double)sum) for 32
// bits, scale * (float)sum in float for 16 and 8 bits; r' = base + d * next.
// Contributions are >= 0, so min and max come from the slot's float bits.
template <int KIND>
__device__ __forceinline__ double quant_scale(const Slot &s)
{
    const float mn = __uint_as_float(~s.cnmin), mx = __uint_as_float(s.cmax);
    if (mn == mx) return 1.0;
    const double d = (double)mx - (double)mn;
    if constexpr (KIND == K_Q32) return d / 4294967295.0;
    else if constexpr (KIND == K_Q16) return (double)(float)(d / 65535.0);
    else return (double)(float)(d / 255.0);
}

template <int KIND>
__device__ __forceinline__ uint32_t quantize(float c, double scale)
{
    const double y = (double)c / scale + 0.5;
    const uint32_t q = y < 9223372036854775808.0 ? (uint32_t)(unsigned long long)(long long)y : 0u;
    if constexpr (KIND == K_Q32) return q;
    else if constexpr (KIND == K_Q16) return q > 65535u ? 65535u : q;
    else return q > 255u ? 255u : q;
}

template <int KIND>
__device__ __forceinline__ float quant_update(float base, float damp, double scale, uint64_t s)
{
    float next;
    if constexpr (KIND == K_Q32) next = (float)(scale * (double)s);
    else next = (float)scale * (float)s;
    return base + damp * next;
}

__device__ __forceinline__ void note_contribution(Partial &p, float c)
{
    p.cmax = max(p.cmax, __float_as_uint(c));
    p.cnmin = max(p.cnmin, ~__float_as_uint(c));
}

// Update of a quantized kind from the exact integer row sum; the next
// contributions are FP32 (quantized when the next iteration loads them).
template <int KIND>
struct QuantEpi {
    float *__restrict__ r;
    float *__restrict__ c_out;
    const uint32_t *__restrict__ outdeg;
    float base, damp;
    double scale, invN, eps;
    Partial part;
    struct In {
        float r_old;
        uint32_t od;
    };
    __device__ __forceinline__ In load(uint32_t i) const { return In{ld_stream(r + i), ld_stream(outdeg + i)}; }
    __device__ __forceinline__ void apply(uint32_t i, uint64_t s, const In &x)
    {
        const float r_new = quant_update<KIND>(base, damp, scale, s);  // pageRank.c:1245,1253
        st_stream(r + i, r_new);
        const float c = x.od ? r_new / (float)x.od : 0.0f;  // pageRank.c:1205-1211
        c_out[i] = c;
        note_contribution(part, c);
        const float diff = fabsf(r_new - x.r_old);
        part.err += (double)diff * invN;
        part.active += ((double)diff >= eps) ? 1u : 0u;
        if (!x.od) part.dang += (double)r_new;
        part.sum += (double)r_new;
    }
    __device__ __forceinline__ void operator()(uint32_t i, uint64_t s) { apply(i, s, load(i)); }
};

template <int KIND>
__device__ __forceinline__ QuantEpi<KIND> make_quant_epi(const uint32_t *outdeg, uint32_t N, const IterView &A,
                                                         const Slot *cur)
{
    QuantEpi<KIND> epi;
    epi.r = A.r;
    epi.c_out = reinterpret_cast<float *>(A.c_out);
    epi.outdeg = outdeg;
    epi.base = 1.0f - A.damp;  // pageRank.c:55
    epi.damp = A.damp;
    epi.scale = quant_scale<KIND>(*cur);
    epi.invN = 1.0 / (double)N;
    epi.eps = A.eps;
    epi.part = {0.0, 0.0, 0.0, 0u, 0u};
    return epi;
}

// ------------------------------------------------------------------ data-driven pull
// pageRankDataDrivenPullGraphCSR (pageRank.c:1711-1870): every row is on the
// worklist in the first iteration, afterwards the rows with an in-neighbour
// updated in the previous iteration.  A worklist row computes r' = base +
// d * sum, adds |r' - r| / N to the error, and only when |r' - r| >= epsilon
// stores r', counts as active and puts its out-neighbours on the next
// worklist.  Its next contribution carries that flag in the sign bit; the row
// accumulator (AccDD) hands each row its sum with the OR of its in-edges'
// flags as the sign, so a row learns from its own in-edges whether it is on
// the worklist.
struct DdEpi {
    float *__restrict__ r;
    float *__restrict__ c_out;
    const uint32_t *__restrict__ outdeg;
    float base, damp;
    double dN, eps;
    bool first;
    Partial part;
    __device__ __forceinline__ float init(uint32_t) const { return 0.0f; }
    __device__ __forceinline__ void operator()(uint32_t i, float acc)
    {
        const float r_old = ld_stream(r + i);
        const uint32_t od = ld_stream(outdeg + i);
        float r_now = r_old;
        uint32_t upd = 0;
        if (first || (__float_as_uint(acc) >> 31)) {
            const float r_new = base + damp * fabsf(acc);  // pageRank.c:1800
            const double e = (double)fabsf(r_new - r_old);
            part.err += e / dN;
            if (e >= eps) {  // pageRank.c:1803-1818
                r_now = r_new;
                upd = 0x80000000u;
                part.active += 1u;
                st_stream(r + i, r_new);
            }
        }
        const float c = od ? r_now / (float)od : 0.0f;  // pageRank.c:1772-1778
        st_stream(c_out + i, __uint_as_float(__float_as_uint(c) | upd));
        if (!od) part.dang += (double)r_now;
        part.sum += (double)r_now;
    }
};

__device__ __forceinline__ DdEpi make_dd_epi(const uint32_t *outdeg, uint32_t N, const IterView &A, const Slot *cur)
{
    DdEpi epi;
    epi.r = A.r;
    epi.c_out = reinterpret_cast<float *>(A.c_out);
    epi.outdeg = outdeg;
    epi.base = 1.0f - A.damp;  // pageRank.c:55
    epi.damp = A.damp;
    epi.dN = (double)N;
    epi.eps = A.eps;
    epi.first = cur->iters == 0;
    epi.part = {0.0, 0.0, 0.0, 0u, 0u};
    return epi;
}

// Update coefficients of one iteration from the previous scalar slot.
template <int KIND, class T>
__device__ __forceinline__ PrEpi<KIND, T> make_epi(const uint32_t *outdeg, uint32_t N, const IterView &A,
                                                   const Slot *cur)
{
    PrEpi<KIND, T> epi;
    epi.r = A.r;
    epi.c_out = reinterpret_cast<T *>(A.c_out);
    epi.outdeg = outdeg;
    epi.tele = A.tele;
    epi.base = 1.0f - A.damp;  // pageRank.c:55
    epi.damp = A.damp;
    epi.invN = 1.0 / (double)N;
    epi.eps = A.eps;
    epi.uni = (float)(1.0 / (double)N);
    epi.fa = epi.fs = epi.fd = 0.0f;
    if constexpr (KIND == K_ADJ) {
        epi.fs = A.damp;
        epi.fd = (float)(cur->sum / (double)N);  // Lambda / N
    } else if constexpr (KIND != K_PR_REF) {
        const double S = cur->sum > 0.0 ? cur->sum : 1.0;
        epi.fa = 1.0f - A.damp;
        epi.fs = (float)((double)A.damp / S);
        epi.fd = (float)(cur->dang / (double)N);
    }
    epi.part = {0.0, 0.0, 0.0, 0u, 0u};
    return epi;
}

// The next scalar slot from the reduced partials, with the stop rule.
template <int KIND>
__device__ __forceinline__ Slot next_slot(const Partial &tot, uint32_t iters, double eps, int stop)
{
    Slot s;
    s.err = tot.err;
    s.dang = tot.dang;
    s.sum = tot.sum;
    s.active = tot.active;
    s.iters = iters + 1;
    s.cmax = tot.cmax;
    s.cnmin = tot.cnmin;
    uint32_t done = 0;
    if (stop == GBX_STOP_ACTIVE) done = tot.active == 0;
    else if (stop == GBX_STOP_L1) done = KIND == K_ADJ ? tot.err < eps * tot.dang : tot.err < eps;
    s.done = done;
    s.conv = done;
    return s;
}

}  // namespace gbx
