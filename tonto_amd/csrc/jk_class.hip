// One angular-momentum class pair of the J/K build per translation unit, so that the
// classes compile in parallel.  Built with -DTGPU_LP=.. -DTGPU_LQ=.. -DTGPU_LR=..
// -DTGPU_LS=.. -DTGPU_HEAVY=0|1.
#include <algorithm>

#include "tgpu_device.h"

#define TGPU_STR2(x) #x
#define TGPU_STR(x) TGPU_STR2(x)
#define TGPU_CAT5(a, b, c, d, e) a##b##c##d##e
#define TGPU_CAT3(a, b, c) a##b##c
#define TGPU_GEN_FILE(a, b, c, d) TGPU_STR(gen/eri_##a##b##c##d.inc)
#define TGPU_GEN_FILE_X(a, b, c, d) TGPU_GEN_FILE(a, b, c, d)
#include TGPU_GEN_FILE_X(TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS)

#define TGPU_LAUNCH_NAME_X(a, b, c, d) TGPU_CAT5(tgpu_launch_jk_, a, b, c, d)
#define TGPU_SCHWARZ_NAME_X(a, b) TGPU_CAT3(tgpu_launch_schwarz_, a, b)
#define TGPU_WS_NAME_X(a, b, c, d) TGPU_CAT5(tgpu_wave_ws_, a, b, c, d)

int TGPU_LAUNCH_NAME_X(TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS)(const JKArgs &a, int variant, bool mixed,
                                                         hipStream_t s)
{
#if !TGPU_HEAVY
    if (variant == 0) {
        // The LDS-row form for single molecules with long ket segments (>= 4 rounds of 64 by
        // default, so that the four wavefronts of one bra pair have work): as many of the
        // bra pair's K rows as fit the work-group's LDS budget next to the Rys table and
        // the ket queue, and at least a quarter of them ((fd|ps) at 4160 functions: 7 of
        // 16 rows in LDS, 120 -> 71 ms against every K update global).  A row holds the
        // columns of the ket shells' two angular momenta only.  Not built for (ff|ss): its FP64 instantiation
        // gives wrong K on every cluster tested while its mixed twin and the global-atomic
        // form are exact (cause not found), so that class pair updates K globally.
#if !(TGPU_LP == 3 && TGPU_LQ == 3 && TGPU_LR == 0 && TGPU_LS == 0)
        constexpr int NP = ncart(TGPU_LP), NQ = ncart(TGPU_LQ);
        constexpr int NRT = EriBody<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, double>::NROOT;
        constexpr size_t kStatic = rys_block_size(NRT) * sizeof(double) + (TGPU_COMPACT ? 4 * 128 * 8 : 0);
        constexpr size_t kMaxLds = 155 * 1024;
        static bool attr_set = false;
        if (!attr_set) {
            // dynamic LDS above the 64 KiB default has to be allowed per kernel
            (void)hipFuncSetAttribute((const void *)jk_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, false, true>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)(kMaxLds - kStatic));
            (void)hipFuncSetAttribute((const void *)jk_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, true, true>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)(kMaxLds - kStatic));
            attr_set = true;
        }
        const size_t budget = a.krow_budget > 0 && (size_t)a.krow_budget < kMaxLds ? (size_t)a.krow_budget : kMaxLds;
        const int nR = a.kcol_lofs[TGPU_LR + 1] - a.kcol_lofs[TGPU_LR];
        const int W = TGPU_LR == TGPU_LS ? nR : nR + a.kcol_lofs[TGPU_LS + 1] - a.kcol_lofs[TGPU_LS];
        const size_t row_bytes = (size_t)W * sizeof(double);
        const int fit = a.kcol_map && W > 0 && budget > kStatic ? (int)((budget - kStatic) / row_bytes) : 0;
        const int nkr = fit < NP + NQ ? fit : NP + NQ;
        // (no K digested, as in the UHF ket-side J pass: no rows)
        if (a.kp != HUGE_VAL && a.n_ket_avg >= a.krow_min_kets && nkr > 0 && 4 * nkr >= NP + NQ) {
            JKArgs b = a;
            b.krow_rows = nkr;
            // (..|ss) class pairs: density rows in the LDS left after the K rows, when all
            // of both fit
            const int pfit = (int)((budget - kStatic - (size_t)nkr * row_bytes) / row_bytes);
            b.prow_rows = TGPU_LR == 0 && TGPU_LS == 0 && a.prow_enable && nkr == NP + NQ && pfit >= NP + NQ
                              ? NP + NQ
                              : 0;
            size_t rows = (size_t)(nkr + b.prow_rows) * row_bytes;
            // the bra pair's two P_max rows, in what is left, unless they lower the number of
            // work-groups a CU holds ((pp|ss) at 4160 functions: 2 -> 1, 99 -> 130 ms; (ds|ps),
            // at one either way: 259 -> 239 ms)
            const size_t pm_bytes = 2 * (size_t)a.n_shell * sizeof(double);
            constexpr size_t kCuLds = 160 * 1024;
            b.pmrow = a.pmrow && a.n_shell > 0 && budget >= kStatic + rows + pm_bytes &&
                              kCuLds / (kStatic + rows + pm_bytes) == kCuLds / (kStatic + rows)
                          ? 1
                          : 0;
            if (b.pmrow) rows += pm_bytes;
            if (mixed) jk_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, true, true><<<a.n_bra, 256, rows, s>>>(b);
            else jk_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, false, true><<<a.n_bra, 256, rows, s>>>(b);
            return 0;
        }
#endif
        const int blocks = (a.n_bra + 3) / 4;
        if (mixed) jk_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, true, false><<<blocks, 256, 0, s>>>(a);
        else jk_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, false, false><<<blocks, 256, 0, s>>>(a);
        return 0;
    }
#endif
    (void)variant;
    using C = WaveCfg<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS>;
    // work-group (x, y): bra pairs x, x + gx, ...; of each, ket chunks y, y + gy, ... of
    // a.ket_chunk kets.  gy splits long prefixes when there are few bra pairs, so that the
    // grid holds about kWaveGrid work-groups; workspace classes stay within their slots.
    constexpr long long kWaveGrid = 65536;
    if (a.n_bra <= 0) return 0;
    const int kc = a.ket_chunk > 0 && a.ket_chunk <= 64 ? a.ket_chunk : 64;  // one lane per ket
    const long long nch = a.n_ket_max > 0 ? (a.n_ket_max + kc - 1) / kc : 1;
    long long gx = a.n_bra, gy;
    if (C::GB) {
        const long long slots = ws_slot_count(a.ws_doubles, C::WS);
        // the caller sizes the workspace for at least one slot of every class it launches
        if (slots <= 0 || !a.ws) return TGPU_E_STATE;
        gx = std::min<long long>(gx, slots);
        gy = std::min<long long>(nch, std::max<long long>(1, slots / gx));
    } else {
        gy = std::min<long long>(nch, std::max<long long>(1, (kWaveGrid + gx - 1) / gx));
    }
    gy = std::min<long long>(gy, 65535);
    JKArgs b = a;
    b.ket_chunk = kc;
    // FP64 over the kets before the FP32 split of each prefix (all of it in an FP64 build),
    // then FP32 over the rest
    if (!mixed) b.fp32_thr = 0.0;
    jk_wave_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, double><<<dim3((unsigned)gx, (unsigned)gy), 64, 0, s>>>(b);
    if (mixed) jk_wave_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, float><<<dim3((unsigned)gx, (unsigned)gy), 64, 0, s>>>(b);
    return 0;
}

size_t TGPU_WS_NAME_X(TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS)()
{
    return WaveCfg<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS>::WS;
}

#if TGPU_LP == TGPU_LR && TGPU_LQ == TGPU_LS
void TGPU_SCHWARZ_NAME_X(TGPU_LP, TGPU_LQ)(const DevPairs &pr, const DevPrims &pm, int start, int n,
                                          double *maxI, double *ws, long long ws_doubles, int sph, hipStream_t s)
{
#if TGPU_HEAVY
    using C = WaveCfg<TGPU_LP, TGPU_LQ, TGPU_LP, TGPU_LQ>;
    const long long slots = ws_slot_count(ws_doubles, C::WS);
    const int grid = C::GB ? (int)(n < slots ? n : slots) : n;
    if (grid > 0) schwarz_wave_kernel<TGPU_LP, TGPU_LQ><<<grid, 64, 0, s>>>(pr, pm, start, n, maxI, ws, sph);
#else
    (void)ws;
    (void)ws_doubles;
    schwarz_kernel<TGPU_LP, TGPU_LQ><<<(n + 63) / 64, 64, 0, s>>>(pr, pm, start, n, maxI, sph);
#endif
}
#endif
