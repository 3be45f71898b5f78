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

void TGPU_LAUNCH_NAME_X(TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS)(const JKArgs &a, int variant, bool mixed,
                                                          hipStream_t s)
{
#if !TGPU_HEAVY
    if (variant == 0) {
        // the LDS-row form for (ss|, (ps| bra pairs, and (pp|, (ds| against (ss| and (ps|
        // kets (2-7 K rows, kernels well below the register limit), when the rows fit a
        // 96 KiB work-group next to the Rys table and the ket queues,
        constexpr int NP = ncart(TGPU_LP), NQ = ncart(TGPU_LQ);
#if (TGPU_LP + TGPU_LQ <= 1) || (TGPU_LP + TGPU_LQ == 2 && TGPU_LR + TGPU_LS <= 1)
        constexpr int NRT = EriBody<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, double>::NROOT;
        constexpr size_t kStatic = rys_block_size(NRT) * sizeof(double) + (TGPU_COMPACT ? 4 * 128 * 8 : 0);
        constexpr size_t kBudget = 96 * 1024;
        static bool attr_set = false;
        if (!attr_set) {
            // dynamic LDS above the 64 KiB default has to be allowed per kernel
            (void)hipFuncSetAttribute((const void *)jk_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, false, true>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)(kBudget - kStatic));
            (void)hipFuncSetAttribute((const void *)jk_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, true, true>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)(kBudget - kStatic));
            attr_set = true;
        }
        const size_t rows = (size_t)(NP + NQ) * a.max_nbf * sizeof(double);
        // and when the molecules' ket segments are on average long enough (>= 4 rounds of
        // 64) to keep the four wavefronts of one bra pair busy: batches of small molecules
        // keep four bra pairs per work-group
        if (a.max_nbf > 0 && rows + kStatic <= kBudget && a.n_ket_avg >= 256) {
            if (mixed) jk_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, true, true><<<a.n_bra, 256, rows, s>>>(a);
            else jk_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, false, true><<<a.n_bra, 256, rows, s>>>(a);
            return;
        }
#else
        (void)NP;
        (void)NQ;
#endif
        const int blocks = (a.n_bra + 3) / 4;
        if (mixed) jk_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, true, false><<<blocks, 256, 0, s>>>(a);
        else jk_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, false, false><<<blocks, 256, 0, s>>>(a);
        return;
    }
#endif
    (void)variant;
    (void)mixed;  // the wavefront form runs in FP64
    using C = WaveCfg<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS>;
    // work-group (x, y): bra pairs x, x + gx, ...; of each, ket chunks y, y + gy, ... of
    // a.ket_chunk kets.  gy splits long prefixes when there are few bra pairs, so that the
    // grid holds about kWaveGrid work-groups; workspace classes stay within their slots.
    constexpr long long kWaveGrid = 65536;
    if (a.n_bra <= 0) return;
    const int kc = a.ket_chunk > 0 && a.ket_chunk <= 64 ? a.ket_chunk : 64;  // one lane per ket
    const long long nch = a.n_ket_max > 0 ? (a.n_ket_max + kc - 1) / kc : 1;
    long long gx = a.n_bra, gy;
    if (C::GB) {
        const long long slots = ws_slot_count(a.ws_doubles, C::WS);
        if (slots <= 0 || !a.ws) return;  // unreachable: tgpu_schwarz allocates it first
        gx = std::min<long long>(gx, slots);
        gy = std::min<long long>(nch, std::max<long long>(1, slots / gx));
    } else {
        gy = std::min<long long>(nch, std::max<long long>(1, (kWaveGrid + gx - 1) / gx));
    }
    gy = std::min<long long>(gy, 65535);
    JKArgs b = a;
    b.ket_chunk = kc;
    jk_wave_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS><<<dim3((unsigned)gx, (unsigned)gy), 64, 0, s>>>(b);
}

size_t TGPU_WS_NAME_X(TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS)()
{
    return WaveCfg<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS>::WS;
}

#if TGPU_LP == TGPU_LR && TGPU_LQ == TGPU_LS
void TGPU_SCHWARZ_NAME_X(TGPU_LP, TGPU_LQ)(const DevPairs &pr, const DevPrims &pm, int start, int n,
                                          double *maxI, double *ws, long long ws_doubles, hipStream_t s)
{
#if TGPU_HEAVY
    using C = WaveCfg<TGPU_LP, TGPU_LQ, TGPU_LP, TGPU_LQ>;
    const long long slots = ws_slot_count(ws_doubles, C::WS);
    const int grid = C::GB ? (int)(n < slots ? n : slots) : n;
    if (grid > 0) schwarz_wave_kernel<TGPU_LP, TGPU_LQ><<<grid, 64, 0, s>>>(pr, pm, start, n, maxI, ws);
#else
    (void)ws;
    (void)ws_doubles;
    schwarz_kernel<TGPU_LP, TGPU_LQ><<<(n + 63) / 64, 64, 0, s>>>(pr, pm, start, n, maxI);
#endif
}
#endif
