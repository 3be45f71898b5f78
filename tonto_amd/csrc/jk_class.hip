// One angular-momentum class pair of the J/K build per translation unit, so that the
// classes compile in parallel.  Built with -DTGPU_LP=.. -DTGPU_LQ=.. -DTGPU_LR=..
// -DTGPU_LS=.. -DTGPU_HEAVY=0|1.
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

void TGPU_LAUNCH_NAME_X(TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS)(const JKArgs &a, int variant, bool mixed,
                                                          hipStream_t s)
{
#if !TGPU_HEAVY
    if (variant == 0) {
        const int blocks = (a.n_bra + 3) / 4;
        if (mixed) jk_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, true><<<blocks, 256, 0, s>>>(a);
        else jk_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS, false><<<blocks, 256, 0, s>>>(a);
        return;
    }
#endif
    (void)variant;
    (void)mixed;  // the wavefront form runs in FP64
    // one work-group per (bra pair, chunk of a.ket_chunk kets); chunks past a bra pair's
    // Schwarz frontier exit at once
    const int chunks = (a.ket_chunk > 0 ? (a.n_ket_max + a.ket_chunk - 1) / a.ket_chunk : 1);
    jk_wave_kernel<TGPU_LP, TGPU_LQ, TGPU_LR, TGPU_LS><<<dim3(a.n_bra, chunks > 0 ? chunks : 1), 64, 0, s>>>(a);
}

#if TGPU_LP == TGPU_LR && TGPU_LQ == TGPU_LS
void TGPU_SCHWARZ_NAME_X(TGPU_LP, TGPU_LQ)(const DevPairs &pr, const DevPrims &pm, int start, int n,
                                          double *maxI, hipStream_t s)
{
#if TGPU_HEAVY
    schwarz_wave_kernel<TGPU_LP, TGPU_LQ><<<n, 64, 0, s>>>(pr, pm, start, n, maxI);
#else
    schwarz_kernel<TGPU_LP, TGPU_LQ><<<(n + 63) / 64, 64, 0, s>>>(pr, pm, start, n, maxI);
#endif
}
#endif
