/* tonto_gpu.h -- C ABI of the MI355X direct-SCF J/K library (libtonto_gpu.so).
 *
 * Drop-in boundary for Tonto's restricted Cartesian J/K build,
 *   MOLECULE.FOCK:make_r_JK_engine(J,K,P)   (foofiles/molecule.fock.foo:1365-1505),
 * which Tonto calls from make_r_Fock_mx (molecule.fock.foo:125) inside every SCF
 * iteration (molecule.scf.foo:1562-1620).  A Fortran caller binds these functions with
 * iso_c_binding; the shim that replaces the body of make_r_JK_engine is shown in
 * INTEGRATION.md.
 *
 * Conventions
 *  - Every function returns 0 on success and a non-zero TGPU_E_* code on failure;
 *    tgpu_last_error(ctx) then describes the failure.  (Tonto itself aborts through
 *    DIE/ENSURE, include/macros.in:336-367; a Fortran shim turns a non-zero status
 *    into DIE.)
 *  - Matrices are n_bf x n_bf FP64.  Because P, J and K are symmetric, row-major and
 *    column-major (Fortran) storage are interchangeable.
 *  - Shells and basis functions are numbered as in Tonto: atom by atom, shells in
 *    library order, Cartesian components in Tonto's order
 *    (foofiles/gaussian_data.foo:295-479), every component of an l-shell carrying the
 *    x^l normalisation (gaussian_data.foo:757-767).
 *  - Shell pairs use Tonto's order: pair index ab = a(a+1)/2 + b, 0-based, a >= b
 *    (molecule.set.foo:2108-2222).
 *  - A context owns device memory on one device and issues work on one HIP stream
 *    (tgpu_set_stream); it is not thread-safe.
 */
#ifndef TONTO_GPU_H
#define TONTO_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TGPU_OK = 0,
    TGPU_E_ARG = 1,        /* invalid argument */
    TGPU_E_UNSUPPORTED = 2,/* angular momentum or size beyond this build */
    TGPU_E_HIP = 3,        /* HIP runtime error */
    TGPU_E_STATE = 4       /* call out of order (e.g. no Schwarz table yet) */
};

/* flags for tgpu_make_r_JK* */
#define TGPU_JK_MIXED 1    /* quartets with Schwarz bound below the FP32 threshold run in FP32 */
#define TGPU_JK_GRAPH 2    /* device entry only: capture the build once into a hipGraph and
                              replay it while P/J/K pointers, flags and class forms are unchanged */
#define TGPU_J_ONLY 4      /* J only, screened by Schwarz_J_test (molecule.fock.foo:3538-3560):
                              MOLECULE.FOCK:make_r_J_engine (1638-1757); K is returned as zero */
#define TGPU_K_ONLY 8      /* K only, screened by Schwarz_K_test (3562-): J returned as zero */

typedef struct tgpu_ctx tgpu_ctx;

/* Molecule and basis.  Replaces the MOLECULE state that make_r_JK_engine reads from
 * `self` (atom positions, basis shells: types.foo:7273-, SHELL1 1960-1998).
 *   atom_xyz[3*n_atom]   positions in bohr
 *   atom_Z[n_atom]       nuclear charges (used by the one-electron integrals)
 *   shell_l, shell_atom, shell_nprim [n_shell]
 *   exps, coefs          concatenated primitives, shell by shell; coefs are Tonto's
 *                        internal (un-normalised) contraction coefficients
 *                        (foofiles/shell.foo:492-507). */
typedef struct {
    int n_atom;
    int n_shell;
    const double *atom_xyz;
    const double *atom_Z;
    const int *shell_l;
    const int *shell_atom;
    const int *shell_nprim;
    const double *exps;
    const double *coefs;
} tgpu_system;

/* Counters of the last J/K build (algorithmic work actually done). */
typedef struct {
    unsigned long long quartets;        /* contracted shell quartets computed */
    unsigned long long prim_quartets;   /* primitive quartets (after pair skipping) */
    unsigned long long fp32_quartets;   /* of which evaluated in FP32 */
    double kernel_ms;                   /* device time of the ERI kernels (HIP events) */
} tgpu_stats;

/* Create a context on the current HIP device: builds Tonto's shell-pair list and the
 * primitive-pair data with and without the 1e-6 pair skip (shell1pair.foo:100-156,
 * shell1quartet.foo:131-420) and uploads them. */
int tgpu_create(const tgpu_system *sys, tgpu_ctx **ctx);

/* A batch of n_mol independent molecules whose J/K builds run in the same launches
 * (each molecule's build is Tonto's make_r_JK_engine on that molecule alone).  Matrices
 * (P, J, K, S, T, V) are then packed molecule after molecule, n_bf(m)^2 doubles each;
 * max_I is packed in molecule order.  No Tonto counterpart (Tonto builds one molecule). */
int tgpu_create_batch(const tgpu_system *systems, int n_mol, tgpu_ctx **ctx);
int tgpu_batch_layout(const tgpu_ctx *ctx, int *n_mol, long long *n_elements,
                      long long *mat_off /* [n_mol] */, int *n_bf /* [n_mol] */,
                      long long *pair_off /* [n_mol+1] */);
int tgpu_destroy(tgpu_ctx *ctx);
const char *tgpu_last_error(const tgpu_ctx *ctx);

/* Basis dimensions of the (first) molecule: n_bf, n_shell_pairs. */
int tgpu_dims(const tgpu_ctx *ctx, int *n_bf, long long *n_shell_pairs);

/* Issue all further work on this hipStream_t (NULL = the null stream). */
int tgpu_set_stream(tgpu_ctx *ctx, void *hip_stream);

/* Cutoffs of SCF_DATA (types.foo:6097-6104, include/macros.in:695):
 * ERI_Schwarz_cutoff, ERI_J_density_cutoff, ERI_K_density_cutoff,
 * SHELL1QUARTET pair_cutoff.  Defaults 1e-9, 1e-9, 1e-9, 1e-6.  Changing pair_cutoff
 * rebuilds the skipped primitive-pair lists. */
int tgpu_set_cutoffs(tgpu_ctx *ctx, double schwarz, double j_density, double k_density,
                     double pair);

/* Spherical basis (Tonto's use_spherical_basis= TRUE, one molecule per context): with
 * on != 0 the matrices of tgpu_make_r_JK[_device] and tgpu_one_electron are n_sph x n_sph
 * (real solid harmonics in Molden order M = 0, +1, -1, ... for d and f shells, s and p
 * Cartesian; tgpu_dims reports n_sph), and the build is make_r_Fock_mx's spherical branch
 * MOLECULE.FOCK:make_r_JK_direct (molecule.fock.foo:123-128, 1235-1363): max_I from the
 * solid-harmonic (ab|ab) blocks, P_max over solid-harmonic blocks, Schwarz_inequality_test
 * (3454-3495: the largest of the six density blocks, one decision for J and K), integrals
 * as by SHELL1QUARTET:change_to_spherical (shell1quartet.foo:505-563).  Call before
 * tgpu_schwarz (a later switch redoes the Schwarz table).  UHF, J-only/K-only and the
 * allocation-free entry stay Cartesian. */
int tgpu_set_spherical(tgpu_ctx *ctx, int on);

/* Mixed-precision threshold: with TGPU_JK_MIXED, quartets with
 * max_I(ab)*max_I(cd) < threshold run in FP32 (default 1e-6). */
int tgpu_set_fp32_threshold(tgpu_ctx *ctx, double threshold);

/* Schwarz table max_I(ab) = sqrt(max |(ab|ab)|) over shell pairs, in Tonto's pair
 * order (MOLECULE.FOCK:make_max_abab_integrals, molecule.fock.foo:3712-3752).
 * Computes it on the device (once per geometry) and, if max_I != NULL, copies it to
 * the host array of n_shell_pairs doubles.  Must precede the first J/K build. */
int tgpu_schwarz(tgpu_ctx *ctx, double *max_I);

/* J and K from the density P, host arrays of n_bf*n_bf doubles
 * (MOLECULE.FOCK:make_r_JK_engine).  On return J and K are after Tonto's
 * post-processing (weight_diagonal_blocks, blockwise_symmetric_fold, J = 2J,
 * symmetric_reflect; molecule.fock.foo:1485-1500), so F = h + J - K/2 as in
 * molecule.fock.foo:136.  Synchronous. */
int tgpu_make_r_JK(tgpu_ctx *ctx, const double *P, double *J, double *K, int flags);

/* Same on device pointers (n_bf*n_bf doubles each), asynchronous on the context's
 * stream; P must stay valid until the stream reaches this work.  Allocates nothing once
 * tgpu_schwarz has run (a first call without it runs tgpu_schwarz, which allocates the
 * wavefront-form workspace of f-shell bases), except the hipGraph of TGPU_JK_GRAPH. */
int tgpu_make_r_JK_device(tgpu_ctx *ctx, const double *P, double *J, double *K, int flags);

/* Allocation-free variant of tgpu_make_r_JK_device.  No counterpart in Tonto: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same results and flags (TGPU_JK_MIXED,
 * TGPU_J_ONLY, TGPU_K_ONLY; not TGPU_JK_GRAPH); every device buffer the build writes
 * besides J and K (the X/Y accumulators, the pair-ordered density and ket-side J blocks,
 * shell-pair density maxima, counters and the f-shell workspace) is taken from one
 * library-owned array of 64 MiB per GPU, which each GPU holds until the process ends.
 * Limit: about 3 n_bf^2 doubles plus the shell-pair maxima and one workspace slot must
 * fit 64 MiB (about n_bf <= 1650); larger systems get
 * TGPU_E_UNSUPPORTED.  tgpu_schwarz must have run (TGPU_E_STATE otherwise) and the
 * context's device must be current.  All calls on one GPU share the array: calls that
 * overlap on one GPU (two streams, two threads) race.  Asynchronous on the context's
 * stream; tgpu_get_stats reports the build as usual. */
int tgpu_make_r_JK_device_noalloc(tgpu_ctx *ctx, const double *P, double *J, double *K, int flags);

/* UHF J and K from the alpha and beta densities Pa, Pb, host arrays of n_bf*n_bf
 * doubles (MOLECULE.FOCK:make_u_JK_engine, molecule.fock.foo:2558-2712): Ja, Jb, Ka, Kb
 * after the same post-processing as the RHF build (2688-2706).  As in Tonto, screening
 * uses max(|Pa|, |Pb|) per shell pair (3670-3701), Ka = K(Pa), Kb = K(Pb), and
 * Ja + Jb = J(Pa + Pb) with Tonto's split between them (the engine adds the beta
 * density's ket-side J into Ja, JBcd => JA(cd), :2637).  flags: 0 or TGPU_JK_MIXED.
 * Synchronous; afterwards tgpu_get_stats gives the engine's quartet counters and the device
 * time of all three passes. */
int tgpu_make_u_JK(tgpu_ctx *ctx, const double *Pa, const double *Pb, double *Ja, double *Jb,
                   double *Ka, double *Kb, int flags);

/* Same on device pointers, asynchronous on the context's stream.  Three passes of the
 * build; allocates two n_bf*n_bf scratch matrices on its first call. */
int tgpu_make_u_JK_device(tgpu_ctx *ctx, const double *Pa, const double *Pb, double *Ja, double *Jb,
                          double *Ka, double *Kb, int flags);

/* Statistics of the last build (synchronises the stream). */
int tgpu_get_stats(tgpu_ctx *ctx, tgpu_stats *stats);

/* Counters per angular-momentum class pair of the last build: out[3k+0..2] =
 * quartets, primitive quartets, FP32 quartets of class pair k (k < tgpu_n_class_pairs(),
 * order: bra class (lp(lp+1)/2+lq) ascending, ket class <= bra class).  Needs n >= 3*pairs.
 * No Tonto counterpart: measurement support. */
int tgpu_n_class_pairs(void);
int tgpu_get_class_counters(tgpu_ctx *ctx, unsigned long long *out, int n);

/* Per-class-pair device times (HIP events on the context stream) of the last build;
 * profiling is off by default.  No Tonto counterpart: measurement support. */
int tgpu_set_profiling(tgpu_ctx *ctx, int on);
int tgpu_get_class_times(tgpu_ctx *ctx, double *ms, int n);

/* Per-class-pair kernel form: 0 = one lane per quartet (class pairs with <= 200 integrals
 * and (e0|f0) entries), 1 = one wavefront per quartet (every class pair); and, for the
 * wavefront form, the ket chunk a work-group takes at a time (16, 32, 64; 0 = chosen per
 * launch from the grid size).  tgpu_autotune times both forms on a build of the given
 * device density and keeps the faster one per class pair, then times the ket chunks of
 * the wavefront-form class pairs and keeps the fastest; ms_out (optional, 2 doubles per
 * class pair) receives the two forms' times (-1 where form 0 does not exist).
 * No Tonto counterpart: kernel selection. */
int tgpu_set_variant(tgpu_ctx *ctx, int class_pair, int variant);
int tgpu_get_variant(const tgpu_ctx *ctx, int class_pair);
int tgpu_set_ket_chunk(tgpu_ctx *ctx, int class_pair, int chunk);
int tgpu_get_ket_chunk(const tgpu_ctx *ctx, int class_pair);
int tgpu_autotune(tgpu_ctx *ctx, const double *P_device, int flags, double *ms_out);

/* One-electron matrices S, T, V (n_bf*n_bf doubles, host arrays), the core-Hamiltonian
 * inputs of the SCF (MOLECULE.INTS: make_overlap_mx, make_kinetic_energy_mx,
 * make_nuclear_attraction_mx; molecule.ints.foo:34-415).  Computed on the device. */
int tgpu_one_electron(tgpu_ctx *ctx, double *S, double *T, double *V);

#ifdef __cplusplus
}
#endif

#endif
