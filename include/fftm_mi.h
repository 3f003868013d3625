/*
 * fftm_mi.h -- C ABI of the MI355X-native FFT backend for fftm's single-GPU `ffts`
 * R2C/C2R path.
 *
 * The library replaces the hipFFT calls behind fftm's backend-neutral
 * `external_wrap` boundary.  Every entry point takes plain device pointers, sizes
 * and an opaque HIP stream (`void *`, a hipStream_t; NULL = the null stream).  No
 * C++ or torch types cross this boundary.
 *
 * Status codes use hipfftResult's numbering so that a caller's existing
 * hipfft_error_string / safe-call logic keeps working
 * (source/external_wrap/hipfft_safe_call.h:16-54).  On failure
 * fftm_mi_last_error() returns a message for the calling thread.
 *
 * Synchronisation mirrors the reference: with sync != 0 an exec ends with
 * hipDeviceSynchronize() and reports asynchronous errors
 * (FFTM_HIPFFT_SAFE_CALL, hipfft_safe_call.h:58-75); with sync == 0 it only checks
 * the launch (hipfft_check_no_sync, hipfft_wrap.h:31-40).
 */
#ifndef FFTM_MI_H
#define FFTM_MI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* --- status codes (hipfftResult numbering) -------------------------------- */
#define FFTM_MI_SUCCESS 0
#define FFTM_MI_INVALID_PLAN 1
#define FFTM_MI_ALLOC_FAILED 2
#define FFTM_MI_INVALID_TYPE 3
#define FFTM_MI_INVALID_VALUE 4
#define FFTM_MI_INTERNAL_ERROR 5
#define FFTM_MI_EXEC_FAILED 6
#define FFTM_MI_SETUP_FAILED 7
#define FFTM_MI_INVALID_SIZE 8
#define FFTM_MI_UNALIGNED_DATA 9
#define FFTM_MI_INVALID_DEVICE 11
#define FFTM_MI_NO_WORKSPACE 13
#define FFTM_MI_NOT_IMPLEMENTED 14
#define FFTM_MI_NOT_SUPPORTED 16

/* --- directions: same values as fftm::direction (source/fft_direction.h:9-15) */
#define FFTM_MI_R2C 1
#define FFTM_MI_C2R 2
#define FFTM_MI_C2CF 3
#define FFTM_MI_C2CB 4

/* --- precisions ----------------------------------------------------------- */
#define FFTM_MI_F32 0 /* float real / float2 complex (hipfftReal / hipfftComplex)          */
#define FFTM_MI_F64 1 /* double real / double2 complex (hipfftDoubleReal / ...Complex)     */
/* FP32 arrays and arithmetic; the intermediate half spectra between the passes of an
 * ffts transform are stored in FP16 with one power-of-two exponent per stored line
 * (4 bytes per bin instead of 8).  ffts 3D layout only; lengths 128..2048 (powers of
 * two).  No counterpart in the reference.                                             */
#define FFTM_MI_F32_H16 2
/* As FFTM_MI_F32_H16, with every stored bin a hi + lo pair of FP16 pairs (8 bytes, value
 * (hi + lo 2^-11) 2^e): FP32-level precision at FP32's bytes.  Lengths 128..1024.      */
#define FFTM_MI_F32_H16X2 3

/* --- plan flags ----------------------------------------------------------- */
/* C2R of rank >= 2 keeps its input intact by default and needs a work area of one
 * input-sized complex buffer.  With this flag the input is used as scratch
 * (hipFFT's documented out-of-place C2R behaviour) and the work size is 0. */
#define FFTM_MI_FLAG_OVERWRITE_INPUT 1
/* ffts handles: keep the real axis and its partner axis in separate passes instead of
 * one plane pass (used to measure the fusion; results are the same within rounding). */
#define FFTM_MI_FLAG_NO_PLANE_FUSION 2

/*
 * Plan descriptor: the fields of fftm::wrap::fft_plan_descriptor
 * (source/external_wrap/fft_plan_descriptor.h:15-32) that define a transform.
 * default_layout != 0 means tight row-major layout (inembed/onembed ignored),
 * as the fft(n, batch) constructor does (hipfft_wrap.h:582-602).
 */
typedef struct fftm_mi_plan_desc
{
    int32_t dir;            /* FFTM_MI_R2C / _C2R / _C2CF / _C2CB             */
    int32_t precision;      /* FFTM_MI_F32 / FFTM_MI_F64                       */
    int32_t rank;           /* 1..3                                            */
    int32_t default_layout; /* 0: use inembed/istride/idist/onembed/...       */
    int64_t n[3];
    int64_t inembed[3];
    int64_t istride;
    int64_t idist;
    int64_t onembed[3];
    int64_t ostride;
    int64_t odist;
    int64_t batch;
    int32_t flags;          /* FFTM_MI_FLAG_*                                   */
    int32_t reserved;
} fftm_mi_plan_desc;

typedef struct fftm_mi_plan_s *fftm_mi_plan;

/* Host-only validation of a descriptor and its work size (no GPU call).
 * Replaces the argument checks hipfftMakePlanMany64 performs
 * (hipfft_wrap.h:1263-1266). */
int fftm_mi_plan_validate( const fftm_mi_plan_desc *desc, size_t *work_size );

/* Create a plan: hipfftCreate + hipfftSetAutoAllocation(0) + hipfftMakePlanMany64
 * (source/external_wrap/hipfft_wrap.h:1253-1275, called from the fft<T,D>
 * constructors :520-605).  Uploads the plan's twiddle tables to the current device. */
int fftm_mi_plan_create( const fftm_mi_plan_desc *desc, fftm_mi_plan *plan );

/* Work size in bytes: fft<T,D>::get_work_size (hipfft_wrap.h:608-611). */
int fftm_mi_plan_get_work_size( fftm_mi_plan plan, size_t *bytes );

/* Bind an external work area: hipfftSetWorkArea (hipfft_wrap.h:614-618). */
int fftm_mi_plan_set_work_area( fftm_mi_plan plan, void *work_area );

/* Bind a stream: hipfftSetStream (hipfft_wrap.h:620-624). */
int fftm_mi_plan_set_stream( fftm_mi_plan plan, void *stream );

/* Execute: hipfftExecR2C/D2Z/C2R/Z2D/C2C/Z2Z (hipfft_wrap.h:303-417) through
 * fft<T,D>::exec / exec_no_sync / exec_direction(_no_sync) (:665-734).
 * dir must equal the plan's direction, except that a C2C plan runs either C2C
 * direction (exec_typed_direction, :686-712).  sync selects the safe-call
 * device synchronisation. */
int fftm_mi_plan_exec( fftm_mi_plan plan, int32_t dir, const void *in, void *out, int32_t sync );

/* The descriptor the plan was created with: fft<T,D>::descriptor (:1226-1229). */
int fftm_mi_plan_get_descriptor( fftm_mi_plan plan, fftm_mi_plan_desc *desc );

/* Number of HBM passes (kernel launches) one exec runs (observability). */
int fftm_mi_plan_num_passes( fftm_mi_plan plan, int32_t *passes );

/* Destroy: hipfftDestroy (fft<T,D>::~fft, hipfft_wrap.h:600-606). */
int fftm_mi_plan_destroy( fftm_mi_plan plan );

/*
 * 4D axis permutation dst = transpose(src, perm) of memory arrays (dims listed
 * slowest..fastest; dst axis k is src axis perm[k]); elem_bytes in {4, 8, 16}.
 * Replaces direct_transpose_4d (source/detail/direct_transpose_4d.h:42-120) and
 * runtime_memcpy_4d_slab_transposer (source/detail/runtime_memcpy_4d_slab_transposer.h:24-198).
 */
int fftm_mi_transpose_4d( int32_t elem_bytes, const int64_t src_dims[4], const int32_t perm[4], const void *src,
                          void *dst, void *stream, int32_t sync );

/*
 * Fused ffts transforms (source/ffts.hpp:356-383, 809-867): the whole forward or
 * backward transform of one ffts instance, with the 4D stage transposes folded
 * into the FFT passes (one HBM read + write per transformed axis).
 *
 *   layout FFTM_MI_LAYOUT_3D           real [x][y][z]    <-> spectrum [x][y][z/2+1]
 *   layout FFTM_MI_LAYOUT_4D_PENCIL    real [w][z][y][x] <-> spectrum [y][z][w/2+1][x]
 *   layout FFTM_MI_LAYOUT_4D_SLAB      real [x][y][z][w] <-> spectrum [x][y][z][w/2+1]
 *   layout FFTM_MI_LAYOUT_4D_XZWY      real [x][y][z][w] <-> spectrum [y][z][w/2+1][x]
 *     (the one-rank fftm native_xzwy spectrum, fftm.hpp:2941-2961, 4519-4560)
 *
 * n = {nx, ny, nz[, nw]}.  backward is unnormalised (x N) and keeps its input when a
 * work area of fftm_mi_ffts_get_work_size bytes is bound; with
 * FFTM_MI_FLAG_OVERWRITE_INPUT it uses the input as scratch (ffts 4D backward is
 * destructive, ffts.hpp:377) and needs no work area.
 */
#define FFTM_MI_LAYOUT_3D 0
#define FFTM_MI_LAYOUT_4D_PENCIL 1
#define FFTM_MI_LAYOUT_4D_SLAB 2
#define FFTM_MI_LAYOUT_4D_XZWY 3

typedef struct fftm_mi_ffts_s *fftm_mi_ffts;

int fftm_mi_ffts_validate( int32_t layout, const int64_t *n, int32_t precision, int32_t flags, size_t *work_size );
int fftm_mi_ffts_create( int32_t layout, const int64_t *n, int32_t precision, int32_t flags, fftm_mi_ffts *h );
int fftm_mi_ffts_get_work_size( fftm_mi_ffts h, size_t *bytes );
int fftm_mi_ffts_set_work_area( fftm_mi_ffts h, void *work_area );
int fftm_mi_ffts_set_stream( fftm_mi_ffts h, void *stream );
int fftm_mi_ffts_forward( fftm_mi_ffts h, const void *in, void *out, int32_t sync );
int fftm_mi_ffts_backward( fftm_mi_ffts h, const void *in, void *out, int32_t sync );
int fftm_mi_ffts_num_passes( fftm_mi_ffts h, int32_t *forward_passes, int32_t *backward_passes );
/* Run one pass (0-based) of the forward (backward = 0) or backward (backward = 1)
 * transform alone, on the handle's stream.  The same buffers as the full call are
 * passed; the pass reads and writes exactly what it does inside that call.  Used to
 * time individual kernels; it has no counterpart in the reference. */
int fftm_mi_ffts_run_pass( fftm_mi_ffts h, int32_t backward, int32_t index, const void *in, void *out, int32_t sync );
/* Variant number (0..95) of one pass's kernel: a tile shape 0..7 plus 8 * the tile
 * order 0..3 plus 32 * the radix schedule 0..2.  Orders: 0 one contiguous range of tiles per XCD, 1 the tiles dealt to the
 * XCDs round-robin, 2 and 3 chunks of 8 and of 32 consecutive tiles dealt
 * round-robin.  Shapes:
 * 0 default, 1 half the columns per workgroup, 2 twice the threads per column, 3 the
 * tile of 4 exchanged through LDS one component at a time, 4 twice the columns
 * (256-byte row segments, up to 160 KB of LDS), 5 twice the columns and the threads
 * per column, 6 the tile of 0 exchanged through LDS one component at a time (half the
 * LDS per workgroup), 7 as 6 with registers capped for 4 waves per SIMD.  Radix
 * schedules (strided FP32 C2C passes of length 512 and 1024, shapes 0 and 6, and shape 2
 * for schedule 1): 0 the
 * default factorisation (8 8 8, 16 8 8), 1 the alternative 8 16 4 / 8 16 8, 2 that
 * alternative with its radix-16 stage as FP16 matrix-core products on hi + lo operand
 * splits with FP32 accumulation.  Within one radix schedule results are bitwise
 * identical across shapes and orders; another schedule rounds differently (FP32
 * accuracy either way).  Set by the tuner or by hand.
 * FP16-storage passes (FFTM_MI_F32_H16) have shapes 0, 2, 6 and 7, hi + lo storage passes
 * (FFTM_MI_F32_H16X2) shapes 0, 2 and 6.  Lengths compiled
 * with the default shape only (below 64, and the 3/5/7-smooth lengths) run the default
 * shape whatever the variant; plane and any-radix passes also keep their own tile order. */
int fftm_mi_ffts_set_pass_variant( fftm_mi_ffts h, int32_t backward, int32_t index, int32_t variant );
int fftm_mi_ffts_get_pass_variant( fftm_mi_ffts h, int32_t backward, int32_t index, int32_t *variant );
/* Algorithmic HBM bytes one pass moves (reads + writes of its input and output). */
int fftm_mi_ffts_pass_bytes( fftm_mi_ffts h, int32_t backward, int32_t index, int64_t *bytes );
int fftm_mi_ffts_destroy( fftm_mi_ffts h );

/*
 * Adjoints and diagonal spectral operators (a capability the reference leaves to its
 * callers' functors, e.g. solve_poisson_3d_functor in
 * source/tests/detail/versioned_ffts_3d_test_driver.h:128-154; these entry points have no
 * counterpart in the reference API).  The spectrum is treated as a real vector space with
 * <X, Y> = Re sum conj(X_k) Y_k.  With w_k = 1 on the DC and (even n) Nyquist planes of
 * the real axis and w_k = 2 elsewhere, C2R(X) = Re sum_k w_k X_k e^{+i k.x}, hence
 *   R2C^T(G) = C2R(G / w)     (a Hermitian-weighted C2R)
 *   C2R^T(v) = w * R2C(v).
 * They run the same passes as backward / forward with the weights folded into the
 * real-axis pass, so each needs the work area when its counterpart does: r2c_adjoint
 * like backward, c2r_adjoint like forward (FFTM_MI_F32_H16 handles keep the FP16
 * spectrum there in both directions).
 */
int fftm_mi_ffts_r2c_adjoint( fftm_mi_ffts h, const void *spectrum_in, void *real_out, int32_t sync );
int fftm_mi_ffts_c2r_adjoint( fftm_mi_ffts h, const void *real_in, void *spectrum_out, int32_t sync );

/*
 * out_k = H_k in_k on the handle's spectrum layout (in == out allowed).  H is real:
 * either per bin (filter: one T per complex bin, spectrum layout) or radial
 * (shell_coeffs[nshells]: H_k = shell_coeffs[shell(k)], shell(k) = min(S-1,
 * floor(S |k| / kmax)) evaluated exactly (the largest s with s^2 kmax^2 <= S^2 |k|^2),
 * kmax = sqrt(sum_a (n_a/2)^2), k = i <= n/2 ? i : i - n on full
 * axes and k = i on the real axis).  Exactly one of filter / shell_coeffs is non-NULL.
 * The operator is self-adjoint, so it also serves as its own adjoint.
 */
int fftm_mi_ffts_spectral_multiply( fftm_mi_ffts h, const void *in, void *out, const void *filter,
                                    const void *shell_coeffs, int32_t nshells, int32_t sync );

/*
 * Analytic diagonal operators on the half spectrum (k as for the shell filter; k_a the
 * wavenumber along logical axis a, Nyquist positive as in the reference drivers):
 *   FFTM_MI_OP_POISSON          H = -scale / |k|^2, H(0) = 0   (periodic inverse
 *                               Laplacian, solve_poisson_3d_functor of
 *                               source/tests/detail/versioned_ffts_3d_test_driver.h:128-154)
 *   FFTM_MI_OP_GAUSSIAN         H = scale * exp(-|k|^2 / (2 sigma^2))
 *   FFTM_MI_OP_SHELL            H = scale * coeffs[shell(k)]   (coeffs: device T[nshells])
 *   FFTM_MI_OP_DERIVATIVE       H = i scale k_a   (poisson_3d_derivative_spectra_functor,
 *                               versioned_ffts_3d_test_driver.h:156-184)
 *   FFTM_MI_OP_POISSON_GRADIENT H = -i scale k_a / |k|^2, H(0) = 0 (the two above in one)
 *   FFTM_MI_OP_BANDPASS         H = scale where klo <= |k| <= khi, else 0
 */
#define FFTM_MI_OP_POISSON 1
#define FFTM_MI_OP_GAUSSIAN 2
#define FFTM_MI_OP_SHELL 3
#define FFTM_MI_OP_DERIVATIVE 4
#define FFTM_MI_OP_POISSON_GRADIENT 5
#define FFTM_MI_OP_BANDPASS 6
typedef struct fftm_mi_spectral_op
{
    int32_t     kind;
    int32_t     nshells; /* OP_SHELL: 1..1024                                   */
    double      scale;   /* overall factor, e.g. 1/N to normalise a round trip   */
    double      sigma;   /* OP_GAUSSIAN: width in wavenumber units (> 0)         */
    const void *coeffs;  /* OP_SHELL: device array of nshells values of type T  */
    int32_t     axis;    /* OP_DERIVATIVE / OP_POISSON_GRADIENT: logical axis    */
    int32_t     reserved;
    double      klo;     /* OP_BANDPASS: band edges in wavenumber units          */
    double      khi;
} fftm_mi_spectral_op;

/* spectrum_out = H spectrum_in on the handle's spectrum layout (in == out allowed). */
int fftm_mi_ffts_spectral_apply( fftm_mi_ffts h, const void *spectrum_in, void *spectrum_out,
                                 const fftm_mi_spectral_op *op, int32_t sync );

/*
 * real_out = backward(H forward(real_in)) in one call, the spectrum kept in the work
 * area.  The last forward pass, the multiply and the first backward pass run as one
 * kernel along the same lines (one HBM round trip instead of three), and op->scale
 * absorbs the 1/N normalisation, so a periodic Poisson solve or a spectral low-pass is
 * five passes over HBM in 3D.  Replaces the reference's forward / functor / backward /
 * scale sequence (versioned_ffts_3d_test_driver.h:640-655).  Needs the work area (a
 * handle created without FFTM_MI_FLAG_OVERWRITE_INPUT); not for FFTM_MI_F32_H16.
 * real_in == real_out is allowed.
 */
int fftm_mi_ffts_filter( fftm_mi_ffts h, const void *real_in, void *real_out, const fftm_mi_spectral_op *op,
                         int32_t sync );

/* grad[s] = sum_{k in shell s} Re(conj(G_k) X_k) (gradient of a loss w.r.t. the shell
 * coefficients, G = dL/dZ for Z = H X), accumulated in FP64, written as T[nshells]. */
int fftm_mi_ffts_shell_gradient( fftm_mi_ffts h, const void *grad_spectrum, const void *spectrum, void *grad,
                                 int32_t nshells, int32_t sync );

/*
 * fftm_mi_ffts_shell_gradient without allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  The shell partial sums live in the library's
 * static device scratch (512 x 1024 doubles, 4 MiB, one copy per GPU held until the process
 * ends) instead of a buffer allocated on the first call; nshells is limited to 1..1024 as
 * for fftm_mi_ffts_shell_gradient.  Every call on a GPU shares that GPU's scratch: calls
 * that overlap in time on one GPU (other streams, other threads) race and must be ordered
 * by the caller.  Same arguments and results as fftm_mi_ffts_shell_gradient.
 */
int fftm_mi_ffts_shell_gradient_noalloc( fftm_mi_ffts h, const void *grad_spectrum, const void *spectrum, void *grad,
                                         int32_t nshells, int32_t sync );

/*
 * Tuning of the tile-shape variants with the reference's autotune cache protocol
 * (source/fftm_autotune.hpp:126-169, fftm_autotune_measure.hpp:293-400,
 * detail/fftm_autotune_hardware.inc:87-98, 285-337): for every pass of the handle and
 * every variant, regenerate the pass input, check the output is bitwise equal to
 * variant 0, time `iterations` launches after `warmup` (median), keep the fastest valid
 * one.  real_in / spectrum_out are scratch buffers of the forward's shapes (overwritten).
 * With cache_path the per-candidate records and winners are merged into a KEY=VALUE
 * cache ("# FFTM C++ autotune cache", tmp + rename) keyed by problem and by the v2
 * hardware signature (FNV-1a 64 of the canonical device string).
 * load_tuning applies a cache's winners when its signature matches this device; a
 * handle created while FFTM_MI_AUTOTUNE_CACHE names a cache file loads it.
 */
int fftm_mi_ffts_autotune( fftm_mi_ffts h, const void *real_in, void *spectrum_out, int32_t warmup,
                           int32_t iterations, const char *cache_path );
int fftm_mi_ffts_load_tuning( fftm_mi_ffts h, const char *cache_path, int32_t *applied );
int fftm_mi_hardware_signature( char *buf, size_t len );
/* FNV-1a 64 of a string as 16 lowercase hex digits (host only). */
int fftm_mi_fnv1a64( const char *text, char *hex17 );

/*
 * The benchmark harness's deterministic input on the device: fill_random_real_3d_functor /
 * fill_random_real_4d_functor with splitmix64 and unit_random_from_state<T>
 * (source/tests/detail/fft_benchmark_common.h:201-257).  Logical element (x, y, z[, w])
 * (rank 3 or 4; n[a] extents, strides[a] element strides of logical axis a in dst, any
 * arrangement) gets 2 u - 1, u = (splitmix64(state) >> 11) 2^-53 rounded to T, with
 * state = seed ^ gx 0xD6E8FEB86659FD93 ^ gy 0xA5A3564E27F886A5 ^ gz 0x9E3779B97F4A7C15
 * [^ gw 0x94D049BB133111EB] and g = start[a] + index (start may be NULL: zeros; a slab's
 * origin otherwise).  precision: FFTM_MI_F32 or FFTM_MI_F64.  Runs on `stream`.
 */
int fftm_mi_fill_random_real( int32_t precision, int32_t rank, const int64_t *n, const int64_t *strides,
                              const int64_t *start, uint64_t seed, void *dst, void *stream, int32_t sync );

/* Last error message of the calling thread ("" when none). */
const char *fftm_mi_last_error( void );

/* Library version: major * 10000 + minor * 100 + patch. */
int fftm_mi_version( void );

#ifdef __cplusplus
}
#endif

#endif /* FFTM_MI_H */
