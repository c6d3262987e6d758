/*
 * bslstm_tfhe.h -- C ABI of the MI355X (gfx950) TFHE bootstrapping engine.
 *
 * Plain C types only: device pointers are passed as typed pointers, the HIP
 * stream as `void *` (a hipStream_t; NULL = the default stream).  Every entry
 * point is asynchronous on the given stream, allocates no device memory,
 * validates its shape/attribute arguments on the host and returns a status
 * code; bslstm_last_error() gives the message of the last failure on the
 * calling thread.
 *
 * Torus values are integers mod q = 2^log2_q.  Complex128 arrays are
 * interleaved (re, im) doubles.
 *
 * Each entry point names the reference interface it replaces
 * (AMIAD-Research/Blind-Spiking-LSTM, file:line).
 */
#ifndef BSLSTM_TFHE_H
#define BSLSTM_TFHE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Mirrors the XLA FFI error kinds the reference handler
 * returns: ffi::Error::InvalidArgument (csrc/mul_bootstrap.cuh:212,223) and
 * ffi::Error::Internal on a failed launch (csrc/mul_bootstrap.cuh:183-188). */
enum {
    BSLSTM_OK = 0,
    BSLSTM_INVALID_ARGUMENT = 1,
    BSLSTM_INTERNAL = 2
};

/* Message of the last failed call on this thread ("" if none). */
const char *bslstm_last_error(void);

/* ABI version (incremented on any signature change). */
int bslstm_abi_version(void);

/*
 * Blind rotation, plain single-launch FP64 form: one workgroup per
 * ciphertext runs the whole collapsed key-unrolled loop.
 *
 * Replaces the XLA FFI handler `mul_bootstrap` (csrc/mul_bootstrap.cuh:192-243,
 * launcher :138-190, kernel :97-136) registered by cuTFHE/multiply.py:10-14.
 * Operands in the handler's order and layout:
 *   ct_init     int64 (..., 2, N); ct_elems = its element count.  If
 *               ct_elems > 2N every ciphertext has its own test polynomial
 *               (stride 2N), else one is broadcast (:162-164).
 *   bsk         complex128 (n_iter, a=2, l, c=2, m, N/2); bsk_dims = the six
 *               trailing dimensions (:149-156).
 *   all_rot_fft complex128 (2N, N/2) FFT of X^r, or NULL.  Its content is a
 *               fixed function of N (parameters_crypto.py:86-91); the kernel
 *               evaluates the same roots exactly and only checks the shape.
 *   indices     int32 (batch, n_iter, m); idx_elems = its element count.
 *   output      int64 (batch, 2, N).
 * Attributes log2_beta, log2_q, l, degree as in the handler (:238-241).
 * degree must be a power of two in [128, 4096] and l in [1, 4]
 * ("Unsupported degree" / "Unsupported l", :211-224).
 */
int bslstm_mul_bootstrap(void *stream,
                         const int64_t *ct_init, int64_t ct_elems,
                         const double *bsk, const int64_t bsk_dims[6],
                         const double *all_rot_fft, int64_t rot_rows, int64_t rot_cols,
                         const int32_t *indices, int64_t idx_elems,
                         int64_t *output, int64_t batch,
                         int64_t log2_beta, int64_t log2_q, int64_t l, int64_t degree);

/*
 * Blind rotation, batched FP64 form: same operands, attributes, semantics and
 * errors as bslstm_mul_bootstrap, and bit-identical outputs; each workgroup
 * runs two ciphertexts and feeds every bootstrapping-key value it loads into
 * both, halving the key traffic from L2.  This is the engine's form of the
 * reference handler's work (csrc/mul_bootstrap.cuh:97-243); degree 4096 runs
 * one ciphertext per workgroup.
 */
int bslstm_mul_bootstrap_batched(void *stream,
                                 const int64_t *ct_init, int64_t ct_elems,
                                 const double *bsk, const int64_t bsk_dims[6],
                                 const double *all_rot_fft, int64_t rot_rows, int64_t rot_cols,
                                 const int32_t *indices, int64_t idx_elems,
                                 int64_t *output, int64_t batch,
                                 int64_t log2_beta, int64_t log2_q, int64_t l, int64_t degree);

/*
 * Rotation exponents of the collapsed blind rotation.
 * Replaces the index preparation of cuTFHE/multiply.py:67-79 (an XLA fusion
 * in the reference):
 *   val[g,p] = sum_j pk[g*c+j] * bit_j(p)        (bit order of cuTFHE/utils.py:5-8)
 *   idx      = int32(rint(2N*val/(q*n_lut)) * n_lut) mod 2N   (rint half-even,
 *              floor modulo)
 * public_key float64 (batch, n); indices int32 (batch, n/collapse, 2^collapse).
 */
int bslstm_rotation_indices(void *stream, const double *public_key, int64_t batch, int64_t n,
                            int64_t collapse, int64_t log2_q, int64_t degree, int64_t n_lut,
                            int32_t *indices);

/*
 * Many-LUT tail of the programmable bootstrap: rotate each blind-rotated
 * test polynomial by X^(-b_ms), b_ms = rint(2N*b/(q*n_lut))*n_lut, and
 * sample-extract coefficients 0..n_lut-1.
 * Replaces cuboot_merge's tail, schemas/RLWE_jax.py:295-303, with
 * rotate_ciphertext (schemas/text_jax.py:348-362, coef_rotation
 * schemas/polynomial_jax.py:150-169) and sample_extract
 * (schemas/text_jax.py:191-208).
 *   br_out  int64 (batch, 2, N);  body float64 (batch)  (the LWE bodies b)
 *   masks   float64 (batch, n_lut, N);  bodies float64 (batch, n_lut)
 */
int bslstm_rotate_extract(void *stream, const int64_t *br_out, const double *body,
                          int64_t batch, int64_t degree, int64_t log2_q, int64_t n_lut,
                          double *masks, double *bodies);

/*
 * LWE key switch.
 * Replaces key_switch_LWE + gadget_product_LWE (schemas/LWE_jax.py:97-117,
 * 65-82) with the float digit rule of decomposition
 * (schemas/polynomial_jax.py:90-111):
 *   d[j,i] = digits of a[i] (j < l_ks, round half-even, base 2^log2_beta)
 *   a'     = centred_mod(sum_{j,i} d[j,i] * ksk_a[j,i,:], q)
 *   b'     = centred_mod(b + centred_mod(sum_{j,i} d[j,i] * ksk_b[j,i], q), q)
 * a_in float64 (batch, n_in), b_in float64 (batch); ksk_a float64
 * (l_ks, n_in, n_out), ksk_b float64 (l_ks, n_in) as built by
 * get_keyswitch_sk_LWE (schemas/LWE_jax.py:86-93); a_out float64
 * (batch, n_out), b_out float64 (batch).  Requires log2_q <= 32 so that the
 * sums are exact in FP64 (the reference's own tensordot is exact there).
 */
int bslstm_key_switch_lwe(void *stream, const double *a_in, const double *b_in,
                          const double *ksk_a, const double *ksk_b,
                          int64_t batch, int64_t n_in, int64_t n_out,
                          int64_t log2_beta_ks, int64_t l_ks, int64_t log2_q,
                          double *a_out, double *b_out);

#ifdef __cplusplus
}
#endif

#endif /* BSLSTM_TFHE_H */
